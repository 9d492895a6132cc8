// blocked.hip -- set-up of the LDS-staged SpMV layout (see spmv.hpp, k_spmv_blocked).
//
// Non-empty rows are grouped into blocks of rows that touch nearby columns:
//   rays   (rows of A)  : up to RB consecutive rays of one angle; their columns (pixels,
//                         column-major) are bucketed by image column (W = n), where the band of
//                         rays covers one contiguous pixel range per column;
//   tiles  (rows of A^T): the pixels of a TX x TY image tile; their columns (rays, angle-major)
//                         are bucketed by angle (W = p), where the rays crossing the tile form
//                         one contiguous detector range per angle.
// For every block a one-workgroup kernel finds, per bucket, the [min, max] column offset it
// touches; the union of these ranges is the block's footprint (a list of contiguous segments
// copied to LDS at the start of each SpMV), and every entry gets a 16-bit LDS slot.  A block
// whose footprint exceeds the LDS capacity keeps global gathers.  Column indices outside the
// matrix (see raytrace.hip) map to a slot that always holds 0.
#include <algorithm>
#include <cstdlib>
#include <numeric>

#include "fair_internal.hpp"
#include "spmv.hpp"

namespace fair {
namespace {

__device__ int block_exclusive_scan(int v, int *tsum, int *total) {
    const int tid = threadIdx.x;
    tsum[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int add = tid >= off ? tsum[tid - off] : 0;
        __syncthreads();
        tsum[tid] += add;
        __syncthreads();
    }
    int incl = tsum[tid];
    *total = tsum[255];
    __syncthreads();
    return incl - v;
}

// FILL = false: per block the number of segments (or -1 when the footprint + zero slot does
// not fit in fmax entries).  FILL = true: write the segments at binfo.z and the entry slots.
// A bucket's range is cut into segments of at most SEG_CHUNK entries, so that one thread
// stages one segment with a short run of contiguous loads.
template <bool FILL>
__global__ void __launch_bounds__(256) k_footprint(const int4 *__restrict__ binfo,
                                                   const int4 *__restrict__ rows,
                                                   const int32_t *__restrict__ JA, int32_t W,
                                                   int32_t NB, int32_t ncols, int32_t fmax,
                                                   int32_t *nseg, int2 *segs, uint16_t *loc) {
    extern __shared__ int32_t sm[];
    int32_t *bmin = sm, *bmax = sm + NB, *boff = sm + 2 * NB;
    __shared__ int tsum[256];
    const int tid = threadIdx.x;
    const int4 bi = binfo[blockIdx.x];
    if (FILL && bi.z < 0) return;
    for (int q = tid; q < NB; q += 256) {
        bmin[q] = INT32_MAX;
        bmax[q] = -1;
    }
    __syncthreads();
    for (int r = bi.x; r < bi.y; ++r) {
        const int4 ri = rows[r];
        for (int k = ri.y + tid; k < ri.z; k += 256) {
            const int c = JA[k];
            if ((uint32_t)c < (uint32_t)ncols) {
                const int q = c / W, o = c - q * W;
                atomicMin(&bmin[q], o);
                atomicMax(&bmax[q], o);
            }
        }
    }
    __syncthreads();
    const int per = (NB + 255) / 256;
    const int q0 = min(tid * per, NB), q1 = min(q0 + per, NB);
    int len = 0, cnt = 0;
    for (int q = q0; q < q1; ++q)
        if (bmax[q] >= 0) {
            const int l = bmax[q] - bmin[q] + 1;
            len += l;
            cnt += (l + SEG_CHUNK - 1) / SEG_CHUNK;
        }
    int F, S;
    int lenoff = block_exclusive_scan(len, tsum, &F);
    int segoff = block_exclusive_scan(cnt, tsum, &S);
    if (!FILL) {
        if (tid == 0) nseg[blockIdx.x] = (F + 1 <= fmax) ? S : -1;
        return;
    }
    for (int q = q0; q < q1; ++q) {
        if (bmax[q] >= 0) {
            const int l = bmax[q] - bmin[q] + 1;
            boff[q] = lenoff;
            for (int c0 = 0; c0 < l; c0 += SEG_CHUNK) {
                const int cl = min(SEG_CHUNK, l - c0);
                segs[bi.z + segoff] = make_int2(q * W + bmin[q] + c0, (cl << 16) | (lenoff + c0));
                ++segoff;
            }
            lenoff += l;
        }
    }
    __syncthreads();
    const int zero_slot = fmax - 1;
    for (int r = bi.x; r < bi.y; ++r) {
        const int4 ri = rows[r];
        for (int k = ri.y + tid; k < ri.z; k += 256) {
            const int c = JA[k];
            int slot = zero_slot;
            if ((uint32_t)c < (uint32_t)ncols) {
                const int q = c / W;
                slot = boff[q] + (c - q * W - bmin[q]);
            }
            loc[k] = (uint16_t)slot;
        }
    }
}

// Group the non-empty rows by key (stable), split groups into blocks of at most rb rows and
// build the footprints.
void build_blocked(Csr &A, const std::vector<int64_t> &key_of_row, int32_t rb, int32_t W, int32_t NB,
                   int32_t lds_elems) {
    Blocked &B = A.blk;
    B = Blocked();
    if (A.n_rows_nz < 0) build_row_list(A);
    const int32_t nr = A.n_rows_nz;
    // 16-bit LDS slots; the FP64 footprint must fit the workgroup's LDS
    if (nr <= 0 || NB > 8192 || lds_elems > 65535 || rb < 1 ||
        (int64_t)lds_elems * (int64_t)sizeof(double) > (int64_t)lds_per_block())
        return;
    std::vector<int4> desc(nr);
    FAIR_HIP(hipMemcpyAsync(desc.data(), A.rows_nz.p, sizeof(int4) * nr, hipMemcpyDeviceToHost, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
    std::vector<int32_t> order(nr);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return key_of_row[desc[a].x] < key_of_row[desc[b].x];
    });
    std::vector<int4> rows(nr), binfo;
    int64_t prev_key = -1;
    int in_block = 0;
    for (int32_t i = 0; i < nr; ++i) {
        rows[i] = desc[order[i]];
        const int64_t k = key_of_row[rows[i].x];
        if (binfo.empty() || k != prev_key || in_block == rb) {
            if (!binfo.empty()) binfo.back().y = i;
            binfo.push_back(make_int4(i, i, 0, 0));
            in_block = 0;
        }
        prev_key = k;
        ++in_block;
    }
    binfo.back().y = nr;
    const int32_t nb = (int32_t)binfo.size();
    B.rows.alloc(4 * (size_t)nr);
    B.binfo.alloc(4 * (size_t)nb);
    FAIR_HIP(hipMemcpyAsync(B.rows.p, rows.data(), sizeof(int4) * nr, hipMemcpyHostToDevice, stream()));
    FAIR_HIP(hipMemcpyAsync(B.binfo.p, binfo.data(), sizeof(int4) * nb, hipMemcpyHostToDevice, stream()));
    DBuf<int32_t> nseg(nb);
    const size_t shm = sizeof(int32_t) * (3 * (size_t)NB + 1);
    hipLaunchKernelGGL(k_footprint<false>, dim3(nb), dim3(256), shm, stream(), (const int4 *)B.binfo.p,
                       (const int4 *)B.rows.p, A.JA.p, W, NB, A.n, lds_elems, nseg.p, (int2 *)nullptr,
                       (uint16_t *)nullptr);
    launch_check("footprint count");
    std::vector<int32_t> hn(nb);
    nseg.download(hn.data(), nb);
    FAIR_HIP(hipStreamSynchronize(stream()));
    int64_t total = 0;
    int32_t direct = 0;
    for (int32_t b = 0; b < nb; ++b) {
        if (hn[b] < 0) {
            binfo[b].z = -1;
            binfo[b].w = -1;
            ++direct;
        } else {
            binfo[b].z = (int32_t)total;
            total += hn[b];
            binfo[b].w = (int32_t)total;
        }
    }
    FAIR_CHECK(total < INT32_MAX, FAIR_ERR_OVERFLOW, "too many footprint segments");
    FAIR_HIP(hipMemcpyAsync(B.binfo.p, binfo.data(), sizeof(int4) * nb, hipMemcpyHostToDevice, stream()));
    B.segs.alloc(2 * (size_t)(total > 0 ? total : 1));
    B.loc.alloc((size_t)A.nnz + CSR_PAD);
    FAIR_HIP(hipMemsetAsync(B.loc.p, 0, sizeof(uint16_t) * ((size_t)A.nnz + CSR_PAD), stream()));
    hipLaunchKernelGGL(k_footprint<true>, dim3(nb), dim3(256), shm, stream(), (const int4 *)B.binfo.p,
                       (const int4 *)B.rows.p, A.JA.p, W, NB, A.n, lds_elems, (int32_t *)nullptr,
                       (int2 *)B.segs.p, B.loc.p);
    launch_check("footprint fill");
    FAIR_HIP(hipStreamSynchronize(stream()));
    B.nblocks = nb;
    B.lds_elems = lds_elems;
    B.n_direct = direct;
    B.ok = true;
}

}  // namespace

// Rays are grouped by sinogram tile: DA consecutive angles x DJ consecutive detector
// positions (rays of neighbouring angles through the same detector position nearly coincide,
// so their pixel footprints overlap).
void build_blocked_rays(Csr &A, int32_t img_n, int32_t p, int32_t DA, int32_t DJ, int32_t lds_elems) {
    std::vector<int64_t> key(A.m);
    const int64_t jblocks = (p + DJ - 1) / DJ;
    for (int32_t r = 0; r < A.m; ++r) {
        const int32_t th = r / p, j = r - th * p;
        key[r] = (int64_t)(th / DA) * jblocks + j / DJ;
    }
    build_blocked(A, key, DA * DJ, img_n, img_n, lds_elems);
}

void build_blocked_tiles(Csr &T, int32_t img_n, int32_t p, int32_t nth, int32_t TX, int32_t TY,
                         int32_t lds_elems) {
    std::vector<int64_t> key(T.m);
    const int64_t tiles_per_col = (img_n + TY - 1) / TY;
    for (int32_t c = 0; c < T.m; ++c) {
        const int32_t ix = c / img_n, r = c - ix * img_n;
        key[c] = (int64_t)(ix / TX) * tiles_per_col + r / TY;
    }
    build_blocked(T, key, TX * TY, p, nth, lds_elems);
}

// Build the staged layouts of a ray-traced matrix and of its transpose (once).  The layouts
// can be switched off with FAIR_SPMV_STAGED=0; FAIR_DA / FAIR_DJ / FAIR_LDS_A and FAIR_TX /
// FAIR_TY / FAIR_LDS_T override the default block shapes (fair_autotune_spmv picks them by
// timing).
// Whether ensure_blocked would still build a staged layout.
bool blocked_pending(const fair_csr *h) {
    if (h->g_n <= 0 || h->transposed || h->blk_tuned || !env_int("FAIR_SPMV_STAGED", 1)) return false;
    return !h->A.blk.ok || !h->have_T || !h->T.blk.ok;
}

void ensure_blocked(fair_csr *h) {
    if (h->g_n <= 0 || h->transposed || h->blk_tuned || !env_int("FAIR_SPMV_STAGED", 1)) return;
    const bool built = !h->A.blk.ok || !h->have_T || !h->T.blk.ok;
    if (!h->A.blk.ok)
        build_blocked_rays(h->A, h->g_n, h->g_p, env_int("FAIR_DA", 4), env_int("FAIR_DJ", 4),
                           env_int("FAIR_LDS_A", 6144));
    ensure_T(h);
    if (!h->T.blk.ok) {
        const int tx = env_int("FAIR_TX", 16), ty = env_int("FAIR_TY", 8);
        build_blocked_tiles(h->T, h->g_n, h->g_p, h->g_nth, tx, ty, env_int("FAIR_LDS_T", 6144));
    }
    if (env_int("FAIR_DEBUG", 0) && built)
        fprintf(stderr, "[fair] staged SpMV: A %d blocks (%d direct), A^T %d blocks (%d direct)\n",
                h->A.blk.nblocks, h->A.blk.n_direct, h->T.blk.nblocks, h->T.blk.n_direct);
}

}  // namespace fair

namespace fair {
namespace {

// Average time of one A*x launch of the staged kernel (3 launches after one warm-up).
float time_blocked(const Csr &A, const double *x, double *y, hipEvent_t e0, hipEvent_t e1) {
    auto run = [&] {
        spmv_launch_blocked<double, SrcPtr<double>, EpiStore<double>>(A.blk, A.JA.p, A.vA.p, A.n, SrcPtr<double>{x},
                                                                     EpiStore<double>{y}, stream());
    };
    run();
    FAIR_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < 3; ++r) run();
    FAIR_HIP(hipEventRecord(e1, stream()));
    FAIR_HIP(hipEventSynchronize(e1));
    launch_check("spmv autotune");
    float ms = 0.f;
    FAIR_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms / 3.f;
}

float time_direct(const Csr &A, const double *x, double *y, hipEvent_t e0, hipEvent_t e1) {
    auto run = [&] {
        spmv_launch_raw<double, SrcPtr<double>, EpiStore<double>>((const int4 *)A.rows_nz.p, A.n_rows_nz, A.JA.p,
                                                                  A.vA.p, A.n, SrcPtr<double>{x}, EpiStore<double>{y},
                                                                  stream());
    };
    run();
    FAIR_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < 3; ++r) run();
    FAIR_HIP(hipEventRecord(e1, stream()));
    FAIR_HIP(hipEventSynchronize(e1));
    launch_check("spmv autotune");
    float ms = 0.f;
    FAIR_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms / 3.f;
}

struct Shape {
    int32_t a, b, lds;
};

}  // namespace
}  // namespace fair

using namespace fair;

// Block shapes of the staged SpMV picked by timing on this matrix: sinogram tiles of DA angles x
// DJ detector positions for A, image tiles of TX x TY pixels for A^T, each with an LDS footprint
// capacity.  choice (may be NULL) receives {DA, DJ, LDS_A, TX, TY, LDS_T}.  No counterpart in the
// reference (its products are OpenMP loops over rows, sparse_matrices.f90:349-393).
extern "C" int fair_autotune_spmv(fair_csr *h, int32_t *choice) {
    return guard([&] {
        FAIR_CHECK(h != nullptr, FAIR_ERR_ARG, "null matrix handle");
        FAIR_CHECK(h->g_n > 0 && !h->transposed, FAIR_ERR_ARG,
                   "the staged layout needs a matrix built by fanlineartomo or paralleltomo");
        h->blk_tuned = false;
        ensure_blocked(h);
        FAIR_CHECK(h->A.blk.ok && h->T.blk.ok, FAIR_ERR_ARG, "the staged layout is switched off");
        if (h->graphs_used) drop_graphs(h);   // captured launches refer to the old layouts
        Csr &A = h->A, &T = h->T;
        const int32_t big = std::max(A.m, A.n);
        DBuf<double> x(big > 0 ? big : 1), y(big > 0 ? big : 1);
        fill(x.p, big, 1.0);
        hipEvent_t e0, e1;
        FAIR_HIP(hipEventCreate(&e0));
        FAIR_HIP(hipEventCreate(&e1));
        // {0, 0, 0}: no staging (row-per-wavefront kernel with global gathers)
        const Shape ray_shapes[] = {{4, 4, 6144}, {2, 8, 6144}, {4, 8, 6144}, {8, 4, 6144}, {8, 2, 6144},
                                    {16, 1, 6144}, {16, 2, 6144}, {32, 1, 6144}, {8, 8, 12288}, {0, 0, 0}};
        const Shape tile_shapes[] = {{8, 8, 6144}, {8, 16, 6144}, {16, 8, 6144}, {4, 8, 4096},
                                     {16, 16, 12288}, {8, 16, 12288}, {0, 0, 0}};
        auto time_any = [&](const Csr &M) {
            if (M.blk.ok) return time_blocked(M, x.p, y.p, e0, e1);
            return time_direct(M, x.p, y.p, e0, e1);
        };
        Shape bestA = ray_shapes[0], bestT = tile_shapes[0];
        float tA = 1e30f, tT = 1e30f;
        for (const Shape &c : ray_shapes) {
            if (c.a > 0) {
                build_blocked_rays(A, h->g_n, h->g_p, c.a, c.b, c.lds);
                if (!A.blk.ok) continue;
            } else {
                A.blk = Blocked();
            }
            const float t = time_any(A);
            if (t < tA) {
                tA = t;
                bestA = c;
            }
        }
        for (const Shape &c : tile_shapes) {
            if (c.a > 0) {
                build_blocked_tiles(T, h->g_n, h->g_p, h->g_nth, c.a, c.b, c.lds);
                if (!T.blk.ok) continue;
            } else {
                T.blk = Blocked();
            }
            const float t = time_any(T);
            if (t < tT) {
                tT = t;
                bestT = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (bestA.a > 0) build_blocked_rays(A, h->g_n, h->g_p, bestA.a, bestA.b, bestA.lds);
        else A.blk = Blocked();
        if (bestT.a > 0) build_blocked_tiles(T, h->g_n, h->g_p, h->g_nth, bestT.a, bestT.b, bestT.lds);
        else T.blk = Blocked();
        FAIR_CHECK((bestA.a == 0 || A.blk.ok) && (bestT.a == 0 || T.blk.ok), FAIR_ERR_INTERNAL,
                   "staged layout rebuild failed");
        h->blk_tuned = true;
        if (choice) {
            const int32_t c[6] = {bestA.a, bestA.b, bestA.lds, bestT.a, bestT.b, bestT.lds};
            for (int q = 0; q < 6; ++q) choice[q] = c[q];
        }
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

