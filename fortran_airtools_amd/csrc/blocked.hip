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

int env_int(const char *name, int dflt) {
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

// Group the non-empty rows by key (stable), split groups into blocks of at most rb rows and
// build the footprints.
void build_blocked(Csr &A, const std::vector<int64_t> &key_of_row, int32_t rb, int32_t W, int32_t NB,
                   int32_t lds_elems) {
    Blocked &B = A.blk;
    B = Blocked();
    if (A.n_rows_nz < 0) build_row_list(A);
    const int32_t nr = A.n_rows_nz;
    if (nr <= 0 || NB > 8192 || lds_elems > 65535 || rb < 1) return;
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
// can be switched off with FAIR_SPMV_STAGED=0; FAIR_RB / FAIR_TILE / FAIR_LDS override the
// block shape.
void ensure_blocked(fair_csr *h) {
    if (h->g_n <= 0 || h->transposed || !env_int("FAIR_SPMV_STAGED", 1)) return;
    const bool built = !h->A.blk.ok || !h->have_T || !h->T.blk.ok;
    if (!h->A.blk.ok)
        build_blocked_rays(h->A, h->g_n, h->g_p, env_int("FAIR_DA", 4), env_int("FAIR_DJ", 4),
                           env_int("FAIR_LDS_A", 6144));
    ensure_T(h);
    if (!h->T.blk.ok) {
        const int tx = env_int("FAIR_TX", 8), ty = env_int("FAIR_TY", 8);
        build_blocked_tiles(h->T, h->g_n, h->g_p, h->g_nth, tx, ty, env_int("FAIR_LDS_T", 6144));
    }
    if (env_int("FAIR_DEBUG", 0) && built)
        fprintf(stderr, "[fair] staged SpMV: A %d blocks (%d direct), A^T %d blocks (%d direct)\n",
                h->A.blk.nblocks, h->A.blk.n_direct, h->T.blk.nblocks, h->T.blk.n_direct);
}

}  // namespace fair
