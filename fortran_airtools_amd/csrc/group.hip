// group.hip -- many independent SIRT problems, each with its own matrix, iterated together.
//
// No counterpart in the reference (its to-do "blocking support", README.md:6): problems of uneven
// size -- say 64 slices from N=128 to N=1024 -- are stepped as one work list.  A group holds, for
// every problem, work items over its A and its A^T: the blocks of their LDS-staged layouts
// (blocked.hip), or runs of 8 rows where a matrix has no such layout, and runs of A^T's empty
// rows (never-hit columns still get their box projection).  One SIRT iteration of every problem
// is two passes, each one launch of a persistent grid of 512-thread workgroups that strides over
// the items (one launch per LDS size class when the layouts differ in footprint capacity):
//   pass 0  r_p[i] = (b_p[i] - a_i.x_p) / M_p[i]                   (AIRMs.f90:1090 inner part)
//   pass 1  x_p[j] = box(x_p[j] + relax * (A_p^T r_p)_j / D_p[j])   (:1090 outer part, :1093-1100)
// A block is processed by the same device function as in the single-matrix kernel (spmv.hpp,
// spmv_block), so every problem's iterate is bit-identical to its own sirt run.  Small problems
// no longer leave the device idle, and an iteration costs a few launches however many problems
// there are.  The persistent grid size (workgroups per CU) is autotuned.
#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "fair_internal.hpp"
#include "spmv.hpp"

namespace fair {
namespace {

// One matrix of a problem (A or A^T) as the group kernel sees it.
struct GMat {
    const int4 *binfo, *brows;   // blocked layout (binfo == nullptr: none)
    const int2 *segs;
    const uint16_t *loc;
    const int4 *rows_nz;         // non-empty rows {row, start, end, 0}
    const int32_t *empty;        // empty rows
    const int32_t *JA;
    const double *vA;
    int32_t ncols, zero_slot;
};

struct GroupProb {
    GMat A, T;
    const double *M, *D;   // D == nullptr: Cimmino (D = 1)
    double *r;             // residual workspace (m)
    const double *b;       // per call
    double *x;             // per call, updated in place
};

struct GBox {
    int upb, lwrb;
    double ub, lb;
    __device__ __forceinline__ double apply(double v) const {   // min(ub) first, then max(lb)
        if (upb) v = fmin_ref(v, ub);
        if (lwrb) v = fmax_ref(v, lb);
        return v;
    }
};

struct GEpiResidual {
    const double *b, *M;
    double *r;
    __device__ __forceinline__ void operator()(int32_t i, double acc) const { r[i] = (b[i] - acc) / M[i]; }
};

struct GEpiUpdate {
    double *x;
    const double *D;
    double relax;
    GBox box;
    __device__ __forceinline__ void operator()(int32_t j, double acc) const {
        double u = relax * acc;
        if (D) u = u / D[j];
        x[j] = box.apply(x[j] + u);
    }
};

// item {problem, kind, a, b}: kind 0 = block a of the blocked layout; 1 = rows_nz[a, b);
// 2 = empty rows [a, b) (zero sum)
enum { ITEM_BLOCK = 0, ITEM_ROWS = 1, ITEM_EMPTY = 2 };
constexpr int ROWS_PER_ITEM = SPMV_BLOCK_THREADS / WAVE;
constexpr int EMPTY_PER_ITEM = SPMV_BLOCK_THREADS;

template <int PASS, class Epi>
__device__ __forceinline__ void group_item(const int4 itm, const GMat &g, const double *src, Epi epi, double *lx) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & (WAVE - 1);
    if (itm.y == ITEM_BLOCK) {
        spmv_block<double, Epi>(g.binfo[itm.z], g.brows, g.segs, g.loc, g.JA, g.vA, g.ncols, g.zero_slot, src, epi,
                                lx);
        __syncthreads();   // the next item's staging reuses the LDS
    } else if (itm.y == ITEM_ROWS) {
        for (int q = itm.z + wv; q < itm.w; q += ROWS_PER_ITEM) {
            const int4 ri = g.rows_nz[q];
            const double acc = row_direct<double>(ri.y, ri.z, lane, g.JA, g.vA, src, g.ncols);
            if (lane == 0) epi(ri.x, acc);
        }
    } else {
        for (int q = itm.z + (int)threadIdx.x; q < itm.w; q += SPMV_BLOCK_THREADS) epi(g.empty[q], 0.0);
    }
}

template <int PASS>
__global__ void __launch_bounds__(SPMV_BLOCK_THREADS, 6) k_group_pass(const int4 *__restrict__ items, int64_t nitems,
                                                                   const GroupProb *__restrict__ P, double relax,
                                                                   GBox box) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *lx = (double *)smem_raw;
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int4 itm = items[it];
        const GroupProb &q = P[itm.x];
        if (PASS == 0)
            group_item<0>(itm, q.A, q.x, GEpiResidual{q.b, q.M, q.r}, lx);
        else
            group_item<1>(itm, q.T, q.r, GEpiUpdate{q.x, q.D, relax, box}, lx);
    }
}

}  // namespace
}  // namespace fair

// Items of one pass that need the same LDS footprint capacity (one launch each).
struct fair_group_class {
    size_t lds_bytes = 8;
    fair::DBuf<int32_t> items;   // int4 each
    int64_t n = 0;
};

struct fair_group {
    int32_t nprob = 0, method = 0;
    std::vector<fair_csr *> mats;
    std::vector<fair::GroupProb> host;
    fair::DBuf<fair::GroupProb> probs;
    std::vector<fair_group_class> passA, passT;
    std::vector<fair::DBuf<double>> r;
    int wg_per_cu = 2;
    int ncu = 1;
};

namespace fair {
namespace {

void check_group(const fair_group *g) { FAIR_CHECK(g != nullptr, FAIR_ERR_ARG, "null group handle"); }

void group_iterate(fair_group *g, int maxits, const fair_opts *o) {
    const double relax = o ? o->relaxpar : 1.0;
    GBox box{0, 0, 0.0, 1.0};
    if (o) {
        box.upb = o->upb != 0;
        box.lwrb = o->lwrb != 0;
        box.ub = o->ubound;
        box.lb = o->lbound;
    }
    // wg_per_cu workgroups per CU stride over a class's items; 0: one workgroup per item, left to
    // the hardware dispatcher
    auto grid_of = [&](const fair_group_class &c) {
        const int64_t g_items = std::max<int64_t>(1, c.n);
        return (unsigned)(g->wg_per_cu > 0 ? std::min<int64_t>(g_items, (int64_t)g->ncu * g->wg_per_cu) : g_items);
    };
    for (int it = 0; it < maxits; ++it) {
        for (const fair_group_class &c : g->passA)
            hipLaunchKernelGGL(k_group_pass<0>, dim3(grid_of(c)), dim3(SPMV_BLOCK_THREADS), c.lds_bytes, stream(),
                               (const int4 *)c.items.p, c.n, g->probs.p, relax, box);
        for (const fair_group_class &c : g->passT)
            hipLaunchKernelGGL(k_group_pass<1>, dim3(grid_of(c)), dim3(SPMV_BLOCK_THREADS), c.lds_bytes, stream(),
                               (const int4 *)c.items.p, c.n, g->probs.p, relax, box);
    }
    launch_check("group sirt");
}

void set_vectors(fair_group *g, const double *const *d_b, double *const *d_x) {
    for (int p = 0; p < g->nprob; ++p) {
        FAIR_CHECK(d_b[p] && d_x[p], FAIR_ERR_ARG, "null vector in group call");
        g->host[p].b = d_b[p];
        g->host[p].x = d_x[p];
    }
    g->probs.upload(g->host.data(), g->host.size());
}

GMat gmat(const Csr &A) {
    GMat m{};
    if (A.blk.ok) {
        m.binfo = (const int4 *)A.blk.binfo.p;
        m.brows = (const int4 *)A.blk.rows.p;
        m.segs = (const int2 *)A.blk.segs.p;
        m.loc = A.blk.loc.p;
        m.zero_slot = A.blk.lds_elems - 1;
    }
    m.rows_nz = (const int4 *)A.rows_nz.p;
    m.empty = A.rows_empty.p;
    m.JA = A.JA.p;
    m.vA = A.vA.p;
    m.ncols = A.n;
    return m;
}

// Work items of one matrix of problem p, filed by the LDS bytes they need (rows and empty rows
// need none).
void add_items(std::map<size_t, std::vector<int4>> &cls, int p, const Csr &A, bool with_empty) {
    if (A.blk.ok) {
        std::vector<int4> &items = cls[sizeof(double) * (size_t)A.blk.lds_elems];
        for (int32_t b = 0; b < A.blk.nblocks; ++b) items.push_back(make_int4(p, ITEM_BLOCK, b, 0));
    } else {
        std::vector<int4> &items = cls[0];
        for (int32_t r = 0; r < A.n_rows_nz; r += ROWS_PER_ITEM)
            items.push_back(make_int4(p, ITEM_ROWS, r, std::min(r + ROWS_PER_ITEM, A.n_rows_nz)));
    }
    if (with_empty) {
        std::vector<int4> &items = cls[0];
        for (int32_t r = 0; r < A.n_rows_empty; r += EMPTY_PER_ITEM)
            items.push_back(make_int4(p, ITEM_EMPTY, r, std::min(r + EMPTY_PER_ITEM, A.n_rows_empty)));
    }
}

void upload_classes(const std::map<size_t, std::vector<int4>> &cls, std::vector<fair_group_class> &out) {
    out.clear();
    out.resize(cls.size());
    size_t q = 0;
    for (const auto &kv : cls) {
        fair_group_class &c = out[q++];
        FAIR_CHECK(kv.first <= (size_t)lds_per_block(), FAIR_ERR_ARG, "block footprint exceeds the LDS");
        c.lds_bytes = std::max<size_t>(kv.first, 8);
        c.n = (int64_t)kv.second.size();
        c.items.alloc(4 * std::max<size_t>(kv.second.size(), 1));
        if (!kv.second.empty())
            FAIR_HIP(hipMemcpyAsync(c.items.p, kv.second.data(), sizeof(int4) * kv.second.size(),
                                    hipMemcpyHostToDevice, stream()));
    }
}

}  // namespace
}  // namespace fair

using namespace fair;

extern "C" int fair_group_create(int32_t nprob, fair_csr *const *As, int32_t method, fair_group **out) {
    return guard([&] {
        FAIR_CHECK(nprob >= 1 && As && out, FAIR_ERR_ARG, "group needs at least one matrix");
        FAIR_CHECK(method == FAIR_CIMMINO || method == FAIR_SART, FAIR_ERR_METHOD, "Not a recognised SIRT method.");
        std::unique_ptr<fair_group> g(new fair_group());
        g->nprob = nprob;
        g->method = method;
        g->mats.assign(As, As + nprob);
        g->host.resize(nprob);
        g->r.resize(nprob);
        std::map<size_t, std::vector<int4>> iA, iT;
        for (int p = 0; p < nprob; ++p) {
            fair_csr *h = As[p];
            FAIR_CHECK(h != nullptr, FAIR_ERR_ARG, "null matrix in group");
            FAIR_CHECK(!h->transposed, FAIR_ERR_ARG, "group matrices must not be transposed");
            sirt_weights(h, method);
            ensure_blocked(h);
            const Csr &A = h->A, &T = h->T;
            g->r[p].alloc(A.m > 0 ? A.m : 1);
            GroupProb &q = g->host[p];
            q = GroupProb{gmat(A), gmat(T), h->M[method].p, method == FAIR_SART ? h->D.p : nullptr, g->r[p].p,
                          nullptr, nullptr};
            add_items(iA, p, A, false);
            add_items(iT, p, T, true);
        }
        upload_classes(iA, g->passA);
        upload_classes(iT, g->passT);
        g->probs.alloc(nprob);
        int dev = 0;
        FAIR_HIP(hipGetDevice(&dev));
        FAIR_HIP(hipDeviceGetAttribute(&g->ncu, hipDeviceAttributeMultiprocessorCount, dev));
        FAIR_HIP(hipStreamSynchronize(stream()));
        *out = g.release();
    });
}

extern "C" int fair_group_free(fair_group *g) {
    return guard([&] {
        (void)hipStreamSynchronize(stream());
        delete g;
    });
}

extern "C" int fair_group_sirt_dev(fair_group *g, const double *const *d_b, int32_t maxits, double *const *d_x,
                                   const fair_opts *opts) {
    return guard([&] {
        check_group(g);
        FAIR_CHECK(d_b && d_x, FAIR_ERR_ARG, "null argument");
        FAIR_CHECK(maxits >= 0, FAIR_ERR_ARG, "negative iteration count");
        set_vectors(g, d_b, d_x);
        group_iterate(g, maxits, opts);
    });
}

// Times one iteration for 1 to 4 persistent workgroups per CU on scratch vectors and keeps the
// fastest.
extern "C" int fair_group_autotune(fair_group *g, int32_t *wg_per_cu) {
    return guard([&] {
        check_group(g);
        std::vector<DBuf<double>> bs(g->nprob), xs(g->nprob);
        std::vector<const double *> pb(g->nprob);
        std::vector<double *> px(g->nprob);
        for (int p = 0; p < g->nprob; ++p) {
            const int32_t m = g->mats[p]->A.m, n = g->mats[p]->A.n;
            bs[p].alloc(m > 0 ? m : 1);
            xs[p].alloc(n > 0 ? n : 1);
            fill(bs[p].p, m, 1.0);
            fill(xs[p].p, n, 0.0);
            pb[p] = bs[p].p;
            px[p] = xs[p].p;
        }
        set_vectors(g, pb.data(), px.data());
        hipEvent_t e0, e1;
        FAIR_HIP(hipEventCreate(&e0));
        FAIR_HIP(hipEventCreate(&e1));
        int best = g->wg_per_cu;
        float best_ms = 1e30f;
        for (int c : {0, 1, 2, 3, 4}) {
            g->wg_per_cu = c;
            group_iterate(g, 1, nullptr);   // warm-up
            FAIR_HIP(hipEventRecord(e0, stream()));
            group_iterate(g, 2, nullptr);
            FAIR_HIP(hipEventRecord(e1, stream()));
            FAIR_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FAIR_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        g->wg_per_cu = best;
        if (wg_per_cu) *wg_per_cu = best;
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}
