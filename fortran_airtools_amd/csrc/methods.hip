// methods.hip -- the AIRMs methods on the device: SIRT (Cimmino, SART), normest, gradient
// descent and the ART family (Kaczmarz, symmetric, randomized) as a persistent sweep.
//
// SIRT / GD iterate as two fused launches each, all state resident in HBM:
//   launch 1: row-per-wavefront A*x whose epilogue forms the weighted residual
//             (SART/Cimmino: r = (b - Ax)/M ; GD: r = Ax - b)
//   launch 2: row-per-wavefront A^T*r over the device-built CSR of A^T whose epilogue applies
//             the update and the box (SIRT: x + relax*(A^T r)/D ; GD: x - rho/sigma^2 A^T r).
// Gradient descent's stopping rule (||x_new - x|| < tol, AIRMs.f90:972-975) is decided on the
// device: a one-block reduction sets a done flag that gates every later launch, so the whole
// loop is enqueued without host round trips.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include <mutex>

#include "fair_internal.hpp"
#include "spmv.hpp"

namespace fair {
namespace {

struct Box {
    int upb, lwrb;
    double ub, lb;
    // AIRMs.f90:1093-1100 / :963-970: min(ub) first, then max(lb)
    __device__ __forceinline__ double apply(double v) const {
        if (upb) v = fmin_ref(v, ub);
        if (lwrb) v = fmax_ref(v, lb);
        return v;
    }
};

Box make_box(const fair_opts *o) {
    Box b{0, 0, 0.0, 1.0};
    if (o) {
        b.upb = o->upb != 0;
        b.lwrb = o->lwrb != 0;
        b.ub = o->ubound;
        b.lb = o->lbound;
    }
    return b;
}

// r = (b - Ax) / M   (AIRMs.f90:1090, inner part)
struct EpiResidual {
    const double *b, *M;
    double *r;
    __device__ __forceinline__ void operator()(int32_t i, double acc) const { r[i] = (b[i] - acc) / M[i]; }
};

// x_new = box(x + (relax * A^T r) / D)   (AIRMs.f90:1090 outer part + :1093-1100)
struct EpiSirtUpdate {
    const double *x, *D;   // D == nullptr: Cimmino's D = 1
    double *xn;
    double relax;
    Box box;
    __device__ __forceinline__ void operator()(int32_t j, double acc) const {
        double u = relax * acc;
        if (D) u = u / D[j];
        xn[j] = box.apply(x[j] + u);
    }
};

__global__ void k_cimmino_weights(double *M, int32_t m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) M[i] = (double)m * (M[i] * M[i]);   // AIRMs.f90:1072: m * (row_norms(2))**2
}

__global__ void k_copy(const double *a, double *b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_scale_div(double *v, int64_t n, double d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = v[i] / d;
}

__global__ void k_neg(const double *a, double *b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = -a[i];
}

// ---- deterministic reductions -----------------------------------------------------------
// Fixed-shape two-level sum: each of RED_BLOCKS blocks sums a strided slice in a fixed order,
// one block adds the partials.  Same inputs -> same bits.
constexpr int RED_BLOCKS = 512;

template <class F>
__global__ void __launch_bounds__(256) k_red_partial(int64_t n, F f, double *part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * 256)
        s += f(i);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__device__ double block_sum_parts(const double *part) {
    __shared__ double sh[8];
    double s = 0.0;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += 512) s += part[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < 8; ++w) t += sh[w];
    return t;
}

__global__ void __launch_bounds__(512) k_red_final(const double *part, double *out) {
    double t = block_sum_parts(part);
    if (threadIdx.x == 0) *out = t;
}

struct SqOf {
    const double *a;
    __device__ __forceinline__ double operator()(int64_t i) const { return a[i] * a[i]; }
};
struct SqDiff {
    const double *a, *b;
    __device__ __forceinline__ double operator()(int64_t i) const {
        double d = a[i] - b[i];
        return d * d;
    }
};
struct DotOf {
    const double *a, *b;
    __device__ __forceinline__ double operator()(int64_t i) const { return a[i] * b[i]; }
};
// v . (x - y), the difference formed first as the reference's array temporary does
struct DotDiff {
    const double *v, *x, *y;
    __device__ __forceinline__ double operator()(int64_t i) const { return v[i] * (x[i] - y[i]); }
};

// twin_art's mutual-step bookkeeping (AIRMs.f90:694-705): d = new - old; new = new - d
__global__ void k_step_back(double *x, double *v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v[i] = x[i] - v[i];
    x[i] = x[i] - v[i];
}

// a = a / na ; b = b / nb (:716-717)
__global__ void k_div2(double *a, double na, double *b, double nb, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = b[i] / nb;
    a[i] = a[i] / na;
}

// x = x + alpha v ; y = y + beta w (:726-727)
__global__ void k_axpy2(double *x, const double *v, double alpha, double *y, const double *w, double beta,
                        int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = x[i] + alpha * v[i];
    y[i] = y[i] + beta * w[i];
}

// out = .5 * (a + b) (:791-795)
__global__ void k_half_sum(const double *a, const double *b, double *out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = .5 * (a[i] + b[i]);
}

template <class F>
double dev_sum(int64_t n, F f, DBuf<double> &part, DBuf<double> &out) {
    hipLaunchKernelGGL(k_red_partial<F>, dim3(RED_BLOCKS), dim3(256), 0, stream(), n, f, part.p);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(512), 0, stream(), part.p, out.p);
    launch_check("reduction");
    double h = 0.0;
    out.download(&h, 1);
    FAIR_HIP(hipStreamSynchronize(stream()));
    return h;
}

// ---- gradient descent with a device-side stopping flag ------------------------------------
struct GdState {
    int32_t iter;   // accepted updates so far; the current iterate is xbuf[iter & 1]
    int32_t done;
};

struct SrcGated {
    const GdState *st;
    const double *x;
    __device__ __forceinline__ const double *operator()() const { return st->done ? nullptr : x; }
};
struct SrcParity {
    const GdState *st;
    const double *x0, *x1;
    __device__ __forceinline__ const double *operator()() const {
        return st->done ? nullptr : ((st->iter & 1) ? x1 : x0);
    }
};

// xn = box(x - step * A^T r); d2 = (xn - x)^2   (AIRMs.f90:960-970)
struct EpiGdUpdate {
    const GdState *st;
    double *x0, *x1, *d2;
    double step;
    Box box;
    __device__ __forceinline__ void operator()(int32_t j, double acc) const {
        const bool odd = st->iter & 1;
        const double *cur = odd ? x1 : x0;
        double *nxt = odd ? x0 : x1;
        double v = box.apply(cur[j] - step * acc);
        nxt[j] = v;
        double d = v - cur[j];
        d2[j] = d * d;
    }
};

// r = A x - b   (AIRMs.f90:979)
struct EpiGdResidual {
    const double *b;
    double *r;
    __device__ __forceinline__ void operator()(int32_t i, double acc) const { r[i] = acc - b[i]; }
};

struct D2Of {
    const double *d2;
    __device__ __forceinline__ double operator()(int64_t i) const { return d2[i]; }
};

__global__ void __launch_bounds__(256) k_gd_partial(int64_t n, const double *d2, GdState *st, double *part) {
    if (st->done) return;
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * 256)
        s += d2[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void __launch_bounds__(512) k_gd_decide(const double *part, GdState *st, double tol) {
    if (st->done) return;
    double t = block_sum_parts(part);
    if (threadIdx.x == 0) {
        if (sqrt(t) < tol) st->done = 1;
        else st->iter += 1;
    }
}

// ---- ART: persistent row-ordered sweep ------------------------------------------------------
// One wavefront walks the row sequence strictly in order (no barriers: its memory operations
// are issued in program order, so a row's x stores precede the next row's x loads).  Per row:
// dot a_i.x (AIRMs.f90:841) reduced with an xor butterfly, which leaves the identical sum in
// every lane; coeff = relax*(b_i - dot)/nrA_i (:842); then x(JA) += coeff*vA with max(lb) /
// min(ub) on the touched entries (sparse_matrices.f90:808-820).  Column indices are unique
// within a row, so each entry has one owner lane.
// Latency hiding: the entries of row s+1 (up to 64*R of them, R the template argument) are loaded
// into registers while row s is processed -- issued after row s's x gathers, so the in-order load
// counter lets those gathers complete first -- and row descriptors run two steps ahead.  Longer
// rows load the remainder inline.
constexpr int SWEEP_THREADS = WAVE;
constexpr int ART_R = 8;

struct ArtRow {
    int32_t r, beg, end;
    double b, w;
};

// Loads are unconditional (clamped to valid addresses, results selected afterwards) so that the
// compiler can wait for exactly the loads it needs instead of for all outstanding ones.
__device__ __forceinline__ ArtRow art_row(const int32_t *rows, int64_t s, int64_t nsteps, const int32_t *IA,
                                          const double *b, const double *nrA) {
    ArtRow d;
    d.r = rows[s < nsteps ? s : nsteps - 1] - 1;
    d.beg = IA[d.r];
    d.end = s < nsteps ? IA[d.r + 1] : d.beg;
    d.b = b[d.r];
    d.w = nrA[d.r];
    return d;
}

// Up to two independent chains (row sequence, iterate), one workgroup each: twin_art sweeps its
// x and y chains in the same launch.
struct SweepChains {
    const int32_t *rows[2];
    int64_t nsteps[2];
    double *x[2];
};

template <int R>
__global__ void __launch_bounds__(SWEEP_THREADS) k_art_sweep(SweepChains ch, const int32_t *__restrict__ IA,
                                                             const int32_t *__restrict__ JA,
                                                             const double *__restrict__ vA,
                                                             const double *__restrict__ b,
                                                             const double *__restrict__ nrA,
                                                             int32_t ncols, double relax, int lwrb,
                                                             double lb, int upb, double ub) {
    const int lane = threadIdx.x;
    const int32_t *__restrict__ rows = ch.rows[blockIdx.x];
    const int64_t nsteps = ch.nsteps[blockIdx.x];
    double *x = ch.x[blockIdx.x];
    if (nsteps <= 0) return;
    ArtRow cur = art_row(rows, 0, nsteps, IA, b, nrA);
    ArtRow nxt = art_row(rows, 1, nsteps, IA, b, nrA);
    int32_t jc[R];
    double vc[R];
    // JA / vA always have CSR_PAD spare entries, so index 0 is a valid address even for nnz = 0
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int32_t k = cur.beg + lane + WAVE * u;
        const int32_t kk = k < cur.end ? k : 0;
        const int32_t jv = JA[kk];
        const double vv = vA[kk];
        jc[u] = k < cur.end ? jv : -1;
        vc[u] = k < cur.end ? vv : 0.0;
    }
    for (int64_t s = 0; s < nsteps; ++s) {
        // x gathers of this row (after the previous row's stores)
        double xv[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const bool ok = (uint32_t)jc[u] < (uint32_t)ncols;
            const double xx = x[ok ? jc[u] : 0];
            xv[u] = ok ? xx : 0.0;
        }
        // prefetch: entries of the next row, descriptor of the row after
        int32_t jn[R];
        double vn[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int32_t k = nxt.beg + lane + WAVE * u;
            const int32_t kk = k < nxt.end ? k : 0;
            const int32_t jv = JA[kk];
            const double vv = vA[kk];
            jn[u] = k < nxt.end ? jv : -1;
            vn[u] = k < nxt.end ? vv : 0.0;
        }
        const ArtRow nn = art_row(rows, s + 2, nsteps, IA, b, nrA);
        double dot = 0.0;
#pragma unroll
        for (int u = 0; u < R; ++u) dot += xv[u] * vc[u];
        for (int32_t k = cur.beg + lane + WAVE * R; k < cur.end; k += WAVE) {
            const int j = JA[k];
            if ((uint32_t)j < (uint32_t)ncols) dot += x[j] * vA[k];
        }
        dot = wave_sum_dpp(dot);
        const double coeff = (relax * (cur.b - dot)) / cur.w;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            if ((uint32_t)jc[u] < (uint32_t)ncols) {
                double v = xv[u] + coeff * vc[u];
                if (lwrb) v = fmax_ref(v, lb);
                if (upb) v = fmin_ref(v, ub);
                x[jc[u]] = v;
            }
        }
        for (int32_t k = cur.beg + lane + WAVE * R; k < cur.end; k += WAVE) {
            const int j = JA[k];
            if ((uint32_t)j >= (uint32_t)ncols) continue;
            double v = x[j] + coeff * vA[k];
            if (lwrb) v = fmax_ref(v, lb);
            if (upb) v = fmin_ref(v, ub);
            x[j] = v;
        }
        cur = nxt;
        nxt = nn;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            jc[u] = jn[u];
            vc[u] = vn[u];
        }
    }
}

// k_art_sweep with R entries per lane in registers: 8 (512 per row), or 16 for a matrix with
// longer rows (an N = 512 image has rows of about 1000 entries), which then need no loop.  A lane
// adds its entries in increasing order either way, so the iterate does not depend on R.
template <class... Args>
void launch_art_sweep(Csr &A, unsigned chains, Args... args) {
    if (A.n_rows_nz < 0) build_row_list(A);
    if (A.max_row > 8 * WAVE)
        hipLaunchKernelGGL(k_art_sweep<16>, dim3(chains), dim3(SWEEP_THREADS), 0, stream(), args...);
    else
        hipLaunchKernelGGL(k_art_sweep<8>, dim3(chains), dim3(SWEEP_THREADS), 0, stream(), args...);
}

// Row steps of up to two chains without a level schedule: the window sweep (artwin.hip) where it
// applies, else the one-wavefront sweep; both give the same iterate bit for bit.  deferred: the
// window sweep's stop flag is cleared and checked by the caller (art_window_clear / _check), and
// the call does not synchronise.
void launch_art_chains(Csr &A, int chains, const int32_t *r0, int64_t n0, double *x0, const int32_t *r1, int64_t n1,
                       double *x1, const double *b, const double *nrA, double relax, const Box &box,
                       bool deferred = false) {
    if (art_window_usable(A, std::max(n0, n1))) {
        const int32_t *const rows[2] = {r0, r1};
        const int64_t ns[2] = {n0, n1};
        double *const xs[2] = {x0, x1};
        launch_art_window(A, chains, rows, ns, xs, b, nrA, relax, box.lwrb, box.lb, box.upb, box.ub, deferred);
        return;
    }
    launch_art_sweep(A, (unsigned)chains, SweepChains{{r0, r1}, {n0, n1}, {x0, x1}}, A.IA.p, A.JA.p, A.vA.p, b, nrA,
                     A.n, relax, box.lwrb, box.lb, box.upb, box.ub);
    launch_check("art sweep");
}

// ---- level-parallel sweep ----------------------------------------------------------------
// The same row steps as k_art_sweep, in an order that keeps every pixel's update sequence: the
// steps of one level (LevelSched) touch disjoint pixels.  Two executors: one launch per level
// (k_art_level below, replayed from a captured graph; the default), and this one, in which the
// LEV_WAVES wavefronts of one workgroup take a level's steps side by side (wavefront w: steps
// off[l] + w, + LEV_WAVES, ...) and a workgroup barrier separates the levels
// (FAIR_ART_LEVEL_MODE=workgroup; 13.3 us per level against 3.7 us at N = 256, as one CU issues
// every gather and store of a level).  One workgroup per chain keeps each chain's stores and
// loads on one CU, where the barrier orders them.  Each row is computed exactly as k_art_sweep
// computes it (entries of lane q: beg + q + 64 t in increasing t, the same DPP reduction and
// update), so the iterate equals k_art_sweep's bit for bit.  A wavefront loads the entries of its
// next step (possibly a level ahead) while it processes the current one.
constexpr int LEV_WAVES = 16;
constexpr int LEV_THREADS = LEV_WAVES * WAVE;

struct LevelChains {
    const int32_t *rows[2];
    const int32_t *off[2];
    int32_t nlev[2];
    double *x[2];
};

// Steps the wavefront's (level, step) to its next step: the next one of this wavefront in level
// l, else its slot in the next level wide enough to have one; l == nlev when none is left.
__device__ __forceinline__ void lev_advance(const int32_t *__restrict__ off, int32_t nlev, int32_t wv,
                                            int32_t &l, int32_t &s) {
    s += LEV_WAVES;
    while (l < nlev && s >= off[l + 1]) {
        ++l;
        s = off[l] + wv;
    }
}

__device__ __forceinline__ ArtRow lev_row(const int32_t *rows, bool ok, int32_t s, const int32_t *IA,
                                          const double *b, const double *nrA) {
    ArtRow d;
    d.r = rows[ok ? s : 0] - 1;
    d.beg = IA[d.r];
    d.end = ok ? IA[d.r + 1] : d.beg;
    d.b = b[d.r];
    d.w = nrA[d.r];
    return d;
}

__global__ void __launch_bounds__(LEV_THREADS) k_art_levels(LevelChains ch, const int32_t *__restrict__ IA,
                                                            const int32_t *__restrict__ JA,
                                                            const double *__restrict__ vA,
                                                            const double *__restrict__ b,
                                                            const double *__restrict__ nrA,
                                                            int32_t ncols, double relax, int lwrb,
                                                            double lb, int upb, double ub) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int32_t wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int32_t *__restrict__ rows = ch.rows[blockIdx.x];
    const int32_t *__restrict__ off = ch.off[blockIdx.x];
    const int32_t nlev = ch.nlev[blockIdx.x];
    double *x = ch.x[blockIdx.x];
    if (nlev <= 0) return;                 // the whole workgroup leaves together
    int32_t lc = 0, sc = wv - LEV_WAVES;
    lev_advance(off, nlev, wv, lc, sc);
    int32_t ln = lc, sn = sc;
    lev_advance(off, nlev, wv, ln, sn);
    ArtRow cur = lev_row(rows, lc < nlev, sc, IA, b, nrA);
    ArtRow nxt = lev_row(rows, ln < nlev, sn, IA, b, nrA);
    int32_t jc[ART_R];
    double vc[ART_R];
#pragma unroll
    for (int u = 0; u < ART_R; ++u) {
        const int32_t k = cur.beg + lane + WAVE * u;
        const int32_t kk = k < cur.end ? k : 0;
        const int32_t jv = JA[kk];
        const double vv = vA[kk];
        jc[u] = k < cur.end ? jv : -1;
        vc[u] = k < cur.end ? vv : 0.0;
    }
    for (int32_t l = 0; l < nlev; ++l) {
        while (lc == l) {
            double xv[ART_R];
#pragma unroll
            for (int u = 0; u < ART_R; ++u) {
                const bool ok = (uint32_t)jc[u] < (uint32_t)ncols;
                const double xx = x[ok ? jc[u] : 0];
                xv[u] = ok ? xx : 0.0;
            }
            int32_t jn[ART_R];
            double vn[ART_R];
#pragma unroll
            for (int u = 0; u < ART_R; ++u) {
                const int32_t k = nxt.beg + lane + WAVE * u;
                const int32_t kk = k < nxt.end ? k : 0;
                const int32_t jv = JA[kk];
                const double vv = vA[kk];
                jn[u] = k < nxt.end ? jv : -1;
                vn[u] = k < nxt.end ? vv : 0.0;
            }
            int32_t l2 = ln, s2 = sn;
            lev_advance(off, nlev, wv, l2, s2);
            const ArtRow nn = lev_row(rows, l2 < nlev, s2, IA, b, nrA);
            double dot = 0.0;
#pragma unroll
            for (int u = 0; u < ART_R; ++u) dot += xv[u] * vc[u];
            for (int32_t k = cur.beg + lane + WAVE * ART_R; k < cur.end; k += WAVE) {
                const int j = JA[k];
                if ((uint32_t)j < (uint32_t)ncols) dot += x[j] * vA[k];
            }
            dot = wave_sum_dpp(dot);
            const double coeff = (relax * (cur.b - dot)) / cur.w;
#pragma unroll
            for (int u = 0; u < ART_R; ++u) {
                if ((uint32_t)jc[u] < (uint32_t)ncols) {
                    double v = xv[u] + coeff * vc[u];
                    if (lwrb) v = fmax_ref(v, lb);
                    if (upb) v = fmin_ref(v, ub);
                    x[jc[u]] = v;
                }
            }
            for (int32_t k = cur.beg + lane + WAVE * ART_R; k < cur.end; k += WAVE) {
                const int j = JA[k];
                if ((uint32_t)j >= (uint32_t)ncols) continue;
                double v = x[j] + coeff * vA[k];
                if (lwrb) v = fmax_ref(v, lb);
                if (upb) v = fmin_ref(v, ub);
                x[j] = v;
            }
            cur = nxt;
            lc = ln;
            nxt = nn;
            ln = l2;
            sn = s2;
#pragma unroll
            for (int u = 0; u < ART_R; ++u) {
                jc[u] = jn[u];
                vc[u] = vn[u];
            }
        }
        __syncthreads();
    }
}

// One level of a schedule as its own launch: one wavefront per step (steps s0..s1 of the level,
// which touch disjoint pixels), spread over the whole device; the launches of consecutive levels
// are ordered by the stream (a captured graph), which also makes each level's stores visible to
// the next.  Each row is computed as k_art_sweep computes it.
constexpr int LEVK_THREADS = WAVE;
// R entries per lane are held in registers (64 R per row; longer rows finish in a loop).  Each
// level is instantiated for its longest row: R = 8 up to 512 entries, R = 16 (a wavefront alone
// in its workgroup has the registers) above, so the rows of an N = 512 image (up to about 1000
// entries) need no loop while short rows do not pay for empty slots.  A lane adds its entries in
// increasing order either way, so the sums do not depend on R.

template <int LEVK_R>
__global__ void __launch_bounds__(LEVK_THREADS) k_art_level(const ArtParams *__restrict__ P,
                                                            const int4 *__restrict__ desc, int32_t s0,
                                                            int32_t s1, const int32_t *__restrict__ JA,
                                                            const double *__restrict__ vA, int32_t ncols) {
    const int lane = threadIdx.x;
    const int32_t s = s0 + (int32_t)blockIdx.x;
    if (s >= s1) return;
    // the step's row and entry range in one load (the schedule holds them in level order)
    const int4 d = desc[s];
    const int32_t r = d.x, beg = d.y, end = d.z;
    double *x = P->x;
    const double relax = P->relax, lb = P->lb, ub = P->ub;
    const int lwrb = P->lwrb, upb = P->upb;
    const double br = P->b[r], wr = P->nrA[r];
    int32_t jc[LEVK_R];
    double vc[LEVK_R], xv[LEVK_R];
#pragma unroll
    for (int u = 0; u < LEVK_R; ++u) {
        const int32_t k = beg + lane + WAVE * u;
        const int32_t kk = k < end ? k : 0;
        const int32_t jv = JA[kk];
        const double vv = vA[kk];
        jc[u] = k < end ? jv : -1;
        vc[u] = k < end ? vv : 0.0;
    }
#pragma unroll
    for (int u = 0; u < LEVK_R; ++u) {
        const bool ok = (uint32_t)jc[u] < (uint32_t)ncols;
        const double xx = x[ok ? jc[u] : 0];
        xv[u] = ok ? xx : 0.0;
    }
    double dot = 0.0;
#pragma unroll
    for (int u = 0; u < LEVK_R; ++u) dot += xv[u] * vc[u];
    for (int32_t k = beg + lane + WAVE * LEVK_R; k < end; k += WAVE) {
        const int j = JA[k];
        if ((uint32_t)j < (uint32_t)ncols) dot += x[j] * vA[k];
    }
    dot = wave_sum_dpp(dot);
    const double coeff = (relax * (br - dot)) / wr;
#pragma unroll
    for (int u = 0; u < LEVK_R; ++u) {
        if ((uint32_t)jc[u] < (uint32_t)ncols) {
            double v = xv[u] + coeff * vc[u];
            if (lwrb) v = fmax_ref(v, lb);
            if (upb) v = fmin_ref(v, ub);
            x[jc[u]] = v;
        }
    }
    for (int32_t k = beg + lane + WAVE * LEVK_R; k < end; k += WAVE) {
        const int j = JA[k];
        if ((uint32_t)j >= (uint32_t)ncols) continue;
        double v = x[j] + coeff * vA[k];
        if (lwrb) v = fmax_ref(v, lb);
        if (upb) v = fmin_ref(v, ub);
        x[j] = v;
    }
}

hipStream_t g_level_capture = nullptr;

// Writes a replay's arguments into the schedule's parameter block, in stream order (launch
// arguments are copied at launch, so no host buffer has to outlive the call).
__global__ void k_set_art_params(ArtParams *P, ArtParams v) { *P = v; }

int env_flag(const char *name, int dflt) {
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

// Levels of a row sequence (one-based rows) over a CSR pattern whose column c is JA[k] - base
// (columns outside 0..n-1 are ignored, as the sweeps ignore them): step s gets level
// 1 + max(level of the last earlier step touching one of its pixels).  Writes the steps sorted
// by (level, position) to ordered (S) and the level offsets to off (levels + 1 <= S + 1 entries);
// returns the number of levels.
//
// The scan keeps each pixel's last level in `last`, n slots plus a sentinel slot n that columns
// outside the image read (as 0) and write, so the loops have no branches; 16-bit levels while they
// fit (half the cache footprint), 32-bit from the start again when they do not.  Built also for
// AVX2 (8-wide gathers in the max pass) and chosen at run time.
template <class T>
__attribute__((always_inline)) inline int64_t level_scan(int32_t n, const int32_t *IA, const int32_t *JA,
                                                         int32_t base, const int32_t *seq, int64_t S,
                                                         int32_t *lev, T *last, int64_t cap) {
    const uint32_t un = (uint32_t)n;
    int64_t L = 0;
    for (int64_t s = 0; s < S; ++s) {
        const int32_t r = seq[s] - 1;
        const int32_t b = IA[r], e = IA[r + 1];
        uint32_t mx = 0;
        for (int32_t k = b; k < e; ++k) {
            uint32_t j = (uint32_t)(JA[k] - base);
            j = j < un ? j : un;
            const uint32_t v = last[j];
            mx = v > mx ? v : mx;
        }
        ++mx;
        if ((int64_t)mx > cap) return -1;
        for (int32_t k = b; k < e; ++k) {
            uint32_t j = (uint32_t)(JA[k] - base);
            j = j < un ? j : un;
            last[j] = (T)mx;
        }
        last[un] = 0;
        lev[s] = (int32_t)mx;
        L = std::max<int64_t>(L, mx);
    }
    return L;
}

template <class T>
int64_t level_scan_plain(int32_t n, const int32_t *IA, const int32_t *JA, int32_t base, const int32_t *seq,
                         int64_t S, int32_t *lev, T *last, int64_t cap) {
    return level_scan<T>(n, IA, JA, base, seq, S, lev, last, cap);
}

template <class T>
__attribute__((target("avx2"))) int64_t level_scan_avx2(int32_t n, const int32_t *IA, const int32_t *JA,
                                                        int32_t base, const int32_t *seq, int64_t S, int32_t *lev,
                                                        T *last, int64_t cap) {
    return level_scan<T>(n, IA, JA, base, seq, S, lev, last, cap);
}

template <class T>
int64_t level_scan_best(int32_t n, const int32_t *IA, const int32_t *JA, int32_t base, const int32_t *seq,
                        int64_t S, int32_t *lev, int64_t cap) {
    std::vector<T> last((size_t)std::max(n, 0) + 1, 0);
    static const bool avx2 = __builtin_cpu_supports("avx2");
    return avx2 ? level_scan_avx2<T>(n, IA, JA, base, seq, S, lev, last.data(), cap)
                : level_scan_plain<T>(n, IA, JA, base, seq, S, lev, last.data(), cap);
}

int32_t levelize(int32_t n, const int32_t *IA, const int32_t *JA, int32_t base, const int32_t *seq, int64_t S,
                 int32_t *ordered, int32_t *off) {
    std::vector<int32_t> lev((size_t)S);
    int64_t L = level_scan_best<uint16_t>(n, IA, JA, base, seq, S, lev.data(), 65535);
    if (L < 0) L = level_scan_best<int32_t>(n, IA, JA, base, seq, S, lev.data(), INT32_MAX);
    std::fill(off, off + L + 1, 0);
    for (int64_t s = 0; s < S; ++s) ++off[lev[s]];
    for (int64_t l = 0; l < L; ++l) off[l + 1] += off[l];
    std::vector<int32_t> fill(off, off + L);
    for (int64_t s = 0; s < S; ++s) ordered[fill[lev[s] - 1]++] = seq[s];
    return (int32_t)L;
}

// The level schedule of a row sequence (one-based rows), made on the host from the pattern and
// cached in the matrix; nullptr when it is switched off (FAIR_ART_LEVELS=0) or the sequence has
// fewer than FAIR_ART_LEVEL_WIDTH (4) steps per level on average, where one wavefront walking the
// sequence is as fast.  A long sequence whose first eighth is already that narrow is rejected
// before it is planned in full.  The cache holds six schedules and evicts the least recently used
// one; the returned reference keeps its schedule alive even if a later plan evicts it.
std::shared_ptr<LevelSched> art_levels(Csr &A, const int32_t *seq, int64_t S) {
    if (S <= 0 || S > INT32_MAX || A.n <= 0 || !env_flag("FAIR_ART_LEVELS", 1)) return nullptr;
    const int64_t width = std::max(1, env_flag("FAIR_ART_LEVEL_WIDTH", 4));
    for (size_t i = 0; i < A.lev.size(); ++i) {
        std::shared_ptr<LevelSched> c = A.lev[i];
        if ((int64_t)c->seq.size() == S && std::equal(c->seq.begin(), c->seq.end(), seq)) {
            A.lev.erase(A.lev.begin() + (ptrdiff_t)i);   // a hit becomes the most recently used
            A.lev.push_back(c);
            return S >= width * c->nlev ? c : nullptr;
        }
    }
    const bool dbg = env_flag("FAIR_DEBUG", 0) != 0;
    auto now = [] { return std::chrono::steady_clock::now(); };
    const auto t0 = now();
    if (A.hIA.empty()) {
        A.hIA.alloc((size_t)A.m + 1);
        A.hJA.alloc((size_t)std::max<int64_t>(A.nnz, 1));
        A.IA.download(A.hIA.data(), (size_t)A.m + 1);
        A.JA.download(A.hJA.data(), (size_t)A.nnz);
        FAIR_HIP(hipStreamSynchronize(stream()));
    }
    const auto t1 = now();
    // the host copy is kept for later plans up to 64 M entries (256 MB) and released after this
    // plan above that
    struct Release {
        Csr &A;
        ~Release() {
            if (A.nnz > (int64_t(1) << 26)) {
                A.hIA.release();
                A.hJA.release();
            }
        }
    } release{A};
    const int32_t *IA = A.hIA.data(), *JA = A.hJA.data();
    std::vector<int32_t> off((size_t)S + 1), ordered((size_t)S);
    // a long sequence whose first eighth is too narrow is not planned in full (random orders,
    // e.g. row-probability draws that change from call to call, stay about 2 steps per level;
    // a cyclic order's prefix is already wider than its threshold)
    const int64_t P = std::max<int64_t>(4096, S / 8);
    if (P < S) {
        const int32_t Lp = levelize(A.n, IA, JA, 0, seq, P, ordered.data(), off.data());
        if (P < width * Lp) return nullptr;
    }
    auto c = std::make_shared<LevelSched>();
    c->seq.assign(seq, seq + S);
    const int32_t L = levelize(A.n, IA, JA, 0, seq, S, ordered.data(), off.data());
    const auto t2 = now();
    off.resize((size_t)L + 1);
    c->nlev = L;
    c->rows.alloc((size_t)S);
    c->off.alloc((size_t)L + 1);
    std::vector<int4> desc((size_t)S);
    for (int64_t s = 0; s < S; ++s) {
        const int32_t r = ordered[s] - 1;
        desc[s] = make_int4(r, IA[r], IA[r + 1], 0);
    }
    c->hmaxlen.assign((size_t)L, 0);
    for (int32_t l = 0; l < L; ++l)
        for (int32_t s = off[l]; s < off[l + 1]; ++s)
            c->hmaxlen[l] = std::max(c->hmaxlen[l], desc[s].z - desc[s].y);
    c->desc.alloc(4 * (size_t)S);
    c->rows.upload(ordered.data(), (size_t)S);
    c->off.upload(off.data(), (size_t)L + 1);
    FAIR_HIP(hipMemcpyAsync(c->desc.p, desc.data(), sizeof(int4) * S, hipMemcpyHostToDevice, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
    if (dbg) {
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        fprintf(stderr, "[fair] level plan of %lld steps: pattern copy %.1f ms, levels %.1f ms, descriptors %.1f ms\n",
                (long long)S, ms(t0, t1), ms(t1, t2), ms(t2, now()));
    }
    c->hoff = std::move(off);
    if (A.lev.size() >= 6) A.lev.erase(A.lev.begin());
    A.lev.push_back(c);
    return S >= width * L ? c : nullptr;
}

// The schedule as one launch per level: captured into a graph on first use, replayed on the
// library stream after the call's arguments are copied to params.  At N = 256 (about 44 steps per
// level) a level takes 3.8 us this way against 13.3 us in one workgroup, whose single CU issues
// every x gather and store of the level.
void graph_art_levels(const Csr &A, LevelSched *s, double *x, const double *b, const double *nrA, double relax,
                      const Box &box, hipStream_t st) {
    if (!s->graph) {
        const auto tc = std::chrono::steady_clock::now();
        if (!g_level_capture) FAIR_HIP(hipStreamCreateWithFlags(&g_level_capture, hipStreamNonBlocking));
        s->params.alloc(1);
        FAIR_HIP(hipStreamSynchronize(stream()));
        hipGraph_t g;
        FAIR_HIP(hipStreamBeginCapture(g_level_capture, hipStreamCaptureModeThreadLocal));
        for (int32_t l = 0; l < s->nlev; ++l) {
            const int32_t w = s->hoff[l + 1] - s->hoff[l];
            if (s->hmaxlen[l] <= 8 * WAVE)
                hipLaunchKernelGGL(k_art_level<8>, dim3((unsigned)w), dim3(LEVK_THREADS), 0, g_level_capture,
                                   s->params.p, (const int4 *)s->desc.p, s->hoff[l], s->hoff[l + 1], A.JA.p, A.vA.p,
                                   A.n);
            else
                hipLaunchKernelGGL(k_art_level<16>, dim3((unsigned)w), dim3(LEVK_THREADS), 0, g_level_capture,
                                   s->params.p, (const int4 *)s->desc.p, s->hoff[l], s->hoff[l + 1], A.JA.p, A.vA.p,
                                   A.n);
        }
        hipError_t le = hipGetLastError();
        FAIR_HIP(hipStreamEndCapture(g_level_capture, &g));
        FAIR_HIP(le);
        hipGraphExec_t ex;
        hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FAIR_HIP(ie);
        s->graph = ex;
        if (env_flag("FAIR_DEBUG", 0))
            fprintf(stderr, "[fair] level graph of %d launches captured in %.1f ms\n", s->nlev,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc).count());
    }
    hipLaunchKernelGGL(k_set_art_params, dim3(1), dim3(1), 0, st, s->params.p,
                       ArtParams{x, b, nrA, relax, box.lb, box.ub, box.lwrb, box.upb});
    FAIR_HIP(hipGraphLaunch(s->graph, st));
    launch_check("art level graph");
}

hipStream_t g_level_side = nullptr;
hipEvent_t g_level_ev[2] = {nullptr, nullptr};

// One level-parallel sweep per chain (s1 / x1 may be null): one launch per level per chain, or
// with FAIR_ART_LEVEL_MODE=workgroup one workgroup per chain.
void launch_art_levels(const Csr &A, LevelSched *s0, double *x0, LevelSched *s1, double *x1,
                       const double *b, const double *nrA, double relax, const Box &box) {
    const char *mode = getenv("FAIR_ART_LEVEL_MODE");
    if (!(mode && !strcmp(mode, "workgroup"))) {
        if (env_flag("FAIR_DEBUG", 0))
            fprintf(stderr, "[fair] level sweep (graph): %zu steps in %d levels%s\n", s0->seq.size(), s0->nlev,
                    s1 ? " (two chains)" : "");
        if (!s1 || s1 == s0) {   // one chain, or two chains of one sequence: one after the other
            graph_art_levels(A, s0, x0, b, nrA, relax, box, stream());
            if (s1) graph_art_levels(A, s1, x1, b, nrA, relax, box, stream());
            return;
        }
        // two independent chains (twin_art): the second replays on a side stream alongside the
        // first, both ordered after the caller's work and before whatever follows on the stream
        if (!g_level_side) {
            FAIR_HIP(hipStreamCreateWithFlags(&g_level_side, hipStreamNonBlocking));
            for (auto &e : g_level_ev) FAIR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        FAIR_HIP(hipEventRecord(g_level_ev[0], stream()));
        FAIR_HIP(hipStreamWaitEvent(g_level_side, g_level_ev[0], 0));
        graph_art_levels(A, s1, x1, b, nrA, relax, box, g_level_side);
        graph_art_levels(A, s0, x0, b, nrA, relax, box, stream());
        FAIR_HIP(hipEventRecord(g_level_ev[1], g_level_side));
        FAIR_HIP(hipStreamWaitEvent(stream(), g_level_ev[1], 0));
        return;
    }
    LevelChains lc{{s0->rows.p, s1 ? s1->rows.p : nullptr},
                   {s0->off.p, s1 ? s1->off.p : nullptr},
                   {s0->nlev, s1 ? s1->nlev : 0},
                   {x0, x1}};
    if (env_flag("FAIR_DEBUG", 0))
        fprintf(stderr, "[fair] level sweep: %zu steps in %d levels%s\n", s0->seq.size(), s0->nlev,
                s1 ? " (two chains)" : "");
    hipLaunchKernelGGL(k_art_levels, dim3(s1 ? 2 : 1), dim3(LEV_THREADS), 0, stream(), lc, A.IA.p, A.JA.p,
                       A.vA.p, b, nrA, A.n, relax, box.lwrb, box.lb, box.upb, box.ub);
    launch_check("art level sweep");
}

__global__ void k_art_weights(double *nrA, int32_t m, double add) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) nrA[i] = nrA[i] + add;
}

__global__ void k_square(double *v, int32_t m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) v[i] = v[i] * v[i];
}

void check_handle(const fair_csr *A) { FAIR_CHECK(A != nullptr, FAIR_ERR_ARG, "null matrix handle"); }

// SIRT weights (AIRMs.f90:1070-1083), computed once per matrix and cached in the handle.
// Cimmino: M = m*||a_i||^2, D = 1.  SART: M = row sums, then the reference's double transpose
// (which leaves A's rows column-sorted), D = column sums.
void sirt_weights_impl(fair_csr *h, int method) {
    FAIR_CHECK(method == FAIR_CIMMINO || method == FAIR_SART, FAIR_ERR_METHOD, "Not a recognised SIRT method.");
    Csr &A = h->A;
    if (A.n_rows_nz < 0) build_row_list(A);
    const int32_t m = A.m, n = A.n;
    if (!h->w_ok[method]) {
        h->M[method].alloc(m > 0 ? m : 1);
        if (method == FAIR_CIMMINO) {
            row_norms(A, 2, h->M[method].p);
            hipLaunchKernelGGL(k_cimmino_weights, dim3(grid_for(m, 256)), dim3(256), 0, stream(), h->M[method].p, m);
        } else {
            row_norms(A, 1, h->M[method].p);          // AIRMs.f90:1076
            h->rows_sorted = true;                    // :1077-1079 (downloads show sorted rows)
            ensure_T(h);
            h->D.alloc(n > 0 ? n : 1);
            row_norms(h->T, 1, h->D.p);               // column sums
        }
        launch_check("sirt weights");
        h->w_ok[method] = true;
    }
    ensure_T(h);
}

void sirt_setup(fair_csr *h, int method) {
    sirt_weights_impl(h, method);
    const int32_t m = h->A.m, n = h->A.n;
    ensure_blocked(h);
    if (h->ws_r.n < (size_t)(m > 0 ? m : 1)) h->ws_r.alloc(m > 0 ? m : 1);
    if (h->ws_x.n < (size_t)(n > 0 ? n : 1)) h->ws_x.alloc(n > 0 ? n : 1);
}

// maxits SIRT iterations on d_x (in: start, out: result), two launches per iteration; r (m) and
// xw (n) are the workspaces: the residual and the second iterate.  Every entry of r that A^T
// reads (the non-empty rows) and every entry of xw is written before it is read.
void sirt_iterate(fair_csr *h, int method, const double *d_b, int maxits, double *d_x, const fair_opts *o,
                  double *r, double *xw) {
    const Csr &A = h->A, &T = h->T;
    const int32_t n = A.n;
    const double *M = h->M[method].p;
    const double *D = method == FAIR_SART ? h->D.p : nullptr;
    const double relax = o ? o->relaxpar : 1.0;
    const Box box = make_box(o);
    double *cur = d_x, *nxt = xw;
    for (int it = 0; it < maxits; ++it) {
        spmv_launch(A, cur, EpiResidual{d_b, M, r});
        EpiSirtUpdate up{cur, D, nxt, relax, box};
        spmv_launch(T, r, up);
        if (T.n_rows_empty > 0)
            hipLaunchKernelGGL((k_empty_rows<double, SrcPtr<double>, EpiSirtUpdate>),
                               dim3(grid_for(T.n_rows_empty, 256)), dim3(256), 0, stream(),
                               T.rows_empty.p, T.n_rows_empty, SrcPtr<double>{r}, up);
        std::swap(cur, nxt);
    }
    if (cur != d_x)
        hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256)), dim3(256), 0, stream(), cur, d_x, (int64_t)n);
    launch_check("sirt");
}

void sirt_run(fair_csr *h, int method, const double *d_b, int maxits, double *d_x, const fair_opts *o) {
    sirt_setup(h, method);
    sirt_iterate(h, method, d_b, maxits, d_x, o, h->ws_r.p, h->ws_x.p);
}

// ---- SIRT without device allocation --------------------------------------------------------
// Workspaces of fair_sirt_dev_noalloc: the residual (m doubles) and, from the next 256-byte
// boundary, the second iterate (n doubles).  One array per GPU, held until the process ends;
// every call on a GPU shares it, so calls that overlap on one GPU race.  Host code never uses
// the name: its device address is looked up per device and kept in g_scratch_addr.
constexpr size_t SCRATCH_DOUBLES = size_t(4) << 20;   // 32 MiB
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int SCRATCH_MAX_DEVICES = 64;
double *g_scratch_addr[SCRATCH_MAX_DEVICES] = {};
std::mutex g_scratch_mutex;

size_t scratch_r_len(int32_t m) { return ((size_t)std::max(m, 1) + 31) / 32 * 32; }

}  // namespace

// The scratch array's address on the current device (looked up once per device).
double *sirt_scratch_addr() {
    int dev = 0;
    FAIR_HIP(hipGetDevice(&dev));
    FAIR_CHECK(dev >= 0 && dev < SCRATCH_MAX_DEVICES, FAIR_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        FAIR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FAIR_CHECK(p != nullptr, FAIR_ERR_HIP, "no device address for the SIRT scratch array");
        g_scratch_addr[dev] = static_cast<double *>(p);
    }
    return g_scratch_addr[dev];
}

namespace {

}  // namespace

void sirt_setup_public(fair_csr *h, int method) { sirt_setup(h, method); }
void sirt_weights(fair_csr *h, int method) { sirt_weights_impl(h, method); }

}  // namespace fair

using namespace fair;

extern "C" int fair_sirt_dev(fair_csr *h, int32_t method, const double *d_b, int32_t maxits, double *d_x,
                             const fair_opts *opts) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(maxits >= 0, FAIR_ERR_ARG, "negative iteration count");
        sirt_run(h, method, d_b, maxits, d_x, opts);
    });
}

// fair_sirt_dev on the static scratch array: no allocation, so the matrix must already hold what
// sirt sets up (weights, CSR of A^T, staged layouts: fair_sirt_prepare), and m + n (the residual
// rounded up to 32 entries) must fit the array.
extern "C" int fair_sirt_dev_noalloc(fair_csr *h, int32_t method, const double *d_b, int32_t maxits, double *d_x,
                                     const fair_opts *opts) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(maxits >= 0, FAIR_ERR_ARG, "negative iteration count");
        FAIR_CHECK(method == FAIR_CIMMINO || method == FAIR_SART, FAIR_ERR_METHOD, "Not a recognised SIRT method.");
        const size_t rlen = scratch_r_len(h->A.m);
        FAIR_CHECK(rlen + (size_t)std::max(h->A.n, 1) <= SCRATCH_DOUBLES, FAIR_ERR_ARG,
                   "m + n exceeds the static SIRT scratch (" + std::to_string(SCRATCH_DOUBLES) +
                       " doubles); use fair_sirt_dev");
        FAIR_CHECK(h->w_ok[method] && h->have_T && h->A.n_rows_nz >= 0 && h->T.n_rows_nz >= 0 && !blocked_pending(h),
                   FAIR_ERR_ARG, "matrix not prepared for this method: call fair_sirt_prepare first");
        double *s = sirt_scratch_addr();
        if (maxits > 0) FAIR_CHECK(d_b && d_x, FAIR_ERR_ARG, "null argument");
        sirt_iterate(h, method, d_b, maxits, d_x, opts, s, s + rlen);
    });
}

extern "C" int fair_sirt_prepare(fair_csr *h, int32_t method) {
    return guard([&] {
        check_handle(h);
        sirt_setup(h, method);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_sirt(fair_csr *h, int32_t method, const double *b, int32_t maxits, const double *x0,
                         const fair_opts *opts, double *xout) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(maxits >= 0, FAIR_ERR_ARG, "negative iteration count");
        FAIR_CHECK(b && xout, FAIR_ERR_ARG, "null argument");
        const int32_t m = h->A.m, n = h->A.n;
        DBuf<double> db(m > 0 ? m : 1), dx(n > 0 ? n : 1);
        db.upload(b, m);
        if (x0) dx.upload(x0, n);
        else fill(dx.p, n, 0.0);
        sirt_run(h, method, db.p, maxits, dx.p, opts);
        FAIR_HIP(hipStreamSynchronize(stream()));
        dx.download(xout, n);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_normest(fair_csr *h, const double *v0, double tol, int32_t maxit, double *sigma) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(v0 && sigma, FAIR_ERR_ARG, "null argument");
        Csr &A = h->A;
        if (A.n_rows_nz < 0) build_row_list(A);
        ensure_T(h);
        ensure_blocked(h);
        const int32_t m = A.m, n = A.n;
        DBuf<double> v(n > 0 ? n : 1), vn(n > 0 ? n : 1), w(m > 0 ? m : 1), part(RED_BLOCKS), out(1);
        v.upload(v0, n);
        // AIRMs.f90:1159-1167
        for (int i = 1; i <= maxit; ++i) {
            spmv(A, v.p, w.p);
            spmv(h->T, w.p, vn.p);
            double nv = std::sqrt(dev_sum(n, SqOf{vn.p}, part, out));
            hipLaunchKernelGGL(k_scale_div, dim3(grid_for(n, 256)), dim3(256), 0, stream(), vn.p, (int64_t)n, nv);
            double dd = std::sqrt(dev_sum(n, SqDiff{vn.p, v.p}, part, out));
            if (dd < tol) break;
            std::swap(v.p, vn.p);
        }
        spmv(A, v.p, w.p);                                   // :1169-1171
        double ww = dev_sum(m, SqOf{w.p}, part, out);
        double vv = dev_sum(n, SqOf{v.p}, part, out);
        *sigma = std::sqrt(ww / vv);
    });
}

extern "C" int fair_gradient_descent(fair_csr *h, const double *b, int32_t maxiter, const double *x0,
                                     double rho, double tol, double sigma, const fair_opts *opts,
                                     double *xout, int32_t *iters) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(b && xout, FAIR_ERR_ARG, "null argument");
        FAIR_CHECK(tol >= 0.0, FAIR_ERR_ARG, "Must use positive tolerance.");
        Csr &A = h->A;
        if (A.n_rows_nz < 0) build_row_list(A);
        ensure_T(h);
        ensure_blocked(h);
        const Csr &T = h->T;
        const int32_t m = A.m, n = A.n;
        DBuf<double> db(m > 0 ? m : 1), r(m > 0 ? m : 1), x0b(n > 0 ? n : 1), x1b(n > 0 ? n : 1),
            d2(n > 0 ? n : 1), part(RED_BLOCKS);
        DBuf<GdState> st(1);
        db.upload(b, m);
        if (x0) x0b.upload(x0, n);
        else fill(x0b.p, n, 0.0);
        GdState s0{0, 0};
        st.upload(&s0, 1);
        const double s2 = sigma * sigma;                     // AIRMs.f90:953
        const double step = rho / s2;                        // rho/sigma * A^T r, :960
        const Box box = make_box(opts);
        if (m > 0)
            hipLaunchKernelGGL(k_neg, dim3(grid_for(m, 256)), dim3(256), 0, stream(), db.p, r.p, (int64_t)m);
        for (int it = 1; it <= maxiter; ++it) {
            EpiGdUpdate up{st.p, x0b.p, x1b.p, d2.p, step, box};
            SrcGated src{st.p, r.p};
            spmv_launch_src(T, src, up);
            if (T.n_rows_empty > 0)
                hipLaunchKernelGGL((k_empty_rows<double, SrcGated, EpiGdUpdate>),
                                   dim3(grid_for(T.n_rows_empty, 256)), dim3(256), 0, stream(),
                                   T.rows_empty.p, T.n_rows_empty, src, up);
            hipLaunchKernelGGL(k_gd_partial, dim3(RED_BLOCKS), dim3(256), 0, stream(), (int64_t)n, d2.p, st.p, part.p);
            hipLaunchKernelGGL(k_gd_decide, dim3(1), dim3(512), 0, stream(), part.p, st.p, tol);
            if (it < maxiter)
                spmv_launch_src(A, SrcParity{st.p, x0b.p, x1b.p}, EpiGdResidual{db.p, r.p});
        }
        launch_check("gradient descent");
        GdState fin;
        st.download(&fin, 1);
        FAIR_HIP(hipStreamSynchronize(stream()));
        ((fin.iter & 1) ? x1b : x0b).download(xout, n);
        FAIR_HIP(hipStreamSynchronize(stream()));
        if (iters) *iters = fin.iter;
    });
}

// ---- ART --------------------------------------------------------------------------------------
namespace {

struct ArtPlan {
    int method;
    std::vector<int32_t> Kuse, ord;
    int nord = 0;
    bool isrand = false, rowprob = false;
    std::vector<double> nrA, cumul;
    int maxiter = 0;
};

// Ordering, damping, row filter and row-probability table of art (AIRMs.f90:217-312).
ArtPlan plan_art(fair_csr *h, int method, int nk, const int32_t *Kin, const int32_t *order, int nord_in,
                 const fair_opts *o, DBuf<double> &d_nrA) {
    FAIR_CHECK(method >= FAIR_KACZMARZ && method <= FAIR_RANDKACZMARZ, FAIR_ERR_METHOD,
               "Unknown ART method. Use kaczmarz, symkaczmarz or randkaczmarz");
    check_unique_columns(h);
    FAIR_CHECK(nk >= 1 && Kin, FAIR_ERR_ARG, "Kin must not be empty");
    Csr &A = h->A;
    if (A.n_rows_nz < 0) build_row_list(A);
    const int32_t m = A.m;
    ArtPlan P;
    P.method = method;
    P.isrand = method == FAIR_RANDKACZMARZ;
    P.rowprob = P.isrand && o && o->rowprobability;
    P.Kuse.assign(Kin, Kin + nk);
    if (method == FAIR_SYMKACZMARZ) {
        for (int k = 0; k < nk; ++k) FAIR_CHECK(Kin[k] % 2 == 0, FAIR_ERR_ARG, "Kin should all be multiples of 2.");
        for (auto &k : P.Kuse) k /= 2;
    }
    std::vector<int32_t> ord;
    if (order) {
        FAIR_CHECK(nord_in >= 0 && nord_in <= m, FAIR_ERR_ARG, "order longer than the number of rows");
        ord.assign(order, order + nord_in);
        for (int32_t r : ord) FAIR_CHECK(r >= 1 && r <= m, FAIR_ERR_ARG, "order entry out of range");
    } else {
        ord.resize(m);
        for (int32_t i = 0; i < m; ++i) ord[i] = i + 1;
    }
    // nrA = row_norms(2)**2 + damp*maxval (AIRMs.f90:273-275)
    d_nrA.alloc(m > 0 ? m : 1);
    row_norms(A, 2, d_nrA.p);
    if (m > 0) hipLaunchKernelGGL(k_square, dim3(grid_for(m, 256)), dim3(256), 0, stream(), d_nrA.p, m);
    P.nrA.resize(m);
    d_nrA.download(P.nrA.data(), m);
    FAIR_HIP(hipStreamSynchronize(stream()));
    const double dmp = o ? o->damp : 0.0;
    if (dmp != 0.0 && m > 0) {
        double mx = P.nrA[0];
        for (double v : P.nrA) mx = v > mx ? v : mx;
        const double add = dmp * mx;
        for (double &v : P.nrA) v = v + add;
        d_nrA.upload(P.nrA.data(), m);
    }
    // keep rows with nrA > 0 (:279-287)
    std::vector<int32_t> kept;
    kept.reserve(ord.size());
    for (int32_t r : ord)
        if (P.nrA[r - 1] > 0) kept.push_back(r);
    P.nord = (int)kept.size();
    P.ord = kept;
    if (P.rowprob) {                                   // :302-308
        P.cumul.assign(m, 1.0);
        double cs = 0.0, tot = 0.0;
        for (int i = 0; i < P.nord; ++i) { cs = cs + P.nrA[i]; P.cumul[i] = cs; }
        for (int i = 0; i < m; ++i) tot = tot + P.nrA[i];
        for (int i = 0; i < P.nord; ++i) P.cumul[i] = P.cumul[i] / tot;
    }
    P.maxiter = 0;
    for (int k : P.Kuse) P.maxiter = std::max(P.maxiter, k);
    return P;
}

int64_t uniforms_needed(const ArtPlan &P) {
    if (!P.isrand) return 0;
    return 2 * (int64_t)P.nord + (int64_t)P.maxiter * (2 * (int64_t)P.nord + (P.rowprob ? P.nord : 0));
}

// scramble (AIRMs.f90:1175-1212): two passes of "swap i with a uniform j", values 1..nord
void scramble(int nord, const double *u, int32_t *out) {
    for (int i = 0; i < nord; ++i) out[i] = i + 1;
    int64_t q = 0;
    for (int k = 0; k < 2; ++k)
        for (int i = 1; i <= nord; ++i) {
            int j = 1 + (int)std::floor((double)nord * u[q++]);
            std::swap(out[j - 1], out[i - 1]);
        }
}

}  // namespace

extern "C" int fair_art_uniform_count(fair_csr *h, int32_t method, int32_t nk, const int32_t *Kin,
                                      const int32_t *order, int32_t nord, const fair_opts *opts,
                                      int64_t *count) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(count != nullptr, FAIR_ERR_ARG, "null argument");
        DBuf<double> d_nrA;
        ArtPlan P = plan_art(h, method, nk, Kin, order, nord, opts, d_nrA);
        *count = uniforms_needed(P);
    });
}

// Upper bound of the draws fair_twin_art can consume (it stops early by its rule).
extern "C" int fair_twin_art_uniform_bound(fair_csr *h, int32_t method, int32_t nk, const int32_t *Kin,
                                           const int32_t *order, int32_t nord, const fair_opts *opts,
                                           int64_t *count) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(count != nullptr, FAIR_ERR_ARG, "null argument");
        DBuf<double> d_nrA;
        ArtPlan P = plan_art(h, method, nk, Kin, order, nord, opts, d_nrA);
        *count = P.isrand ? ((int64_t)P.maxiter + 1) * (2 * (int64_t)P.nord + (P.rowprob ? 2 * (int64_t)P.nord : 0))
                          : 0;
    });
}

// The level schedule of a row sequence over a host CSR pattern (one-based JA, as at the boundary);
// host work only, no device call.
extern "C" int fair_level_schedule(int32_t m, int32_t n, const int32_t *IA, const int32_t *JA, const int32_t *seq,
                                   int64_t nsteps, int32_t *rows_out, int32_t *off_out, int32_t *nlev) {
    return guard([&] {
        FAIR_CHECK(IA && nlev && (nsteps == 0 || (seq && rows_out && off_out)) && m >= 0 && n >= 0 &&
                       nsteps >= 0 && nsteps <= INT32_MAX, FAIR_ERR_ARG, "null argument or bad sizes");
        FAIR_CHECK(IA[0] == 0 && (IA[m] == 0 || JA), FAIR_ERR_ARG, "IA(0) must be 0");
        for (int32_t r = 0; r < m; ++r) FAIR_CHECK(IA[r + 1] >= IA[r], FAIR_ERR_ARG, "IA must be non-decreasing");
        for (int64_t s = 0; s < nsteps; ++s)
            FAIR_CHECK(seq[s] >= 1 && seq[s] <= m, FAIR_ERR_ARG, "row index out of range in the sequence");
        *nlev = nsteps > 0 ? levelize(n, IA, JA, 1, seq, nsteps, rows_out, off_out) : 0;
        if (nsteps == 0 && off_out) off_out[0] = 0;
    });
}

// basic_sweep, AIRMs.f90:813-858, with host arrays: the row steps over rows[0..nsteps) (one-based
// rows, chosen by the caller), x updated in place.  Returns after x is copied back.
extern "C" int fair_art_steps(fair_csr *h, const int32_t *rows, int64_t nsteps, const double *b, const double *nrA,
                              const fair_opts *o, double *x) {
    return guard([&] {
        check_handle(h);
        check_unique_columns(h);
        FAIR_CHECK(x && b && nrA && (rows || nsteps == 0) && nsteps >= 0, FAIR_ERR_ARG, "null argument");
        const int32_t m = h->A.m, n = h->A.n;
        for (int64_t s = 0; s < nsteps; ++s)
            FAIR_CHECK(rows[s] >= 1 && rows[s] <= m, FAIR_ERR_ARG, "row index out of range in the sweep");
        if (nsteps == 0 || n == 0) return;
        DBuf<int32_t> dr((size_t)nsteps);
        DBuf<double> db(m > 0 ? m : 1), dw(m > 0 ? m : 1), dx(n);
        dr.upload(rows, (size_t)nsteps);
        db.upload(b, m);
        dw.upload(nrA, m);
        dx.upload(x, n);
        const Box box = make_box(o);
        const double relax = o ? o->relaxpar : 1.0;
        // a caller repeating one sequence (cyclic or symmetric sweeps) reuses its level schedule
        const std::shared_ptr<LevelSched> lev = art_levels(h->A, rows, nsteps);
        if (lev) {
            launch_art_levels(h->A, lev.get(), dx.p, nullptr, nullptr, db.p, dw.p, relax, box);
        } else {
            launch_art_chains(h->A, 1, dr.p, nsteps, dx.p, nullptr, 0, nullptr, db.p, dw.p, relax, box);
        }
        dx.download(x, n);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_art_steps_dev(fair_csr *h, const int32_t *d_rows, int64_t nsteps, const double *d_b,
                                  const double *d_nrA, const fair_opts *o, double *d_x) {
    return guard([&] {
        check_handle(h);
        check_unique_columns(h);
        FAIR_CHECK(nsteps >= 0 && (nsteps == 0 || (d_rows && d_b && d_nrA && d_x)), FAIR_ERR_ARG,
                   "null argument");
        const Box box = make_box(o);
        const double relax = o ? o->relaxpar : 1.0;
        if (nsteps > 0 && h->A.n > 0)
            launch_art_chains(h->A, 1, d_rows, nsteps, d_x, nullptr, 0, nullptr, d_b, d_nrA, relax, box);
    });
}

namespace {

// Row sequence of one basic_sweep call (AIRMs.f90:813-858): ord forward or reversed, or, with
// rowprobability, one draw per step searched in the cumulative table (first ri with cumul(ri) >= u).
int64_t sweep_rows(const ArtPlan &P, const std::vector<int32_t> &ord, bool rev, const double *u, int64_t nu,
                   int64_t q, int32_t m, int32_t *out) {
    for (int ii = 0; ii < P.nord; ++ii) {
        if (P.rowprob) {
            FAIR_CHECK(q < nu, FAIR_ERR_ARG, "too few uniforms for twin_art");
            auto it = std::lower_bound(P.cumul.begin(), P.cumul.end(), u[q++]);
            out[ii] = std::min((int)(it - P.cumul.begin()) + 1, (int)m);
        } else {
            out[ii] = rev ? ord[P.nord - 1 - ii] : ord[ii];
        }
    }
    return q;
}

}  // namespace

// twin_art, AIRMs.f90:371-811: chains x (row order, plus its reverse for symkaczmarz) and y
// (reversed order, plus the forward one) run as two workgroups of one sweep launch per
// iteration; stoprule 0 = 'errorgauge' (stop after `slack` iterations without a new minimum of
// ||x - y||, result (x + y)/2 at the minimum), 1 = 'mutualstep' (2x2 least-squares step along
// both chains' last steps, stop when |alpha| + |beta| < 1e-5).  Draws as for fair_art, consumed
// in the reference's order; *consumed receives their number.  Xstore / Ystore: nk x n (row k =
// iterate at Kuse(k)), EG: maxval(Kin) entries (zero where not reached).
extern "C" int fair_twin_art(fair_csr *h, int32_t method, int32_t stoprule, const double *b, int32_t nk,
                             const int32_t *Kin, const double *x0, const double *y0, const int32_t *order,
                             int32_t nord, int32_t slack, const fair_opts *opts, const double *uniforms, int64_t nu,
                             double *xout, int32_t *final_it, double *Xstore, double *Ystore, double *EG,
                             int64_t *consumed) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(b && xout && final_it && Xstore && Ystore && EG, FAIR_ERR_ARG, "null argument");
        FAIR_CHECK(stoprule == 0 || stoprule == 1, FAIR_ERR_METHOD,
                   "Unknown stopping rule. Use errorgauge or mutualstep.");
        DBuf<double> d_nrA;
        ArtPlan P = plan_art(h, method, nk, Kin, order, nord, opts, d_nrA);
        FAIR_CHECK(!P.isrand || uniforms || nu == 0, FAIR_ERR_ARG, "null uniforms");
        const int32_t m = h->A.m, n = h->A.n;
        const Box box = make_box(opts);
        const double relax = opts ? opts->relaxpar : 1.0;
        const bool sym = P.method == FAIR_SYMKACZMARZ;
        int maxK = 0;
        for (int k = 0; k < nk; ++k) maxK = std::max(maxK, Kin[k]);
        for (int k = 0; k < maxK; ++k) EG[k] = 0.0;
        DBuf<double> db(m > 0 ? m : 1), x(n > 0 ? n : 1), y(n > 0 ? n : 1), v(n > 0 ? n : 1), w(n > 0 ? n : 1),
            part(RED_BLOCKS), out(1);
        db.upload(b, m);
        fill(v.p, n, 0.0);
        fill(w.p, n, 0.0);
        const int per_chain = sym ? 2 * P.nord : P.nord;
        std::vector<int32_t> sx(std::max(1, per_chain)), sy(std::max(1, per_chain));
        DBuf<int32_t> dsx(sx.size()), dsy(sy.size());
        std::vector<int32_t> ord = P.ord;
        int64_t q = 0;
        auto scramble_now = [&] {
            FAIR_CHECK(q + 2 * (int64_t)P.nord <= nu, FAIR_ERR_ARG, "too few uniforms for twin_art");
            scramble(P.nord, uniforms + q, ord.data());
            q += 2 * (int64_t)P.nord;
        };
        const unsigned nb = grid_for(n, 256);
        std::shared_ptr<LevelSched> lx, ly;   // held for the whole call
        bool planned = false;
        // both chains' sweeps of one iteration, draws taken in the reference's call order
        auto sweep_both = [&] {
            q = sweep_rows(P, ord, false, uniforms, nu, q, m, sx.data());
            if (sym) q = sweep_rows(P, ord, true, uniforms, nu, q, m, sx.data() + P.nord);
            q = sweep_rows(P, ord, true, uniforms, nu, q, m, sy.data());
            if (sym) q = sweep_rows(P, ord, false, uniforms, nu, q, m, sy.data() + P.nord);
            if (per_chain <= 0) return;
            if (!planned && !P.isrand) {   // without draws both chains repeat their sequences
                lx = art_levels(h->A, sx.data(), per_chain);
                ly = lx ? art_levels(h->A, sy.data(), per_chain) : nullptr;
            }
            planned = true;
            if (lx && ly) {
                launch_art_levels(h->A, lx.get(), x.p, ly.get(), y.p, db.p, d_nrA.p, relax, box);
                return;
            }
            dsx.upload(sx.data(), per_chain);
            dsy.upload(sy.data(), per_chain);
            launch_art_chains(h->A, 2, dsx.p, per_chain, x.p, dsy.p, per_chain, y.p, db.p, d_nrA.p, relax, box);
        };
        auto norm_diff = [&](const double *a, const double *c) { return std::sqrt(dev_sum(n, SqDiff{a, c}, part, out)); };
        if (P.isrand) scramble_now();                                   // :620-622
        int startiter;
        if (x0 && y0) {                                                 // :650-654
            x.upload(x0, n);
            y.upload(y0, n);
            FAIR_CHECK(norm_diff(x.p, y.p) >= 1e-10, FAIR_ERR_ARG, "Initial conditions x0 and y0 too close.");
            startiter = 1;
        } else {                                                        // :634-649, 655-666
            const double *s0 = x0 ? x0 : y0;
            if (s0) {
                x.upload(s0, n);
                y.upload(s0, n);
            } else {
                fill(x.p, n, 0.0);
                fill(y.p, n, 0.0);
            }
            sweep_both();
            EG[0] = norm_diff(x.p, y.p);
            startiter = 2;
        }
        int maxiter = 0;
        for (int k : P.Kuse) maxiter = std::max(maxiter, k);
        size_t rec = 0;
        int trck = 0, iter;
        double minEG = HUGE_VAL;
        *final_it = 0;
        for (iter = startiter; iter <= maxiter; ++iter) {               // :672-770
            if (P.isrand) scramble_now();
            if (stoprule == 1) {
                hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, stream(), x.p, v.p, (int64_t)n);
                hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, stream(), y.p, w.p, (int64_t)n);
            }
            sweep_both();
            if (stoprule == 0) {
                const double eg = norm_diff(x.p, y.p);
                EG[iter - 1] = eg;
                if (eg < minEG) {
                    minEG = eg;
                    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, stream(), x.p, v.p, (int64_t)n);
                    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, stream(), y.p, w.p, (int64_t)n);
                    trck = 0;
                    *final_it = iter;
                } else if (++trck >= slack) {
                    break;
                }
            } else {
                hipLaunchKernelGGL(k_step_back, dim3(nb), dim3(256), 0, stream(), x.p, v.p, (int64_t)n);
                hipLaunchKernelGGL(k_step_back, dim3(nb), dim3(256), 0, stream(), y.p, w.p, (int64_t)n);
                const double nw = std::sqrt(dev_sum(n, SqOf{w.p}, part, out));
                const double nv = std::sqrt(dev_sum(n, SqOf{v.p}, part, out));
                hipLaunchKernelGGL(k_div2, dim3(nb), dim3(256), 0, stream(), v.p, nv, w.p, nw, (int64_t)n);
                const double m1 = dev_sum(n, DotOf{v.p, v.p}, part, out);
                const double m2 = dev_sum(n, DotOf{w.p, w.p}, part, out);
                const double m3 = -dev_sum(n, DotOf{v.p, w.p}, part, out);
                const double r1 = -dev_sum(n, DotDiff{v.p, x.p, y.p}, part, out);
                const double r2 = dev_sum(n, DotDiff{w.p, x.p, y.p}, part, out);
                const double det = m1 * m2 - m3 * m3;
                const double alph = (r1 * m2 - r2 * m3) / det;
                const double bet = (m1 * r2 - m3 * r1) / det;
                hipLaunchKernelGGL(k_axpy2, dim3(nb), dim3(256), 0, stream(), x.p, v.p, alph, y.p, w.p, bet,
                                   (int64_t)n);
                EG[iter - 1] = norm_diff(x.p, y.p);
                if (std::fabs(alph) + std::fabs(bet) < 1e-5) break;
            }
            if (rec < P.Kuse.size() && iter == P.Kuse[rec]) {           // :761-764
                x.download(Xstore + rec * (size_t)n, n);
                y.download(Ystore + rec * (size_t)n, n);
                ++rec;
            }
        }
        DBuf<double> dout(n > 0 ? n : 1);
        if (stoprule == 0) {
            hipLaunchKernelGGL(k_half_sum, dim3(nb), dim3(256), 0, stream(), v.p, w.p, dout.p, (int64_t)n);
        } else {
            hipLaunchKernelGGL(k_half_sum, dim3(nb), dim3(256), 0, stream(), x.p, y.p, dout.p, (int64_t)n);
            *final_it = iter;   // the DO variable after the loop (maxiter + 1 without an exit)
        }
        launch_check("twin_art");
        dout.download(xout, n);
        FAIR_HIP(hipStreamSynchronize(stream()));
        if (consumed) *consumed = q;
    });
}

namespace {

// Two pinned host buffers and two device buffers for row sequences queued on the library's stream:
// queue() waits until the upload last queued from the same host buffer has completed, copies the
// sequence in, queues its upload and returns the device copy.  A device buffer is reused two
// sequences later, after the sweep that read it (stream order).
struct PinnedSeqs {
    int32_t *host[2] = {nullptr, nullptr};
    DBuf<int32_t> dev[2];
    hipEvent_t done[2] = {nullptr, nullptr};
    int next = 0;
    void init(size_t count) {
        for (int i = 0; i < 2; ++i) {
            FAIR_HIP(hipHostMalloc((void **)&host[i], count * sizeof(int32_t), hipHostMallocDefault));
            dev[i].alloc(count);
            FAIR_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
        }
    }
    const int32_t *queue(const int32_t *seq, size_t count) {
        const int i = next;
        next ^= 1;
        FAIR_HIP(hipEventSynchronize(done[i]));
        std::copy(seq, seq + count, host[i]);
        FAIR_HIP(hipMemcpyAsync(dev[i].p, host[i], count * sizeof(int32_t), hipMemcpyHostToDevice, stream()));
        FAIR_HIP(hipEventRecord(done[i], stream()));
        return dev[i].p;
    }
    ~PinnedSeqs() {
        (void)hipStreamSynchronize(stream());
        for (int i = 0; i < 2; ++i) {
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (host[i]) (void)hipHostFree(host[i]);
        }
    }
};

// art, AIRMs.f90:59-369.  With xref (the reference's oracle), the iterate before each iteration is
// kept and the run stops once the distance to xref grows: Xout(:, size(Kin)) = that iterate
// (:322, 330-333).  norm2(xp - xo) of one iteration is norm2(x - xo) of the one before, so one
// device reduction per iteration suffices.  Rows of Xout that are not reached stay zero.
void art_run(fair_csr *h, int32_t method, const double *b, int32_t nk, const int32_t *Kin, const double *x0,
             const double *xref, const int32_t *order, int32_t nord, const fair_opts *opts, const double *uniforms,
             int64_t nu, double *Xout, int32_t *exit_iter, int64_t *consumed) {
    check_handle(h);
    FAIR_CHECK(b && Xout, FAIR_ERR_ARG, "null argument");
    DBuf<double> d_nrA;
    ArtPlan P = plan_art(h, method, nk, Kin, order, nord, opts, d_nrA);
    // Without xref every draw of the run is needed up front; with xref the run may stop early,
    // so the draws are checked as each iteration takes them.
    const int64_t need = uniforms_needed(P);
    const int64_t per_iter = P.isrand ? 2 * (int64_t)P.nord + (P.rowprob ? P.nord : 0) : 0;
    FAIR_CHECK((xref ? std::min(need, 2 * (int64_t)P.nord + per_iter) : need) <= nu && (need == 0 || uniforms),
               FAIR_ERR_ARG, "too few uniforms: need " + std::to_string(need));
    const int32_t m = h->A.m, n = h->A.n;
    const Box box = make_box(opts);
    const double relax = opts ? opts->relaxpar : 1.0;
    for (size_t k = 0; k < (size_t)nk * (size_t)n; ++k) Xout[k] = 0.0;
    DBuf<double> db(m > 0 ? m : 1), dx(n > 0 ? n : 1);
    db.upload(b, m);
    if (x0) dx.upload(x0, n);
    else fill(dx.p, n, 0.0);
    DBuf<double> dxr, dxp, part, out;
    double prev = 0.0;
    if (xref) {
        dxr.alloc(n > 0 ? n : 1);
        dxp.alloc(n > 0 ? n : 1);
        part.alloc(RED_BLOCKS);
        out.alloc(1);
        dxr.upload(xref, n);
        prev = std::sqrt(dev_sum(n, SqDiff{dx.p, dxr.p}, part, out));
    }
    const int steps_per_iter = P.method == FAIR_SYMKACZMARZ ? 2 * P.nord : P.nord;
    std::vector<int32_t> seq(std::max(1, steps_per_iter));
    DBuf<int32_t> dseq(seq.size());
    // Random orders in the window sweep run as a pipeline: the host draws the next iteration's
    // sequence into one of two pinned buffers and queues its upload while the device still sweeps
    // the one before; the sweeps' stop flag is checked once, after the last.
    const bool piped = P.isrand && steps_per_iter > 0 && art_window_usable(h->A, steps_per_iter);
    PinnedSeqs pipe;
    if (piped) {
        pipe.init((size_t)steps_per_iter);
        art_window_clear(h->A);
    }
    std::vector<int32_t> ord = P.ord;
    int64_t q = 0;
    if (P.isrand) {                                   // :297
        scramble(P.nord, uniforms + q, ord.data());
        q += 2 * (int64_t)P.nord;
    }
    // the row sequence of one iteration, taking its draws
    auto build_seq = [&](int32_t *out) {
        if (P.isrand) {                               // :320
            scramble(P.nord, uniforms + q, ord.data());
            q += 2 * (int64_t)P.nord;
        }
        for (int ii = 0; ii < P.nord; ++ii) {
            if (P.rowprob) {                          // :828-835 (first ri with cumul(ri) >= u)
                double u = uniforms[q++];
                auto it = std::lower_bound(P.cumul.begin(), P.cumul.end(), u);
                int ri = (int)(it - P.cumul.begin()) + 1;
                out[ii] = std::min(ri, (int)m);
            } else {
                out[ii] = ord[ii];
            }
        }
        if (P.method == FAIR_SYMKACZMARZ)             // :326-328
            for (int ii = 0; ii < P.nord; ++ii) out[P.nord + ii] = ord[P.nord - 1 - ii];
    };
    size_t rec = 0;
    if (exit_iter) *exit_iter = 0;
    const unsigned nb = grid_for(n, 256);
    // after an iteration: the oracle stop (:330-333) and the recorded iterates (:337-340); true to stop
    auto after_sweep = [&](int iter) {
        if (xref) {
            const double cur = std::sqrt(dev_sum(n, SqDiff{dx.p, dxr.p}, part, out));
            if (cur > prev) {
                dxp.download(Xout + (size_t)(nk - 1) * n, n);
                FAIR_HIP(hipStreamSynchronize(stream()));
                if (exit_iter) *exit_iter = iter;
                return true;
            }
            prev = cur;
        }
        if (rec < P.Kuse.size() && iter == P.Kuse[rec]) {
            dx.download(Xout + rec * (size_t)n, n);
            ++rec;
        } else if (piped) {
            return false;   // nothing to read back: the next iteration is queued behind this one
        }
        FAIR_HIP(hipStreamSynchronize(stream()));
        return false;
    };
    std::shared_ptr<LevelSched> lev;   // held for the whole call
    for (int iter = 1; iter <= P.maxiter; ++iter) {
        FAIR_CHECK(q + per_iter <= nu || !P.isrand, FAIR_ERR_ARG, "too few uniforms: need " + std::to_string(need));
        build_seq(seq.data());
        if (xref && n > 0) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, stream(), dx.p, dxp.p, (int64_t)n);
        // without draws every iteration sweeps the same sequence: one level schedule serves all
        if (iter == 1 && !P.isrand) lev = art_levels(h->A, seq.data(), steps_per_iter);
        if (lev) {
            launch_art_levels(h->A, lev.get(), dx.p, nullptr, nullptr, db.p, d_nrA.p, relax, box);
        } else if (piped) {
            const int32_t *d_rows = pipe.queue(seq.data(), (size_t)steps_per_iter);
            launch_art_chains(h->A, 1, d_rows, steps_per_iter, dx.p, nullptr, 0, nullptr, db.p, d_nrA.p, relax, box,
                              true);
        } else if (steps_per_iter > 0) {
            dseq.upload(seq.data(), steps_per_iter);
            launch_art_chains(h->A, 1, dseq.p, steps_per_iter, dx.p, nullptr, 0, nullptr, db.p, d_nrA.p, relax,
                              box);
        }
        if (after_sweep(iter)) break;
    }
    if (piped) art_window_check(h->A);
    if (consumed) *consumed = q;
}

}  // namespace

extern "C" int fair_art(fair_csr *h, int32_t method, const double *b, int32_t nk, const int32_t *Kin,
                        const double *x0, const int32_t *order, int32_t nord, const fair_opts *opts,
                        const double *uniforms, int64_t nu, double *Xout) {
    return guard([&] {
        art_run(h, method, b, nk, Kin, x0, nullptr, order, nord, opts, uniforms, nu, Xout, nullptr, nullptr);
    });
}

extern "C" int fair_art_oracle(fair_csr *h, int32_t method, const double *b, int32_t nk, const int32_t *Kin,
                               const double *x0, const double *xref, const int32_t *order, int32_t nord,
                               const fair_opts *opts, const double *uniforms, int64_t nu, double *Xout,
                               int32_t *exit_iter, int64_t *consumed) {
    return guard([&] {
        FAIR_CHECK(xref != nullptr, FAIR_ERR_ARG, "null xref");
        art_run(h, method, b, nk, Kin, x0, xref, order, nord, opts, uniforms, nu, Xout, exit_iter, consumed);
    });
}
