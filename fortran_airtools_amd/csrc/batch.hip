// batch.hip -- precision variants and batched SIRT.
//
// FP32 single right-hand side: the same blocked SpMV kernels instantiated for float (FP32 matrix
// values, FP32 iterate, FP32 accumulation); the SIRT weights M, D stay FP64 and the per-row /
// per-column scaling is applied in FP64 before rounding the result to FP32.
//
// Batched SIRT (k right-hand sides, e.g. k slices sharing one geometry): X is n x k and B, R are
// m x k, row-major (the k values of one pixel / ray are contiguous).  The SpMM kernels spread the
// right-hand sides over the lanes, so every gathered matrix entry pulls a contiguous row segment
// of X; the row order, the number of rows per wavefront (and with it the right-hand sides per
// pass) and the tiled variant for A^T are tuned per matrix.  The iteration loop can be captured
// once into a HIP graph and replayed.
#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <tuple>
#include <utility>

#include "fair_internal.hpp"
#include "spmv.hpp"

namespace fair {
namespace {

__global__ void k_to_f32(const double *a, float *b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

void make_f32(Csr &A) {
    if (A.vA32.p) return;
    A.vA32.alloc((size_t)A.nnz + CSR_PAD);
    FAIR_HIP(hipMemsetAsync(A.vA32.p + A.nnz, 0, sizeof(float) * CSR_PAD, stream()));
    if (A.nnz > 0)
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A.nnz, 256)), dim3(256), 0, stream(), A.vA.p, A.vA32.p, A.nnz);
    launch_check("fp32 values");
}

struct BoxF {
    int upb, lwrb;
    float ub, lb;
    __device__ __forceinline__ float apply(float v) const {
        if (upb) v = fmin_ref(v, ub);
        if (lwrb) v = fmax_ref(v, lb);
        return v;
    }
};

BoxF make_boxf(const fair_opts *o) {
    BoxF b{0, 0, 0.f, 1.f};
    if (o) {
        b.upb = o->upb != 0;
        b.lwrb = o->lwrb != 0;
        b.ub = (float)o->ubound;
        b.lb = (float)o->lbound;
    }
    return b;
}

// r = (b - Ax) / M with M in FP64
struct EpiResidualF {
    const float *b;
    const double *M;
    float *r;
    __device__ __forceinline__ void operator()(int32_t i, float acc) const {
        r[i] = (float)(((double)b[i] - (double)acc) / M[i]);
    }
};

// x_new = box(x + relax * (A^T r) / D), scaling in FP64
struct EpiSirtUpdateF {
    const float *x;
    const double *D;
    float *xn;
    double relax;
    BoxF box;
    __device__ __forceinline__ void operator()(int32_t j, float acc) const {
        double u = relax * (double)acc;
        if (D) u = u / D[j];
        xn[j] = box.apply((float)((double)x[j] + u));
    }
};

template <class Src, class Epi>
void spmv_f32(const Csr &A, Src src, Epi epi) {
    if (A.blk.ok)
        spmv_launch_blocked<float, Src, Epi>(A.blk, A.JA.p, A.vA32.p, A.n, src, epi, stream());
    else
        spmv_launch_raw<float, Src, Epi>((const int4 *)A.rows_nz.p, A.n_rows_nz, A.JA.p, A.vA32.p, A.n,
                                         src, epi, stream());
    launch_check("spmv fp32");
}

__global__ void k_copy_f32(const float *a, float *b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

void check_handle(const fair_csr *A) { FAIR_CHECK(A != nullptr, FAIR_ERR_ARG, "null matrix handle"); }

// ---- batched: SpMM ----------------------------------------------------------------------------
// One wavefront per matrix row.  Each lane holds VEC consecutive right-hand sides (VEC = 4 when k
// is a multiple of 4), so one pass covers 64*VEC of them (all 256 of a k = 256 batch) and every
// matrix entry gathers one contiguous, coalesced row segment of X (1 KB at k = 256): the matrix
// is read once per product.  A row's entries are loaded 64 at a time, one per lane, and
// broadcast with v_readlane; SPMM_U gathers are issued before their products are summed.
// Columns outside [0, ncols) and the lanes past k read a zero row instead (their weight is 0).
//
// Rows run in a tuned order (spmm_order below): rays grouped by sinogram tiles (angles x
// detectors) for A, pixels by image tiles for A^T, so that the rows in flight on one XCD share
// their footprint in its L2.  The hardware deals block b to XCD b % 8; logical block
// (b % 8) * per_xcd + b / 8 gives every XCD a contiguous stretch of the order.
constexpr int NXCD = 8;
constexpr int SPMM_U = 8;
constexpr int SPMM_WAVES = 4;   // rows (wavefronts) per 256-thread block

template <int V>
struct FV;
template <>
struct FV<1> {
    float v[1];
    __device__ __forceinline__ static FV load(const float *p) { return FV{{*p}}; }
};
template <>
struct FV<4> {
    float v[4];
    __device__ __forceinline__ static FV load(const float *p) {
        const float4 t = *reinterpret_cast<const float4 *>(p);
        return FV{{t.x, t.y, t.z, t.w}};
    }
};
template <int V>
__device__ __forceinline__ void store_fv(float *p, const float (&o)[V]) {
    if constexpr (V == 4) *reinterpret_cast<float4 *>(p) = make_float4(o[0], o[1], o[2], o[3]);
    else p[0] = o[0];
}

template <int VEC, class Epi>
__global__ void __launch_bounds__(256) k_spmm_rows(const int4 *__restrict__ rows, int32_t nrows, int32_t per_xcd,
                                                   const int32_t *__restrict__ JA, const float *__restrict__ vA,
                                                   int32_t ncols, const float *__restrict__ X, int32_t k,
                                                   const float *__restrict__ zrow, Epi epi) {
    const int b = blockIdx.x;
    const int lb = (b % NXCD) * per_xcd + b / NXCD;
    const int wv = __builtin_amdgcn_readfirstlane(lb * SPMM_WAVES + (int)(threadIdx.x >> 6));
    if (wv >= nrows) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int4 ri = rows[wv];
    for (int c0 = 0; c0 < k; c0 += WAVE * VEC) {
        const int col = c0 + lane * VEC;
        const bool lane_in = col < k;
        // lanes past k read the start of the gathered row (their sums are dropped); masked
        // entries read the zero row at the same lane offset
        const uint32_t offb = 4u * (uint32_t)(lane_in ? col : c0);   // byte offset in a gathered row
        const float *zbase = zrow - c0;
        float acc[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[e] = 0.f;
        // the next 64 entries are loaded (unconditionally, clamped to the row's last entry) while
        // the current ones are gathered
        const int last = max(ri.z - 1, ri.y);
        int cn = JA[min(ri.y + lane, last)];
        float vn = vA[min(ri.y + lane, last)];
        for (int base = ri.y; base < ri.z; base += WAVE) {
            const bool in = base + lane < ri.z;
            int cj = in ? cn : -1;
            float vj = vn;
            cn = JA[min(base + WAVE + lane, last)];
            vn = vA[min(base + WAVE + lane, last)];
            if ((uint32_t)cj >= (uint32_t)ncols) {
                cj = -1;
                vj = 0.f;
            }
            const int cnt = min(WAVE, ri.z - base);
            for (int t = 0; t < cnt; t += SPMM_U) {
                FV<VEC> xv[SPMM_U];
                float w[SPMM_U];
#pragma unroll
                for (int u = 0; u < SPMM_U; ++u) {
                    const int c = __builtin_amdgcn_readlane(cj, t + u);
                    w[u] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vj), t + u));
                    const float *rowp = X + (int64_t)max(c, 0) * k;
                    rowp = c >= 0 ? rowp : zbase;
                    xv[u] = FV<VEC>::load(reinterpret_cast<const float *>(reinterpret_cast<const char *>(rowp) + offb));
                }
#pragma unroll
                for (int u = 0; u < SPMM_U; ++u)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) acc[e] += w[u] * xv[u].v[e];
            }
        }
        if (lane_in) epi.template apply<VEC>(ri.x, col, acc);
    }
}

// ---- batched: SpMM with several rows per wavefront ---------------------------------------------
// k_spmm_groups gives each row a group of G = 16 or 32 lanes (64 / G rows per wavefront) and each
// lane four right-hand sides, so one pass covers 4G of them; the grid's y index picks the chunk.
// The hardware dispatches the y slices one after the other, so each slice is a full pass over the
// rows (the matrix is read once per slice) whose rows in flight on an XCD touch 256 / 4G times
// fewer cache lines of the panel than one pass over 256, while one gather instruction still moves
// 64 / G row segments of 16G bytes (1 KB in all, as in k_spmm_rows).  Each 16-lane row of the
// wavefront loads 16 consecutive entries of its group's matrix row and hands entry t to all its
// lanes with the DPP row broadcast.  The panel is read through a buffer resource: an entry past its
// row's end, or with a column outside [0, ncols), gets an offset beyond the resource's range,
// which loads zeros, and weight 0.  Each output sums its row's entries in the stored order with the
// same expression as k_spmm_rows, whose padding adds +0 the same way, so both give the same bits.
constexpr uint32_t SPMM_OOB = 0x80000000u;   // a panel offset beyond every resource's range

template <int N>
__device__ __forceinline__ uint32_t row_bcast(uint32_t v) {   // lane N of each 16-lane row, to the row
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + N, 0xF, 0xF, false);
}

// entries T0 .. T0 + SPMM_U - 1 of the 16 held by each 16-lane row
template <int T0, int... U>
__device__ __forceinline__ void spmm_span(__amdgpu_buffer_rsrc_t px, uint32_t oj, uint32_t wj, uint32_t offb,
                                          float (&acc)[4], std::integer_sequence<int, U...>) {
    const uint32_t o[SPMM_U] = {(row_bcast<T0 + U>(oj) + offb)...};
    const uint32_t w[SPMM_U] = {row_bcast<T0 + U>(wj)...};
    decltype(__builtin_amdgcn_raw_buffer_load_b128(px, 0u, 0, 0)) t[SPMM_U];
#pragma unroll
    for (int u = 0; u < SPMM_U; ++u) t[u] = __builtin_amdgcn_raw_buffer_load_b128(px, o[u], 0, 0);
#pragma unroll
    for (int u = 0; u < SPMM_U; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] += __uint_as_float(w[u]) * __uint_as_float(t[u][e]);
}

// entries T0 .. T0 + 7 of a 16-entry half window held by each 16-lane group: LDS row offset and
// weight broadcast from lane t of the group; the lane reads its float4 of the staged row
template <int T0, int... U>
__device__ __forceinline__ void spmm_slice_span(const float4 *buf, int off, uint32_t wbits, int gl, float (&acc)[4],
                                                std::integer_sequence<int, U...>) {
    const int o[sizeof...(U)] = {(int)row_bcast<T0 + U>((uint32_t)off)...};
    const uint32_t wv[sizeof...(U)] = {row_bcast<T0 + U>(wbits)...};
#pragma unroll
    for (int u = 0; u < (int)sizeof...(U); ++u) {
        const float4 xv = buf[o[u] + gl];
        const float ww = __uint_as_float(wv[u]);
        acc[0] += ww * xv.x;
        acc[1] += ww * xv.y;
        acc[2] += ww * xv.z;
        acc[3] += ww * xv.w;
    }
}

template <int G, class Epi>
__global__ void __launch_bounds__(256) k_spmm_groups(const int4 *__restrict__ rows, int32_t nrows, int32_t per_xcd,
                                                     const int32_t *__restrict__ JA, const float *__restrict__ vA,
                                                     int32_t ncols, const float *__restrict__ X, int32_t k, Epi epi) {
    static_assert(G == 16 || G == 32, "16 or 32 lanes per row");
    constexpr int RPW = WAVE / G;
    const int b = blockIdx.x;
    const int lb = (b % NXCD) * per_xcd + b / NXCD;
    const int w0 = __builtin_amdgcn_readfirstlane((lb * SPMM_WAVES + (int)(threadIdx.x >> 6)) * RPW);
    if (w0 >= nrows) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int g = lane / G, rl = lane & 15;
    const int row = w0 + g;
    const int4 ri = row < nrows ? rows[row] : make_int4(0, 0, 0, 0);
    const int len = ri.z - ri.y;
    int maxlen = 0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) maxlen = max(maxlen, __builtin_amdgcn_readlane(len, i * G));
    const int c0 = (int)blockIdx.y * 4 * G;
    const int col = c0 + 4 * (lane % G);
    const uint32_t offb = 4u * (uint32_t)(col < k ? col : c0);   // lanes past k re-read the chunk's start
    const uint32_t k4 = 4u * (uint32_t)k;
    const __amdgpu_buffer_rsrc_t px =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(X), (short)0, (int)((uint32_t)ncols * k4), 0x00020000);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    // The next 16 entries are loaded while the current ones are gathered.  The loads are
    // unconditional (an entry past the row's end re-reads the row's last one, or entry 0 of the
    // matrix for an empty group) so that waiting for the current ones never waits for them.
    const int last = ri.y + max(len - 1, 0);
    int cn = JA[min(ri.y + rl, last)];
    float vn = vA[min(ri.y + rl, last)];
    for (int base = 0; base < maxlen; base += 16) {
        const bool in = base + rl < len;
        const int cj = in ? cn : -1;
        const float vj = vn;
        cn = JA[min(ri.y + base + 16 + rl, last)];
        vn = vA[min(ri.y + base + 16 + rl, last)];
        const bool ok = (uint32_t)cj < (uint32_t)ncols;
        const uint32_t oj = ok ? (uint32_t)cj * k4 : SPMM_OOB;
        const uint32_t wj = ok ? __float_as_uint(vj) : 0u;
        spmm_span<0>(px, oj, wj, offb, acc, std::make_integer_sequence<int, SPMM_U>{});
        if (maxlen - base > SPMM_U) spmm_span<SPMM_U>(px, oj, wj, offb, acc, std::make_integer_sequence<int, SPMM_U>{});
    }
    if (row < nrows && col < k) epi.template apply<4>(ri.x, col, acc);
}

template <class Epi>
__global__ void k_spmm_empty(const int32_t *__restrict__ rows, int32_t nrows, int32_t k, Epi epi) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nrows * k) return;
    const float z[1] = {0.f};
    epi.template apply<1>(rows[t / k], (int)(t % k), z);
}

struct EpiStoreB {
    float *Y;
    int32_t k;
    template <int V>
    __device__ __forceinline__ void apply(int32_t i, int j, const float (&acc)[V]) const {
        store_fv<V>(Y + (int64_t)i * k + j, acc);
    }
};

// r = (b - Ax) / M, FP64 scaling
struct EpiResidualB {
    const float *B;
    const double *M;
    float *R;
    int32_t k;
    template <int V>
    __device__ __forceinline__ void apply(int32_t i, int j, const float (&acc)[V]) const {
        const int64_t q = (int64_t)i * k + j;
        const double mi = M[i];
        const FV<V> bv = FV<V>::load(B + q);
        float o[V];
#pragma unroll
        for (int e = 0; e < V; ++e) o[e] = (float)(((double)bv.v[e] - (double)acc[e]) / mi);
        store_fv<V>(R + q, o);
    }
};

// x_new = box(x + relax * (A^T r) / D), FP64 scaling
struct EpiUpdateB {
    const float *X;
    const double *D;
    float *Xn;
    int32_t k;
    double relax;
    BoxF box;
    template <int V>
    __device__ __forceinline__ void apply(int32_t p, int j, const float (&acc)[V]) const {
        const int64_t q = (int64_t)p * k + j;
        const double dp = D ? D[p] : 1.0;
        const FV<V> xv = FV<V>::load(X + q);
        float o[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            double u = relax * (double)acc[e];
            if (D) u = u / dp;
            o[e] = box.apply((float)((double)xv.v[e] + u));
        }
        store_fv<V>(Xn + q, o);
    }
};

__device__ float g_zero_row[WAVE * 4];   // zero-initialised: the gather target of masked entries

const float *zero_row() {
    static const float *p = nullptr;
    if (!p) {
        void *q = nullptr;
        FAIR_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(g_zero_row)));
        p = static_cast<const float *>(q);
    }
    return p;
}

// Row orders of the SpMM.  Code 0 keeps the compacted row list's order; codes 1.. group rays
// (rows of a ray-traced A: angle a = r / p, detector d = r % p) into BA x BD sinogram tiles or
// pixels (rows of A^T: image column x = j / N, row y = j % N) into TX x TY image tiles.  Tiles
// run angle-block (column-block) major, or with the third entry 1 detector-band (row-band) major
// (a band of detectors through all angles); rows inside a tile run row-major.
constexpr int SPMM_ORDERS = 12;
constexpr int ALL = 1 << 20;   // one block spanning every angle
constexpr int RAY_TILES[SPMM_ORDERS][3] = {{0, 0, 0},   {16, 16, 0}, {32, 8, 0},   {32, 16, 0},
                                           {64, 8, 0},  {32, 32, 0}, {64, 16, 0},  {16, 32, 0},
                                           {32, 16, 1}, {64, 8, 1},  {ALL, 16, 1}, {ALL, 8, 1}};
constexpr int PIX_TILES[SPMM_ORDERS][3] = {{0, 0, 0},   {8, 8, 0},   {16, 16, 0}, {32, 32, 0},
                                           {16, 32, 0}, {32, 16, 0}, {16, 8, 0},  {8, 16, 0},
                                           {32, 8, 0},  {64, 16, 0}, {16, 16, 1}, {64, 64, 0}};

// Code c + SPMM_GROUP_CODE * s (s = 1, 2) runs k_spmm_groups with SPMM_GROUP_LANES[s] lanes per row
// in row order c; + SPMM_CAP_CODE * b (b = 2..6) lets at most b of the row kernels' workgroups be
// resident on a compute unit (each reserves 1 / b of its LDS), which bounds the rows in flight on
// an XCD and with them the panel lines they share its L2 with.
constexpr int SPMM_GROUP_CODE = 1000;
constexpr int SPMM_GROUPINGS = 3;
constexpr int SPMM_GROUP_LANES[SPMM_GROUPINGS] = {WAVE, 32, 16};
constexpr int SPMM_CAP_CODE = 10000;
constexpr int SPMM_CAPS[] = {6, 5, 4, 3, 2};

int code_order(int code) { return code % SPMM_GROUP_CODE; }
int code_grouping(int code) { return (code / SPMM_GROUP_CODE) % (SPMM_CAP_CODE / SPMM_GROUP_CODE); }

// LDS bytes a row-kernel workgroup reserves so that at most `cap` fit on a compute unit (0: none)
size_t cap_lds(int code) {
    const int cap = code / SPMM_CAP_CODE;
    if (cap <= 0) return 0;
    const size_t b = ((size_t)lds_per_cu() / (size_t)cap) & ~(size_t)1023;
    return std::min(b, (size_t)lds_per_block());
}

// Whether grouping s (> 0) applies to k right-hand sides over ncols panel rows: k a multiple of 4
// and wider than one pass, the panel addressable by the buffer resource with room for SPMM_OOB.
bool groups_apply(int s, int32_t k, int32_t ncols) {
    if (s <= 0 || s >= SPMM_GROUPINGS || k % 4 != 0 || k <= 4 * SPMM_GROUP_LANES[s]) return false;
    return (uint64_t)ncols * (uint64_t)k * 4u + 4u * (uint64_t)k <= SPMM_OOB;
}

// The row descriptors of M in order `code` (built once per code and cached in M); rays says
// whether M's rows are rays (A of a ray-traced matrix) or pixels (its transpose).
const int4 *spmm_order(const fair_csr *h, Csr &M, bool rays, int code) {
    code %= SPMM_GROUP_CODE;
    const bool geo = h->g_n > 0 && h->g_p > 0 && h->g_nth > 0;
    if (code <= 0 || code >= SPMM_ORDERS || !geo || M.n_rows_nz <= 0) return (const int4 *)M.rows_nz.p;
    if (M.spmm_code == code && M.spmm_rows.p) return (const int4 *)M.spmm_rows.p;
    const int64_t nr = M.n_rows_nz;
    std::vector<int4> d((size_t)nr);
    FAIR_HIP(hipMemcpyAsync(d.data(), M.rows_nz.p, sizeof(int4) * nr, hipMemcpyDeviceToHost, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
    const int64_t N = h->g_n, P = h->g_p, NTH = h->g_nth;
    const int *tile = rays ? RAY_TILES[code] : PIX_TILES[code];
    const int64_t t0 = std::min<int64_t>(tile[0], rays ? NTH : N), t1 = tile[1];
    const bool band_major = tile[2] != 0;
    const int64_t n0 = rays ? (NTH + t0 - 1) / t0 : (N + t0 - 1) / t0;   // blocks along u
    const int64_t n1 = rays ? (P + t1 - 1) / t1 : (N + t1 - 1) / t1;     // blocks along v
    std::vector<std::pair<int64_t, int32_t>> key((size_t)nr);
    for (int64_t s = 0; s < nr; ++s) {
        const int64_t r = d[s].x;
        const int64_t u = rays ? r / P : r / N, v = rays ? r % P : r % N;
        const int64_t blk = band_major ? (v / t1) * n0 + u / t0 : (u / t0) * n1 + v / t1;
        key[s] = {blk * (t0 * t1) + (u % t0) * t1 + v % t1, (int32_t)s};
    }
    std::sort(key.begin(), key.end());
    std::vector<int4> o((size_t)nr);
    for (int64_t s = 0; s < nr; ++s) o[s] = d[key[s].second];
    if (h->graphs_used) drop_graphs(h);   // captured launches refer to the order being replaced
    M.spmm_rows.alloc(4 * (size_t)nr);
    FAIR_HIP(hipMemcpyAsync(M.spmm_rows.p, o.data(), sizeof(int4) * nr, hipMemcpyHostToDevice, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
    M.spmm_code = code;
    return (const int4 *)M.spmm_rows.p;
}

// ---- batched: tiled SpMM for rows that share their columns (pixels of A^T) ------------------------
// A block of 16 wavefronts takes the rows of one tile (pixels of a small image tile, which are
// crossed by largely the same rays), one, two or four rows per wavefront (16-, 32- or 64-pixel
// tiles).  The union of the tile's columns, sorted, is staged through LDS in chunks of CH panel
// rows, double-buffered: chunk c + 1 is loaded into registers while chunk c is summed and written to the
// other buffer at the end of the step, and one barrier per chunk orders both.  Each row's entries
// of a chunk are a prefix of a 64-entry window (LDS slot and value, one per lane; the slots
// ascend), found with ballot + popcount.  Gathered rows are read from LDS (1 KB per wavefront read
// at k = 256) instead of through the texture data path, which bounds the row kernels.  Each lane
// sums its row's entries in the stored order with the same expression as k_spmm_rows, so both
// kernels give the same bits.  One pass over k <= 256 right-hand sides (k a multiple of 4).
constexpr int STILE_ROWS = 16;
constexpr int STILE_SHAPES = 9;   // 16, 32 or 64 pixels (x by y): one, two or four per wavefront
constexpr int STILE[STILE_SHAPES][2] = {{4, 4}, {8, 2}, {2, 8}, {16, 1}, {1, 16}, {8, 4}, {4, 8}, {8, 8}, {16, 4}};
constexpr int SRAY_SHAPES = 6;    // 16, 32 or 64 rays (angles by detectors)
constexpr int SRAY[SRAY_SHAPES][2] = {{8, 8}, {16, 4}, {4, 16}, {8, 4}, {4, 8}, {4, 4}};
constexpr int STILE_CHUNKS = 4;
constexpr int STILE_CH[STILE_CHUNKS] = {64, 32, 128, 64};
constexpr int STILE_SLICE = 64;   // right-hand sides per pass of the sliced kernel (chunk kinds 2, 3)
// order codes STILE_CODE0 + 10 * c + i select the tiled kernel, shape i, chunk STILE_CH[c]; c = 0, 1
// stage whole panel rows (k_spmm_tiles), c = 2, 3 slices of 64 right-hand sides (k_spmm_tiles_sliced)
constexpr int STILE_CODE0 = 100;
bool is_tiled(int code) {
    return code >= STILE_CODE0 && code < STILE_CODE0 + 10 * STILE_CHUNKS && code % 10 < STILE_SHAPES;
}
int tiled_shape(int code) { return (code - STILE_CODE0) % 10; }
int tiled_chunk(int code) { return STILE_CH[(code - STILE_CODE0) / 10]; }
bool tiled_sliced(int code) { return (code - STILE_CODE0) / 10 >= 2; }
size_t tiled_lds(int code, int32_t k) {
    return tiled_sliced(code) ? 2 * ((size_t)tiled_chunk(code) + 1) * STILE_SLICE * sizeof(float)
                              : 2 * (size_t)tiled_chunk(code) * (size_t)k * sizeof(float);
}

template <int CH, int RPW, class Epi>
__global__ void __launch_bounds__(1024) k_spmm_tiles(const int4 *__restrict__ binfo, int32_t nblocks, int32_t per_xcd,
                                                     const int4 *__restrict__ rows, const int32_t *__restrict__ uni,
                                                     const uint16_t *__restrict__ slot, const float *__restrict__ vA,
                                                     const float *__restrict__ X, int32_t k, Epi epi) {
    extern __shared__ float4 stage[];   // two buffers of CH rows of k / 4 float4
    constexpr int PER_WAVE = CH / STILE_ROWS;
    const int b = blockIdx.x;
    const int lb = (b % NXCD) * per_xcd + b / NXCD;
    if (lb >= nblocks) return;          // the whole block leaves together
    const int4 bi = binfo[lb];          // first row, rows, union start, union size
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    const int kv = k >> 2, kl = min(lane, kv - 1);
    // this wavefront's rows of the tile: w, w + 16, ...
    bool has[RPW];
    int4 ri[RPW];
    int last[RPW];   // entry loads are clamped here (entry 0 without a row)
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
        has[j] = w + STILE_ROWS * j < bi.y;
        ri[j] = has[j] ? rows[bi.x + w + STILE_ROWS * j] : make_int4(0, 0, 0, 0);
        last[j] = max(ri[j].z - 1, ri[j].y);
    }
    // the union rows of chunk c0 (all union numbers first, then all gathers; rows past the union
    // re-read its last one, lanes past k / 4 the row's last float4, neither is stored)
    FV<4> pre[PER_WAVE];   // plain floats: the rows stay in registers
    int col[PER_WAVE];
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) col[i] = uni[bi.z + min(w + STILE_ROWS * i, bi.w - 1)];
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) pre[i] = FV<4>::load(X + (int64_t)col[i] * k + 4 * kl);
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i)
        if (w + STILE_ROWS * i < min(CH, bi.w) && lane < kv)
            stage[(w + STILE_ROWS * i) * kv + lane] = make_float4(pre[i].v[0], pre[i].v[1], pre[i].v[2], pre[i].v[3]);
    // Each row's entries from q0 on, one per lane; q is the next one to sum.  A chunk holds at most
    // CH <= 64 of a row's entries (the layout is built only where every row's slots strictly
    // ascend), so a window loaded at q holds all entries of the next 64 / CH chunks; the windows are
    // reloaded that often, all rows together, at the end of a step, after the gathers of the next
    // chunk were waited for: the waits for a window never wait for gathers issued after it.  Every
    // step issues the same gathers (the last step's re-read the union's last row, unused) and a
    // missing row loads entry 0, so that the waits between them are exact.
    int q0[RPW], q[RPW], sl[RPW];
    float va[RPW];
    float acc[RPW][4];
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
        q0[j] = q[j] = ri[j].y;
        sl[j] = (int)slot[min(q0[j] + lane, last[j])];
        va[j] = vA[min(q0[j] + lane, last[j])];
        acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0.f;
    }
    const int nch = (bi.w + CH - 1) / CH;
    for (int c = 0; c < nch; ++c) {
        const int c0 = c * CH, cn = min(CH, bi.w - c0);
        const float4 *buf = stage + (c & 1) * CH * kv;
        // chunk c is in its buffer; every wavefront is done with chunk c - 1, whose buffer takes
        // chunk c + 1 at the end of this step
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) col[i] = uni[bi.z + min(c0 + CH + w + STILE_ROWS * i, bi.w - 1)];
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) pre[i] = FV<4>::load(X + (int64_t)col[i] * k + 4 * kl);
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            // entries of the window below the chunk's end (a prefix: the slots ascend; lanes past
            // the row's end hold a clamped copy and do not count), from q on
            const int nin = has[j] ? __popcll(__ballot(sl[j] < c0 + cn && q0[j] + lane < ri[j].z)) : 0;
            int t = q[j] - q0[j];
            for (; t + 4 <= nin; t += 4) {
                float4 xv[4];
                float wv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int su = __builtin_amdgcn_readlane(sl[j], t + u) - c0;
                    wv[u] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(va[j]), t + u));
                    xv[u] = buf[su * kv + kl];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc[j][0] += wv[u] * xv[u].x;
                    acc[j][1] += wv[u] * xv[u].y;
                    acc[j][2] += wv[u] * xv[u].z;
                    acc[j][3] += wv[u] * xv[u].w;
                }
            }
            for (; t < nin; ++t) {
                const int su = __builtin_amdgcn_readlane(sl[j], t) - c0;
                const float wv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(va[j]), t));
                const float4 xv = buf[su * kv + kl];
                acc[j][0] += wv * xv.x;
                acc[j][1] += wv * xv.y;
                acc[j][2] += wv * xv.z;
                acc[j][3] += wv * xv.w;
            }
            q[j] = q0[j] + max(nin, q[j] - q0[j]);   // the rest of the row lies in later chunks
        }
        float4 *nb = stage + ((c + 1) & 1) * CH * kv;
        const int nn = min(CH, bi.w - c0 - CH);
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i)
            if (w + STILE_ROWS * i < nn && lane < kv)
                nb[(w + STILE_ROWS * i) * kv + lane] = make_float4(pre[i].v[0], pre[i].v[1], pre[i].v[2], pre[i].v[3]);
        if (2 * CH > WAVE || (c & 1)) {   // a window holds the next WAVE / CH chunks' entries of a row
#pragma unroll
            for (int j = 0; j < RPW; ++j) {
                q0[j] = q[j];
                sl[j] = (int)slot[min(q0[j] + lane, last[j])];
                va[j] = vA[min(q0[j] + lane, last[j])];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < RPW; ++j)
        if (has[j] && lane < kv) epi.template apply<4>(ri[j].x, lane * 4, acc[j]);
}

// The sliced tiled kernel: the same tiles and layout, the right-hand sides in passes of 64 (the
// grid's y index), so a staged union row is 256 B and a chunk of CH = 64 or 128 of them holds four
// times the entries of a row-staged chunk per barrier.  Each 16-lane group of a wavefront takes one
// row of the tile (64 rows per workgroup), each lane four right-hand sides; a group holds a 32-entry
// window of its row's entries from the next one to sum (slot and value, two per lane), reloaded per
// chunk, and hands entry t to its lanes with the DPP row broadcast.  Window positions outside the
// chunk or the row read the buffer's zero row with weight 0, which adds +0: each output sums its row's entries in the stored order with the same
// expression as the row kernels, so all give the same bits.
template <int CH, class Epi>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_spmm_tiles_sliced(const int4 *__restrict__ binfo, int32_t nblocks,
                                                            int32_t per_xcd, const int4 *__restrict__ rows,
                                                            const int32_t *__restrict__ uni,
                                                            const uint16_t *__restrict__ slot,
                                                            const float *__restrict__ vA, const float *__restrict__ X,
                                                            int32_t k, Epi epi) {
    extern __shared__ float4 stage[];   // two buffers of CH + 1 rows of 16 float4; row CH stays zero
    constexpr int RW = STILE_SLICE / 4;   // float4 per staged row
    constexpr int PT = CH * RW / 1024;    // staged float4 per thread and chunk
    static_assert(PT >= 1 && CH * RW == PT * 1024, "a chunk is a whole number of float4 per thread");
    const int b = blockIdx.x;
    const int lb = (b % NXCD) * per_xcd + b / NXCD;
    if (lb >= nblocks) return;          // the whole block leaves together
    const int4 bi = binfo[lb];          // first row, rows, union start, union size
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & (WAVE - 1)), gl = lane & 15;
    const int ri_i = 4 * w + (lane >> 4);   // this lane group's row of the tile
    const bool has = ri_i < bi.y;
    const int4 ri = has ? rows[bi.x + ri_i] : make_int4(0, 0, 0, 0);
    const int last = max(ri.z - 1, ri.y);
    const int c0s = (int)blockIdx.y * STILE_SLICE;
    const int col = c0s + 4 * gl;           // this lane's right-hand sides
    // staging: thread t copies float4 (t % 16) of rows t / 16 + 64 i; rows past the union re-read its
    // last one (not stored)
    const int sr = (int)threadIdx.x >> 4, sc = (int)threadIdx.x & 15;
    const int scol = min(c0s + 4 * sc, k - 4);
    FV<4> pre[PT];
    int ucol[PT];
#pragma unroll
    for (int i = 0; i < PT; ++i) ucol[i] = uni[bi.z + min(sr + 64 * i, bi.w - 1)];
#pragma unroll
    for (int i = 0; i < PT; ++i) pre[i] = FV<4>::load(X + (int64_t)ucol[i] * k + scol);
#pragma unroll
    for (int i = 0; i < PT; ++i)
        if (sr + 64 * i < min(CH, bi.w))
            stage[(sr + 64 * i) * RW + sc] = make_float4(pre[i].v[0], pre[i].v[1], pre[i].v[2], pre[i].v[3]);
    if (threadIdx.x < 2 * RW) stage[((threadIdx.x / RW) * (CH + 1) + CH) * RW + threadIdx.x % RW] = make_float4(0.f, 0.f, 0.f, 0.f);
    // the row's 32-entry window from w0 (its next entry to sum): entries w0 + gl and w0 + 16 + gl
    int w0 = ri.y;
    int sl0 = (int)slot[min(w0 + gl, last)], sl1 = (int)slot[min(w0 + 16 + gl, last)];
    float va0 = vA[min(w0 + gl, last)], va1 = vA[min(w0 + 16 + gl, last)];
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    const int nch = (bi.w + CH - 1) / CH;
    for (int c = 0; c < nch; ++c) {
        const int c0 = c * CH, cn = min(CH, bi.w - c0);
        const float4 *buf = stage + (c & 1) * (CH + 1) * RW;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PT; ++i) ucol[i] = uni[bi.z + min(c0 + CH + sr + 64 * i, bi.w - 1)];
#pragma unroll
        for (int i = 0; i < PT; ++i) pre[i] = FV<4>::load(X + (int64_t)ucol[i] * k + scol);
        for (;;) {
            // the window starts at q (w0 == q): its positions in this chunk are a prefix (the slots
            // ascend), of at most 32 entries per group; 8-entry spans no group needs are skipped
            const int e0 = w0 + gl, e1 = w0 + 16 + gl;
            const bool in0 = has && e0 < ri.z && sl0 < c0 + cn;
            const bool in1 = has && e1 < ri.z && sl1 < c0 + cn;
            const int o0 = (in0 ? sl0 - c0 : CH) * RW, o1 = (in1 ? sl1 - c0 : CH) * RW;
            const uint32_t u0 = in0 ? __float_as_uint(va0) : 0u, u1 = in1 ? __float_as_uint(va1) : 0u;
            const uint64_t m0 = __ballot(in0), m1 = __ballot(in1);
            constexpr uint64_t LO = 0x00ff00ff00ff00ffull, HI = 0xff00ff00ff00ff00ull;
            if (m0 & LO) spmm_slice_span<0>(buf, o0, u0, gl, acc, std::make_integer_sequence<int, 8>{});
            if (m0 & HI) spmm_slice_span<8>(buf, o0, u0, gl, acc, std::make_integer_sequence<int, 8>{});
            if (m1 & LO) spmm_slice_span<0>(buf, o1, u1, gl, acc, std::make_integer_sequence<int, 8>{});
            if (m1 & HI) spmm_slice_span<8>(buf, o1, u1, gl, acc, std::make_integer_sequence<int, 8>{});
            const int sh = lane & ~15;
            const int cnt = __popc((uint32_t)(m0 >> sh) & 0xffffu) + __popc((uint32_t)(m1 >> sh) & 0xffffu);
            w0 += cnt;   // the next window starts at the next entry
            const bool more = has && cnt == 32 && w0 < ri.z;   // the window lay wholly in the chunk
            sl0 = (int)slot[min(w0 + gl, last)];
            sl1 = (int)slot[min(w0 + 16 + gl, last)];
            va0 = vA[min(w0 + gl, last)];
            va1 = vA[min(w0 + 16 + gl, last)];
            if (!__ballot(more)) break;
        }
        float4 *nb = stage + ((c + 1) & 1) * (CH + 1) * RW;
        const int nn = min(CH, bi.w - c0 - CH);
#pragma unroll
        for (int i = 0; i < PT; ++i)
            if (sr + 64 * i < nn) nb[(sr + 64 * i) * RW + sc] = make_float4(pre[i].v[0], pre[i].v[1], pre[i].v[2], pre[i].v[3]);
    }
    if (has && col < k) epi.template apply<4>(ri.x, col, acc);
}

// The tiled layout of M for tile shape i: rows are pixels (j = x * N + y, tiles of STILE[i] pixels)
// or rays (r = angle * p + detector, tiles of SRAY[i] angles x detectors).  Built on the host from
// M's pattern once per shape and cached in M; false when it does not apply (a column outside the
// matrix, or a row that repeats a column).
bool build_stiles(const fair_csr *h, Csr &M, bool rays, int shape) {
    if (M.stile.code == shape) return M.stile.nblocks > 0;
    M.stile.code = shape;
    M.stile.nblocks = 0;
    // rows r = u * V + v: pixels j = x * N + y of an N x N image, or rays r = angle * p + detector
    const int64_t U = rays ? h->g_nth : h->g_n, V = rays ? h->g_p : h->g_n;
    if (U <= 0 || V <= 0 || (int64_t)M.m != U * V || M.nnz <= 0) return false;
    if (shape < 0 || shape >= (rays ? SRAY_SHAPES : STILE_SHAPES)) return false;
    std::vector<int32_t> IA((size_t)M.m + 1), JA((size_t)M.nnz);
    M.IA.download(IA.data(), (size_t)M.m + 1);
    M.JA.download(JA.data(), (size_t)M.nnz);
    FAIR_HIP(hipStreamSynchronize(stream()));
    const int TU = rays ? SRAY[shape][0] : STILE[shape][0], TV = rays ? SRAY[shape][1] : STILE[shape][1];
    std::vector<int4> binfo, rows;
    std::vector<int32_t> uni;
    std::vector<uint16_t> slot((size_t)M.nnz);
    std::vector<int32_t> mark((size_t)std::max(M.n, 1), -1), pos((size_t)std::max(M.n, 1), 0);
    std::vector<int32_t> u, indeg, first, succ, order, fill;
    int32_t stamp = 0;
    // One block of the layout from rows rw[0, n) (non-empty, in tile order): the union of their
    // columns in an order every row follows (sorted for pixels, whose rays ascend; Kahn's algorithm
    // on the precedence of consecutive entries for rays).  Where there is no such order, or the
    // union is too large, the rows are split in two halves (the earlier angles first) and placed
    // separately.  0: the layout does not apply (a column outside the matrix, or a single row that
    // repeats a column).
    std::function<int(const int64_t *, size_t)> place = [&](const int64_t *rw, size_t n) -> int {
        ++stamp;
        u.clear();
        for (size_t i = 0; i < n; ++i)
            for (int32_t e = IA[rw[i]]; e < IA[rw[i] + 1]; ++e) {
                const int32_t c = JA[e];
                if ((uint32_t)c >= (uint32_t)M.n) return 0;
                if (mark[c] != stamp) {
                    mark[c] = stamp;
                    pos[c] = (int32_t)u.size();
                    u.push_back(c);
                }
            }
        bool ok = u.size() <= 65535;
        if (ok && !rays) {
            std::sort(u.begin(), u.end());
        } else if (ok) {
            const size_t nu = u.size();
            indeg.assign(nu, 0);
            first.assign(nu + 1, 0);
            for (size_t i = 0; i < n; ++i)
                for (int32_t e = IA[rw[i]] + 1; e < IA[rw[i] + 1]; ++e) ++first[(size_t)pos[JA[e - 1]] + 1];
            for (size_t i = 0; i < nu; ++i) first[i + 1] += first[i];
            succ.assign((size_t)first[nu], 0);
            fill.assign(first.begin(), first.end() - 1);
            for (size_t i = 0; i < n; ++i)
                for (int32_t e = IA[rw[i]] + 1; e < IA[rw[i] + 1]; ++e) {
                    const int32_t to = pos[JA[e]];
                    succ[(size_t)fill[(size_t)pos[JA[e - 1]]]++] = to;
                    ++indeg[(size_t)to];
                }
            order.clear();
            for (size_t i = 0; i < nu; ++i)
                if (indeg[i] == 0) order.push_back((int32_t)i);
            for (size_t q = 0; q < order.size(); ++q)
                for (int32_t f = first[(size_t)order[q]]; f < first[(size_t)order[q] + 1]; ++f)
                    if (--indeg[(size_t)succ[(size_t)f]] == 0) order.push_back(succ[(size_t)f]);
            ok = order.size() == nu;
            if (ok) {
                std::vector<int32_t> ou(nu);
                for (size_t i = 0; i < nu; ++i) ou[i] = u[(size_t)order[i]];
                u.swap(ou);
            }
        }
        if (ok) {
            for (size_t i = 0; i < u.size(); ++i) pos[u[i]] = (int32_t)i;
            // the kernel takes a chunk's entries of a row as a prefix of at most one per union
            // row: every row's slots must strictly ascend
            for (size_t i = 0; i < n && ok; ++i)
                for (int32_t e = IA[rw[i]]; e < IA[rw[i] + 1]; ++e) {
                    slot[e] = (uint16_t)pos[JA[e]];
                    if (e > IA[rw[i]] && slot[e] <= slot[e - 1]) ok = false;
                }
        }
        if (!ok) {
            if (n == 1) return 0;
            return place(rw, n / 2) && place(rw + n / 2, n - n / 2);
        }
        binfo.push_back(make_int4((int32_t)rows.size(), (int32_t)n, (int32_t)uni.size(), (int32_t)u.size()));
        for (size_t i = 0; i < n; ++i) rows.push_back(make_int4((int32_t)rw[i], IA[rw[i]], IA[rw[i] + 1], 0));
        uni.insert(uni.end(), u.begin(), u.end());
        return 1;
    };
    std::vector<int64_t> rw;
    for (int64_t u0 = 0; u0 < U; u0 += TU)
        for (int64_t v0 = 0; v0 < V; v0 += TV) {
            rw.clear();
            for (int64_t a = u0; a < std::min(U, u0 + TU); ++a)
                for (int64_t b = v0; b < std::min(V, v0 + TV); ++b)
                    if (IA[a * V + b + 1] > IA[a * V + b]) rw.push_back(a * V + b);
            if (!rw.empty() && !place(rw.data(), rw.size())) return false;
        }
    if (h->graphs_used) drop_graphs(h);   // captured launches refer to the layout being replaced
    auto up = [](DBuf<int32_t> &d, const void *src, size_t ints) {
        d.alloc(ints > 0 ? ints : 1);
        if (ints) FAIR_HIP(hipMemcpyAsync(d.p, src, ints * sizeof(int32_t), hipMemcpyHostToDevice, stream()));
    };
    up(M.stile.binfo, binfo.data(), 4 * binfo.size());
    up(M.stile.rows, rows.data(), 4 * rows.size());
    up(M.stile.uni, uni.data(), uni.size());
    M.stile.slot.alloc((size_t)M.nnz);
    M.stile.slot.upload(slot.data(), (size_t)M.nnz);
    FAIR_HIP(hipStreamSynchronize(stream()));
    M.stile.nblocks = (int32_t)binfo.size();
    return M.stile.nblocks > 0;
}

// Whether order code `code` selects the tiled kernel for M with k right-hand sides (and its layout
// is built).
bool stiles_apply(const fair_csr *h, Csr &M, bool rays, int code, int32_t k) {
    if (!is_tiled(code) || k % 4 != 0 || (!tiled_sliced(code) && k > 4 * WAVE)) return false;
    if (tiled_lds(code, k) > (size_t)lds_per_block()) return false;
    return build_stiles(h, M, rays, tiled_shape(code));
}

template <class Epi>
void spmm_launch_tiles(const Csr &M, bool rays, const float *X, int32_t k, int code, Epi epi, hipStream_t s) {
    const int64_t blocks = M.stile.nblocks;
    const int64_t per_xcd = (blocks + NXCD - 1) / NXCD;
    const size_t lds = tiled_lds(code, k);
    const int *shp = rays ? SRAY[tiled_shape(code)] : STILE[tiled_shape(code)];
    if (tiled_sliced(code)) {
        auto kern = tiled_chunk(code) == 128 ? k_spmm_tiles_sliced<128, Epi> : k_spmm_tiles_sliced<64, Epi>;
        hipLaunchKernelGGL(kern, dim3((unsigned)(per_xcd * NXCD), (unsigned)((k + STILE_SLICE - 1) / STILE_SLICE)),
                           dim3(1024), lds, s, (const int4 *)M.stile.binfo.p, (int32_t)blocks, (int32_t)per_xcd,
                           (const int4 *)M.stile.rows.p, M.stile.uni.p, M.stile.slot.p, M.vA32.p, X, k, epi);
        return;
    }
    const int rpw = shp[0] * shp[1] / STILE_ROWS;
    auto kern = tiled_chunk(code) == 64
                    ? (rpw == 4 ? k_spmm_tiles<64, 4, Epi> : rpw == 2 ? k_spmm_tiles<64, 2, Epi> : k_spmm_tiles<64, 1, Epi>)
                    : (rpw == 4 ? k_spmm_tiles<32, 4, Epi> : rpw == 2 ? k_spmm_tiles<32, 2, Epi> : k_spmm_tiles<32, 1, Epi>);
    hipLaunchKernelGGL(kern, dim3((unsigned)(per_xcd * NXCD)), dim3(64 * STILE_ROWS), lds, s,
                       (const int4 *)M.stile.binfo.p, (int32_t)blocks, (int32_t)per_xcd, (const int4 *)M.stile.rows.p,
                       M.stile.uni.p, M.stile.slot.p, M.vA32.p, X, k, epi);
}

template <int VEC, class Epi>
void spmm_launch_v(const int4 *rows, int32_t nrows, const Csr &M, const float *X, int32_t k, Epi epi, size_t lds,
                   hipStream_t s) {
    const int64_t blocks = (nrows + SPMM_WAVES - 1) / SPMM_WAVES;
    const int64_t per_xcd = (blocks + NXCD - 1) / NXCD;
    hipLaunchKernelGGL((k_spmm_rows<VEC, Epi>), dim3((unsigned)(per_xcd * NXCD)), dim3(256), lds, s, rows, nrows,
                       (int32_t)per_xcd, M.JA.p, M.vA32.p, M.n, X, k, zero_row(), epi);
}

template <int G, class Epi>
void spmm_launch_g(const int4 *rows, int32_t nrows, const Csr &M, const float *X, int32_t k, Epi epi, size_t lds,
                   hipStream_t s) {
    const int64_t waves = (nrows + WAVE / G - 1) / (WAVE / G);
    const int64_t blocks = (waves + SPMM_WAVES - 1) / SPMM_WAVES;
    const int64_t per_xcd = (blocks + NXCD - 1) / NXCD;
    const int64_t slices = (k + 4 * G - 1) / (4 * G);
    hipLaunchKernelGGL((k_spmm_groups<G, Epi>), dim3((unsigned)(per_xcd * NXCD), (unsigned)slices), dim3(256), lds, s,
                       rows, nrows, (int32_t)per_xcd, M.JA.p, M.vA32.p, M.n, X, k, epi);
}

// Everything the products of a batched run build lazily (row orders, tiled layouts), made ahead.
void spmm_prepare(const fair_csr *h, int32_t k) {
    Csr &A = const_cast<fair_csr *>(h)->A, &T = const_cast<fair_csr *>(h)->T;
    const bool a_rays = !h->transposed;
    if (!stiles_apply(h, A, a_rays, h->spmm_tile[0], k)) spmm_order(h, A, a_rays, h->spmm_tile[0]);
    if (!stiles_apply(h, T, !a_rays, h->spmm_tile[1], k)) spmm_order(h, T, !a_rays, h->spmm_tile[1]);
}

// Y = M * X (FP32) through Epi, rows of M in order `code`.
template <class Epi>
void spmm_launch(const fair_csr *h, Csr &M, bool rays, const float *X, int32_t k, int code, Epi epi,
                 hipStream_t s) {
    if (M.n_rows_nz <= 0 || k <= 0) return;
    if (stiles_apply(h, M, rays, code, k)) {
        spmm_launch_tiles(M, rays, X, k, code, epi, s);
        return;
    }
    const int order = code_order(code), grouping = code_grouping(code);
    const size_t lds = cap_lds(code);
    const int4 *rows = spmm_order(h, M, rays, is_tiled(order) ? 0 : order);
    if (groups_apply(grouping, k, M.n)) {
        if (SPMM_GROUP_LANES[grouping] == 32) spmm_launch_g<32>(rows, M.n_rows_nz, M, X, k, epi, lds, s);
        else spmm_launch_g<16>(rows, M.n_rows_nz, M, X, k, epi, lds, s);
    } else if (k % 4 == 0) spmm_launch_v<4>(rows, M.n_rows_nz, M, X, k, epi, lds, s);
    else spmm_launch_v<1>(rows, M.n_rows_nz, M, X, k, epi, lds, s);
}

// Cached HIP graphs of whole batched runs, keyed by the run's arguments.
// A captured run also bakes in handle-owned pointers (weights, FP32 values, layouts, workspaces):
// whatever replaces them drops the handle's graphs.  The tuned row orders are part of the key.
struct GraphKey {
    const void *h, *B, *X;
    int32_t method, k, maxits, upb, lwrb, ord0, ord1;
    double relax, ub, lb;
    bool operator<(const GraphKey &o) const {
        return std::tie(h, B, X, method, k, maxits, upb, lwrb, ord0, ord1, relax, ub, lb) <
               std::tie(o.h, o.B, o.X, o.method, o.k, o.maxits, o.upb, o.lwrb, o.ord0, o.ord1, o.relax, o.ub,
                        o.lb);
    }
};
std::map<GraphKey, hipGraphExec_t> g_graphs;
hipStream_t g_capture = nullptr;

}  // namespace

// Cached SIRT weights and FP32 values (shared with methods.hip's FP64 set-up).
void sirt_setup_public(fair_csr *h, int method);

namespace {

void batch_iterations(fair_csr *h, int method, const float *dB, int32_t k, int32_t maxits, float *dX,
                      float *dR, float *dXn, const fair_opts *o, hipStream_t s) {
    Csr &A = h->A, &T = h->T;
    const bool a_rays = !h->transposed;   // A's rows are the rays unless the caller transposed it
    const double *M = h->M[method].p;
    const double *D = method == FAIR_SART ? h->D.p : nullptr;
    const double relax = o ? o->relaxpar : 1.0;
    const BoxF box = make_boxf(o);
    float *cur = dX, *nxt = dXn;
    const int oa = h->spmm_tile[0], ot = h->spmm_tile[1];   // orders built by the caller
    for (int it = 0; it < maxits; ++it) {
        spmm_launch(h, A, a_rays, cur, k, oa, EpiResidualB{dB, M, dR, k}, s);
        EpiUpdateB up{cur, D, nxt, k, relax, box};
        spmm_launch(h, T, !a_rays, dR, k, ot, up, s);
        if (T.n_rows_empty > 0) {
            const int64_t t = (int64_t)T.n_rows_empty * k;
            hipLaunchKernelGGL((k_spmm_empty<EpiUpdateB>), dim3(grid_for(t, 256)), dim3(256), 0, s,
                               T.rows_empty.p, T.n_rows_empty, k, up);
        }
        std::swap(cur, nxt);
    }
    if (cur != dX) {
        const int64_t t = (int64_t)A.n * k;
        hipLaunchKernelGGL(k_copy_f32, dim3(grid_for(t, 256)), dim3(256), 0, s, cur, dX, t);
    }
}

}  // namespace
}  // namespace fair

using namespace fair;

extern "C" int fair_sirt_f32_dev(fair_csr *h, int32_t method, const float *d_b, int32_t maxits, float *d_x,
                                 const fair_opts *o) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(maxits >= 0, FAIR_ERR_ARG, "negative iteration count");
        sirt_setup_public(h, method);
        make_f32(h->A);
        make_f32(h->T);
        const Csr &A = h->A, &T = h->T;
        const int32_t m = A.m, n = A.n;
        if (h->ws_r32.n < (size_t)(m > 0 ? m : 1)) h->ws_r32.alloc(m > 0 ? m : 1);
        if (h->ws_x32.n < (size_t)(n > 0 ? n : 1)) h->ws_x32.alloc(n > 0 ? n : 1);
        const double *M = h->M[method].p;
        const double *D = method == FAIR_SART ? h->D.p : nullptr;
        const double relax = o ? o->relaxpar : 1.0;
        const BoxF box = make_boxf(o);
        float *cur = d_x, *nxt = h->ws_x32.p, *r = h->ws_r32.p;
        for (int it = 0; it < maxits; ++it) {
            spmv_f32(A, SrcPtr<float>{cur}, EpiResidualF{d_b, M, r});
            EpiSirtUpdateF up{cur, D, nxt, relax, box};
            spmv_f32(T, SrcPtr<float>{r}, up);
            if (T.n_rows_empty > 0)
                hipLaunchKernelGGL((k_empty_rows<float, SrcPtr<float>, EpiSirtUpdateF>),
                                   dim3(grid_for(T.n_rows_empty, 256)), dim3(256), 0, stream(),
                                   T.rows_empty.p, T.n_rows_empty, SrcPtr<float>{r}, up);
            std::swap(cur, nxt);
        }
        if (cur != d_x)
            hipLaunchKernelGGL(k_copy_f32, dim3(grid_for(n, 256)), dim3(256), 0, stream(), cur, d_x, (int64_t)n);
        launch_check("sirt fp32");
    });
}

extern "C" int fair_sirt_f32(fair_csr *h, int32_t method, const float *b, int32_t maxits, const float *x0,
                             const fair_opts *o, float *xout) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(b && xout, FAIR_ERR_ARG, "null argument");
        const int32_t m = h->A.m, n = h->A.n;
        DBuf<float> db(m > 0 ? m : 1), dx(n > 0 ? n : 1);
        db.upload(b, m);
        if (x0) dx.upload(x0, n);
        else FAIR_HIP(hipMemsetAsync(dx.p, 0, sizeof(float) * (n > 0 ? n : 1), stream()));
        int st = fair_sirt_f32_dev(h, method, db.p, maxits, dx.p, o);
        if (st != FAIR_OK) throw Error{st, fair_last_error()};
        FAIR_HIP(hipStreamSynchronize(stream()));
        dx.download(xout, n);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_sirt_batch_dev(fair_csr *h, int32_t method, const float *d_B, int32_t k, int32_t maxits,
                                   float *d_X, const fair_opts *o, int32_t use_graph) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(k >= 1 && maxits >= 0, FAIR_ERR_ARG, "k must be >= 1 and maxits >= 0");
        sirt_setup_public(h, method);
        make_f32(h->A);
        make_f32(h->T);
        const int32_t m = h->A.m, n = h->A.n;
        const size_t need_r = (size_t)m * k, need_x = (size_t)n * k;
        if (h->ws_rb.n < need_r || h->ws_xb.n < need_x) {
            // captured graphs refer to the old workspaces
            FAIR_HIP(hipStreamSynchronize(stream()));
            if (h->graphs_used) drop_graphs(h);
            if (h->ws_rb.n < need_r) h->ws_rb.alloc(need_r);
            if (h->ws_xb.n < need_x) h->ws_xb.alloc(need_x);
        }
        // FAIR_SPMM_ORDERS=a,t fixes the two products' row orders (tests and experiments)
        if (const char *e = getenv("FAIR_SPMM_ORDERS")) {
            int oa = 0, ot = 0;
            if (sscanf(e, "%d,%d", &oa, &ot) == 2) {
                h->spmm_tile[0] = oa;
                h->spmm_tile[1] = ot;
            }
        }
        // everything a run allocates or looks up is made before a capture can start: the row
        // orders and tiled layouts (cached in the matrices) and the zero row's address
        spmm_prepare(h, k);
        (void)zero_row();
        if (!use_graph) {
            batch_iterations(h, method, d_B, k, maxits, d_X, h->ws_rb.p, h->ws_xb.p, o, stream());
            launch_check("sirt batch");
            return;
        }
        // graph path: capture the whole run once on a private stream, replay afterwards
        if (!g_capture) FAIR_HIP(hipStreamCreateWithFlags(&g_capture, hipStreamNonBlocking));
        GraphKey key{h,        d_B,          d_X,          method,       k,
                     maxits,   o ? o->upb : 0, o ? o->lwrb : 0, h->spmm_tile[0], h->spmm_tile[1],
                     o ? o->relaxpar : 1.0, o ? o->ubound : 0.0, o ? o->lbound : 1.0};
        auto it = g_graphs.find(key);
        if (it == g_graphs.end()) {
            FAIR_HIP(hipStreamSynchronize(stream()));
            hipGraph_t graph;
            FAIR_HIP(hipStreamBeginCapture(g_capture, hipStreamCaptureModeThreadLocal));
            batch_iterations(h, method, d_B, k, maxits, d_X, h->ws_rb.p, h->ws_xb.p, o, g_capture);
            hipError_t le = hipGetLastError();
            FAIR_HIP(hipStreamEndCapture(g_capture, &graph));
            FAIR_HIP(le);
            hipGraphExec_t exec;
            FAIR_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            FAIR_HIP(hipGraphDestroy(graph));
            it = g_graphs.emplace(key, exec).first;
            h->graphs_used = true;
        }
        // order the replay after the caller's work on the library stream
        FAIR_HIP(hipStreamSynchronize(stream()));
        FAIR_HIP(hipGraphLaunch(it->second, g_capture));
        FAIR_HIP(hipStreamSynchronize(g_capture));
    });
}

// Release the graphs that captured this handle (called by fair_csr_free).
namespace fair {
void drop_graphs(const fair_csr *h) {
    for (auto it = g_graphs.begin(); it != g_graphs.end();) {
        if (it->first.h == h) {
            (void)hipGraphExecDestroy(it->second);
            it = g_graphs.erase(it);
        } else {
            ++it;
        }
    }
}
}  // namespace fair

// Per right-hand side relative error ||X_j - Xref_j|| / ||Xref_j|| with FP64 sums (X: n x k FP32,
// Xref: n x k FP64, both row-major).
namespace {
__global__ void __launch_bounds__(256) k_batch_err(const float *X, const double *Xr, int32_t n, int32_t k,
                                                  double *num, double *den) {
    const int j = blockIdx.y;
    double a = 0.0, b = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) {
        const double r = Xr[p * k + j], d = (double)X[p * k + j] - r;
        a += d * d;
        b += r * r;
    }
    a = fair::wave_sum(a);
    b = fair::wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&num[j], a);
        atomicAdd(&den[j], b);
    }
}
}  // namespace

extern "C" int fair_batch_relerr_dev(const float *d_X, const double *d_Xref, int32_t n, int32_t k, double *err) {
    return guard([&] {
        FAIR_CHECK(d_X && d_Xref && err && n >= 0 && k >= 1, FAIR_ERR_ARG, "bad arguments");
        DBuf<double> num(k), den(k);
        FAIR_HIP(hipMemsetAsync(num.p, 0, sizeof(double) * k, stream()));
        FAIR_HIP(hipMemsetAsync(den.p, 0, sizeof(double) * k, stream()));
        hipLaunchKernelGGL(k_batch_err, dim3(64, k), dim3(256), 0, stream(), d_X, d_Xref, n, k, num.p, den.p);
        launch_check("batch error");
        std::vector<double> a(k), b(k);
        num.download(a.data(), k);
        den.download(b.data(), k);
        FAIR_HIP(hipStreamSynchronize(stream()));
        for (int j = 0; j < k; ++j) err[j] = std::sqrt(a[j]) / (b[j] > 0 ? std::sqrt(b[j]) : 1.0);
    });
}

// Autotuner of the batched SpMM: times every row order (sinogram tiles for A, image tiles for
// A^T; only the stored order for a matrix without geometry), each with one row per wavefront and,
// where k allows, with two and four rows per wavefront (passes of 128 and 64 right-hand sides),
// and the LDS-staged tiled kernel's tile shapes and chunk sizes for both products, on
// this matrix with k right-hand sides with HIP events (two launches each, after a warm-up); then
// times the pairs of the SPMM_FINALISTS fastest per product as whole iterations, then the
// residency cap of each row-kernel finalist and the pairs of the capped finalists, and keeps the
// fastest choice in the handle.
constexpr int SPMM_FINALISTS = 3;
constexpr int SPMM_PAIR_ITS = 4;   // even: the iterate ends in the buffer it started in
extern "C" int fair_autotune_batch(fair_csr *h, int32_t k, int32_t *order_out) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(k >= 1, FAIR_ERR_ARG, "k must be >= 1");
        sirt_setup_public(h, FAIR_SART);
        make_f32(h->A);
        make_f32(h->T);
        Csr &A = h->A, &T = h->T;
        const bool a_rays = !h->transposed;
        const bool geo = h->g_n > 0 && h->g_p > 0 && h->g_nth > 0;
        DBuf<float> X((size_t)A.n * k), R((size_t)A.m * k), Y((size_t)std::max(A.m, A.n) * k);
        FAIR_HIP(hipMemsetAsync(X.p, 0, X.bytes(), stream()));
        FAIR_HIP(hipMemsetAsync(R.p, 0, R.bytes(), stream()));
        hipEvent_t e0, e1;
        FAIR_HIP(hipEventCreate(&e0));
        FAIR_HIP(hipEventCreate(&e1));
        const bool dbg = getenv("FAIR_DEBUG") && atoi(getenv("FAIR_DEBUG")) != 0;
        std::vector<std::pair<float, int>> timed[2];   // (ms, code) per product
        for (int dir = 0; dir < 2; ++dir) {
            Csr &M = dir == 0 ? A : T;
            const bool rays = dir == 0 ? a_rays : !a_rays;
            const float *src = dir == 0 ? X.p : R.p;
            std::vector<int> codes;
            for (int gs = 0; gs < SPMM_GROUPINGS; ++gs)
                if (gs == 0 || groups_apply(gs, k, M.n))
                    for (int c = 0; c < (geo ? SPMM_ORDERS : 1); ++c) codes.push_back(c + SPMM_GROUP_CODE * gs);
            for (int ci = 0; ci < STILE_CHUNKS; ++ci)
                for (int c = 0; c < (rays ? SRAY_SHAPES : STILE_SHAPES); ++c)
                    if (geo && k % 4 == 0 && (ci >= 2 || k <= 4 * WAVE)) codes.push_back(STILE_CODE0 + 10 * ci + c);
            for (int code : codes) {
                if (is_tiled(code) && !stiles_apply(h, M, rays, code, k)) continue;
                auto run = [&] { spmm_launch(h, M, rays, src, k, code, EpiStoreB{Y.p, k}, stream()); };
                run();
                FAIR_HIP(hipEventRecord(e0, stream()));
                run();
                run();
                FAIR_HIP(hipEventRecord(e1, stream()));
                FAIR_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                FAIR_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (dbg && is_tiled(code)) {
                    const int *shp = rays ? SRAY[tiled_shape(code)] : STILE[tiled_shape(code)];
                    fprintf(stderr, "[fair] spmm %s tiled %d (%dx%d %s, LDS-staged union, %d-row chunks%s): %.1f us\n",
                            dir == 0 ? "A" : "A^T", code, shp[0], shp[1], rays ? "rays" : "pixels",
                            tiled_chunk(code), tiled_sliced(code) ? " of 64 right-hand sides" : "", ms * 500.f);
                } else if (dbg) {
                    const int oc = code_order(code), gs = code_grouping(code);
                    const int *tl = rays ? RAY_TILES[oc] : PIX_TILES[oc];
                    fprintf(stderr, "[fair] spmm %s order %d (%dx%d%s), %d lanes per row: %.1f us\n",
                            dir == 0 ? "A" : "A^T", code, tl[0] == ALL ? -1 : tl[0], tl[1],
                            tl[2] ? ", band major" : "", SPMM_GROUP_LANES[gs], ms * 500.f);
                }
                timed[dir].emplace_back(ms, code);
            }
            std::stable_sort(timed[dir].begin(), timed[dir].end());
            if (timed[dir].size() > (size_t)SPMM_FINALISTS) timed[dir].resize(SPMM_FINALISTS);
        }
        // The finalists' pairs run as whole SART iterations (residual and update epilogues, both
        // matrices and four panels streaming through the caches as in a run), which the single
        // products above do not show: a product read in several passes re-reads its matrix from
        // HBM there.  Then each row-kernel finalist's residency cap is tuned against the other
        // product's best code, and the pairs of the capped finalists are timed; the fastest is kept.
        DBuf<float> Bz((size_t)A.m * k);
        FAIR_HIP(hipMemsetAsync(Bz.p, 0, Bz.bytes(), stream()));
        auto iteration_ms = [&](int ca, int ct) {
            h->spmm_tile[0] = ca;
            h->spmm_tile[1] = ct;
            spmm_prepare(h, k);
            batch_iterations(h, FAIR_SART, Bz.p, k, 2, X.p, R.p, Y.p, nullptr, stream());
            FAIR_HIP(hipEventRecord(e0, stream()));
            batch_iterations(h, FAIR_SART, Bz.p, k, SPMM_PAIR_ITS, X.p, R.p, Y.p, nullptr, stream());
            FAIR_HIP(hipEventRecord(e1, stream()));
            FAIR_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FAIR_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (dbg) fprintf(stderr, "[fair] spmm orders %d, %d: %.1f us per iteration\n", ca, ct, ms * 1000.f / SPMM_PAIR_ITS);
            return ms;
        };
        // uncapped pairs
        float best = 1e30f;
        int best_pair[2] = {timed[0].empty() ? 0 : timed[0][0].second, timed[1].empty() ? 0 : timed[1][0].second};
        for (const auto &ta : timed[0])
            for (const auto &tt : timed[1]) {
                const float ms = iteration_ms(ta.second, tt.second);
                if (ms < best) {
                    best = ms;
                    best_pair[0] = ta.second;
                    best_pair[1] = tt.second;
                }
            }
        // each row-kernel finalist's best cap, against the other product's best code
        std::vector<int> capped[2];
        for (int dir = 0; dir < 2; ++dir)
            for (const auto &tc : timed[dir]) {
                int bc = tc.second;
                if (!is_tiled(bc)) {
                    int trial[2] = {best_pair[0], best_pair[1]};
                    trial[dir] = tc.second;
                    float bms = iteration_ms(trial[0], trial[1]);
                    for (int cap : SPMM_CAPS) {
                        trial[dir] = tc.second + SPMM_CAP_CODE * cap;
                        const float ms = iteration_ms(trial[0], trial[1]);
                        if (ms < bms) {
                            bms = ms;
                            bc = trial[dir];
                        }
                    }
                }
                capped[dir].push_back(bc);
            }
        // and the pairs of the capped finalists
        for (int ca : capped[0])
            for (int ct : capped[1]) {
                if (!(ca >= SPMM_CAP_CODE || ct >= SPMM_CAP_CODE)) continue;   // timed uncapped above
                const float ms = iteration_ms(ca, ct);
                if (ms < best) {
                    best = ms;
                    best_pair[0] = ca;
                    best_pair[1] = ct;
                }
            }
        h->spmm_tile[0] = best_pair[0];
        h->spmm_tile[1] = best_pair[1];
        spmm_prepare(h, k);
        FAIR_HIP(hipStreamSynchronize(stream()));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        launch_check("autotune batch");
        if (order_out) {
            order_out[0] = h->spmm_tile[0];
            order_out[1] = h->spmm_tile[1];
        }
    });
}

extern "C" int fair_sirt_batch(fair_csr *h, int32_t method, const float *B, int32_t k, int32_t maxits,
                               const float *X0, const fair_opts *o, int32_t use_graph, float *Xout) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(B && Xout && k >= 1, FAIR_ERR_ARG, "bad arguments");
        const size_t mk = (size_t)h->A.m * k, nk = (size_t)h->A.n * k;
        DBuf<float> dB(mk > 0 ? mk : 1), dX(nk > 0 ? nk : 1);
        dB.upload(B, mk);
        if (X0) dX.upload(X0, nk);
        else FAIR_HIP(hipMemsetAsync(dX.p, 0, sizeof(float) * (nk > 0 ? nk : 1), stream()));
        FAIR_HIP(hipStreamSynchronize(stream()));
        int st = fair_sirt_batch_dev(h, method, dB.p, k, maxits, dX.p, o, use_graph);
        if (st != FAIR_OK) throw Error{st, fair_last_error()};
        // a graph keyed by these temporary buffers would dangle: release it
        if (use_graph) drop_graphs(h);
        FAIR_HIP(hipStreamSynchronize(stream()));
        dX.download(Xout, nk);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}
