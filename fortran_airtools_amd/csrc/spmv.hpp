// spmv.hpp -- the CSR row-per-wavefront SpMV kernel with fused epilogues.
//
// A 64-lane wavefront owns SPMV_RPW consecutive non-empty rows of the compacted
// {row,start,end} list (one descriptor load per wave).  Per row, every lane requests
// SPMV_UNROLL aligned pairs of (value, index) with non-temporal loads -- 512 entries, so the
// whole row is in flight before the first gather and the streamed matrix does not evict the
// gathered vector from the XCD's L2 -- then gathers x[JA] (2 MiB at n = 512^2 in FP64, L2/MALL
// resident), accumulates in two chains and reduces across the wavefront.  Column indices outside [0, ncols) contribute nothing (the reference
// can emit such indices, see raytrace.hip).  Lane 0 hands the row result to an epilogue functor, which is where
// SART / Cimmino / gradient-descent steps fuse their vector work.
#pragma once

#include "fair_internal.hpp"

namespace fair {

// Gather source of a launch: returns the vector to gather from, or nullptr to skip the whole
// launch (device-side stopping rules gate later iterations this way).
template <class V>
struct SrcPtr {
    const V *x;
    __device__ __forceinline__ const V *operator()() const { return x; }
};

template <class V>
struct EpiStore {
    V *y;
    __device__ __forceinline__ void operator()(int32_t r, V acc) const { y[r] = acc; }
};

template <class T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}

template <class V>
__device__ __forceinline__ V gather(const V *__restrict__ x, int32_t c, int32_t ncols) {
    return (uint32_t)c < (uint32_t)ncols ? x[c] : (V)0;
}

template <class V>
struct Vec2;
template <>
struct Vec2<double> {
    typedef double __attribute__((ext_vector_type(2))) type;
};
template <>
struct Vec2<float> {
    typedef float __attribute__((ext_vector_type(2))) type;
};
typedef int32_t __attribute__((ext_vector_type(2))) int2v;

// Rows per wavefront: the wave loads the descriptors of its rows with one load and walks them.
constexpr int SPMV_RPW = 4;
// Pair-chunks in flight per lane: 2 * 64 * SPMV_UNROLL entries of the row are requested before
// the first gather (512 entries: a whole average fan-beam row at N = 512).
constexpr int SPMV_UNROLL = 4;
static_assert(CSR_PAD >= 2, "pair loads need a spare entry after the last row");

template <class V, class Src, class Epi>
__global__ void __launch_bounds__(256) k_spmv_wave(const int4 *__restrict__ rows, int32_t nrows,
                                                   const int32_t *__restrict__ JA,
                                                   const V *__restrict__ vA, int32_t ncols,
                                                   Src src, Epi epi) {
    typedef typename Vec2<V>::type V2;
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (WAVE - 1);
    const int r0 = wv * SPMV_RPW;
    if (r0 >= nrows) return;
    const V *__restrict__ x = src();
    if (!x) return;
    const int nr = min(SPMV_RPW, nrows - r0);
    int4 d = make_int4(0, 0, 0, 0);
    if (lane < nr) d = rows[r0 + lane];
    for (int q = 0; q < nr; ++q) {
        const int row = __shfl(d.x, q, WAVE), s = __shfl(d.y, q, WAVE), e = __shfl(d.z, q, WAVE);
        V acc0 = 0, acc1 = 0;
        // pairs start at even indices: 16-byte aligned FP64 loads, 8-byte aligned index loads
        for (int base = (s & ~1) + 2 * lane; base < e; base += 2 * WAVE * SPMV_UNROLL) {
            V2 v[SPMV_UNROLL];
            int2v c[SPMV_UNROLL];
#pragma unroll
            for (int u = 0; u < SPMV_UNROLL; ++u) {
                const int k = base + 2 * WAVE * u;
                if (k < e) {
                    v[u] = __builtin_nontemporal_load((const V2 *)(vA + k));
                    c[u] = __builtin_nontemporal_load((const int2v *)(JA + k));
                } else {
                    v[u] = V2{0, 0};
                    c[u] = int2v{-1, -1};
                }
            }
#pragma unroll
            for (int u = 0; u < SPMV_UNROLL; ++u) {
                const int k = base + 2 * WAVE * u;
                if (k >= s && k < e) acc0 += v[u].x * gather(x, c[u].x, ncols);
                if (k + 1 < e) acc1 += v[u].y * gather(x, c[u].y, ncols);
            }
        }
        const V acc = wave_sum(acc0 + acc1);
        if (lane == 0) epi(row, acc);
    }
}

template <class V, class Src, class Epi>
inline void spmv_launch_raw(const int4 *rows, int32_t nrows, const int32_t *JA, const V *vA,
                            int32_t ncols, Src src, Epi epi, hipStream_t s) {
    if (nrows <= 0) return;
    const int waves = (nrows + SPMV_RPW - 1) / SPMV_RPW;
    hipLaunchKernelGGL((k_spmv_wave<V, Src, Epi>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                       s, rows, nrows, JA, vA, ncols, src, epi);
}

template <class Src, class Epi>
inline void spmv_launch_src(const Csr &A, Src src, Epi epi) {
    spmv_launch_raw<double, Src, Epi>((const int4 *)A.rows_nz.p, A.n_rows_nz, A.JA.p, A.vA.p, A.n,
                                      src, epi, stream());
    launch_check("spmv");
}

template <class Epi>
inline void spmv_launch(const Csr &A, const double *x, Epi epi) {
    spmv_launch_src(A, SrcPtr<double>{x}, epi);
}

// Rows with no entries still take their epilogue with a zero sum (a column of A that no ray
// hits still gets the box projection in SIRT / gradient descent).
template <class V, class Src, class Epi>
__global__ void k_empty_rows(const int32_t *__restrict__ rows, int32_t nrows, Src src, Epi epi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows || !src()) return;
    epi(rows[i], (V)0);
}

}  // namespace fair
