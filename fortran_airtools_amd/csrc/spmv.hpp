// spmv.hpp -- the CSR row-per-wavefront SpMV kernel with fused epilogues.
//
// One 64-lane wavefront per non-empty row (rows taken from the compacted {row,start,end}
// list).  Lanes stream vA / JA with coalesced, non-temporal loads (so the streamed matrix does
// not evict the gathered vector from the XCD's L2), gather x[JA] (the vector stays L2/MALL
// resident: 2 MiB for n = 512^2 in FP64), accumulate in two independent chains, and reduce
// across the wavefront.  Column indices outside [0, ncols) contribute nothing (the reference
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

template <class V, class Src, class Epi>
__global__ void __launch_bounds__(256) k_spmv_wave(const int4 *__restrict__ rows, int32_t nrows,
                                                   const int32_t *__restrict__ JA,
                                                   const V *__restrict__ vA, int32_t ncols,
                                                   Src src, Epi epi) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (WAVE - 1);
    if (w >= nrows) return;
    const V *__restrict__ x = src();
    if (!x) return;
    const int4 ri = rows[w];
    V acc0 = 0, acc1 = 0;
    int k = ri.y + lane;
    for (; k + WAVE < ri.z; k += 2 * WAVE) {
        int32_t c0 = ldnt(JA + k), c1 = ldnt(JA + k + WAVE);
        V v0 = ldnt(vA + k), v1 = ldnt(vA + k + WAVE);
        acc0 += v0 * gather(x, c0, ncols);
        acc1 += v1 * gather(x, c1, ncols);
    }
    if (k < ri.z) acc0 += ldnt(vA + k) * gather(x, ldnt(JA + k), ncols);
    V acc = wave_sum(acc0 + acc1);
    if (lane == 0) epi(ri.x, acc);
}

template <class V, class Src, class Epi>
inline void spmv_launch_raw(const int4 *rows, int32_t nrows, const int32_t *JA, const V *vA,
                            int32_t ncols, Src src, Epi epi, hipStream_t s) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL((k_spmv_wave<V, Src, Epi>), dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0,
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
