// spmv.hpp -- CSR SpMV kernels with fused epilogues.
//
// Two kernels share the row walk: a 64-lane wavefront owns a row, every lane requests
// SPMV_UNROLL aligned pairs of (value, column) with non-temporal loads (512 entries: a whole
// average fan-beam row at N = 512 is in flight before the first use), accumulates in two chains
// and reduces across the wavefront; lane 0 hands the row result to an epilogue functor, which
// is where SART / Cimmino / gradient-descent steps fuse their vector work.
//
//  k_spmv_wave     rows from the compacted {row,start,end} list, SPMV_RPW rows per wavefront,
//                  columns gathered from global memory (any CSR matrix).
//  k_spmv_blocked  one workgroup per block of rows (blocked.hip): the block's column footprint is
//                  first copied into LDS with coalesced segment loads, then entries read LDS
//                  through a 16-bit local index -- no per-entry global gathers and 10 instead of
//                  12 streamed bytes per FP64 entry.  Blocks whose footprint does not fit fall
//                  back to global gathers inside the same launch.
//
// Column indices outside [0, ncols) contribute nothing (the reference can emit such indices,
// see raytrace.hip).
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

// Rows per wavefront of the direct kernel: one descriptor load serves SPMV_RPW rows.
constexpr int SPMV_RPW = 4;
// Pair-chunks in flight per lane (2 * 64 * SPMV_UNROLL entries of a row requested at once).
constexpr int SPMV_UNROLL = 4;
static_assert(CSR_PAD >= 2, "pair loads need a spare entry after the last row");

// One row, columns gathered from global memory.  Pairs start at even entry indices, so FP64
// pair loads are 16-byte and index pair loads 8-byte aligned; the pair straddling a row end
// reads one entry of the next row (or of the padding) and masks it.
template <class V>
__device__ __forceinline__ V row_direct(int s, int e, int lane, const int32_t *__restrict__ JA,
                                        const V *__restrict__ vA, const V *__restrict__ x,
                                        int32_t ncols) {
    typedef typename Vec2<V>::type V2;
    V acc0 = 0, acc1 = 0;
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
        // Every gather is issued, with an out-of-range column clamped to 0 (a matrix with entries
        // has ncols >= 1), so that all of them are in flight together; a term outside the row or
        // the matrix then adds +0, which leaves a sum that starts at +0 unchanged.
        V g0[SPMV_UNROLL], g1[SPMV_UNROLL];
#pragma unroll
        for (int u = 0; u < SPMV_UNROLL; ++u) {
            g0[u] = x[(uint32_t)c[u].x < (uint32_t)ncols ? c[u].x : 0];
            g1[u] = x[(uint32_t)c[u].y < (uint32_t)ncols ? c[u].y : 0];
        }
#pragma unroll
        for (int u = 0; u < SPMV_UNROLL; ++u) {
            const int k = base + 2 * WAVE * u;
            const bool in0 = k >= s && k < e && (uint32_t)c[u].x < (uint32_t)ncols;
            const bool in1 = k + 1 < e && (uint32_t)c[u].y < (uint32_t)ncols;
            acc0 += in0 ? v[u].x * g0[u] : (V)0;
            acc1 += in1 ? v[u].y * g1[u] : (V)0;
        }
    }
    return wave_sum_dpp(acc0 + acc1);
}

// One chunk (2 * 64 * SPMV_UNROLL entries) of a row held in registers: values and 16-bit
// local indices of the LDS footprint.
template <class V>
struct StagedChunk {
    typename Vec2<V>::type v[SPMV_UNROLL];
    uint32_t l[SPMV_UNROLL];

    __device__ __forceinline__ void load(int base, int e, const uint16_t *__restrict__ loc,
                                         const V *__restrict__ vA) {
        typedef typename Vec2<V>::type V2;
#pragma unroll
        for (int u = 0; u < SPMV_UNROLL; ++u) {
            const int k = base + 2 * WAVE * u;
            if (k < e) {
                v[u] = __builtin_nontemporal_load((const V2 *)(vA + k));
                l[u] = __builtin_nontemporal_load((const uint32_t *)(loc + k));
            } else {
                v[u] = V2{0, 0};
                l[u] = 0;
            }
        }
    }

    // Every LDS slot read here is valid (a pair's slots are below the layout's LDS capacity, also
    // for the entry before an odd row start and past the row end), so the reads are issued
    // together; entries outside the row are then dropped by selects, which keeps each lane's sums
    // exactly those of its own entries, in order.
    __device__ __forceinline__ void use(int base, int s, int e, const V *lx, V &acc0, V &acc1) const {
        V a[SPMV_UNROLL], b[SPMV_UNROLL];
#pragma unroll
        for (int u = 0; u < SPMV_UNROLL; ++u) {
            a[u] = lx[l[u] & 0xffffu];
            b[u] = lx[l[u] >> 16];
        }
#pragma unroll
        for (int u = 0; u < SPMV_UNROLL; ++u) {
            const int k = base + 2 * WAVE * u;
            const V t0 = v[u].x * a[u], t1 = v[u].y * b[u];
            acc0 = (k >= s && k < e) ? acc0 + t0 : acc0;
            acc1 = (k + 1 < e) ? acc1 + t1 : acc1;
        }
    }
};

template <class V, class Src, class Epi>
__global__ void __launch_bounds__(256) k_spmv_wave(const int4 *__restrict__ rows, int32_t nrows,
                                                   const int32_t *__restrict__ JA,
                                                   const V *__restrict__ vA, int32_t ncols,
                                                   Src src, Epi epi) {
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
        const V acc = row_direct<V>(s, e, lane, JA, vA, x, ncols);
        if (lane == 0) epi(row, acc);
    }
}

// Workgroup of SPMV_BLOCK_THREADS: each wavefront first requests the first chunk of its first
// row (independent of LDS), then all threads stage the footprint (one thread per segment of at
// most SEG_CHUNK contiguous columns), and after the barrier the prefetched chunk is consumed,
// so the matrix stream and the footprint loads are in flight together.
constexpr int SPMV_BLOCK_THREADS = 512;

// Copies a block's footprint (segments of at most SEG_CHUNK = 16 contiguous columns) from x into
// LDS.  Each wavefront takes 64 segments per round, one descriptor per lane (one coalesced load),
// and copies them four at a time: 16 consecutive lanes per segment, so a load instruction touches
// a few cache lines instead of one line per lane (the L1 tag lookups are what bounds the
// footprint gathers).  Eight such loads are in flight before their LDS stores.
template <class V>
__device__ __forceinline__ void stage_footprint(const int4 bi, const int2 *__restrict__ segs,
                                                const V *__restrict__ x, V *lx) {
    static_assert(SEG_CHUNK == 16, "16 lanes copy one segment");
    constexpr int NW = SPMV_BLOCK_THREADS / WAVE, PHASE = 8;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & (WAVE - 1);
    const int sub = lane & 15, grp = lane >> 4;
    for (int s0 = bi.z + wv * WAVE; s0 < bi.w; s0 += NW * WAVE) {
        const int2 mine = s0 + lane < bi.w ? segs[s0 + lane] : make_int2(0, 0);
        const int n = min(WAVE, bi.w - s0);             // segments of this round (wave-uniform)
        for (int k0 = 0; 4 * k0 < n; k0 += PHASE) {
            V v[PHASE];
            int dst[PHASE];
#pragma unroll
            for (int k = 0; k < PHASE; ++k) {
                const int j = 4 * (k0 + k) + grp;       // segment of this lane group
                const int cx = __shfl(mine.x, j & (WAVE - 1), WAVE);
                const int cy = __shfl(mine.y, j & (WAVE - 1), WAVE);
                const bool on = j < n && sub < (cy >> 16);
                dst[k] = on ? (cy & 0xffff) + sub : -1;
                if (on) v[k] = x[cx + sub];
            }
#pragma unroll
            for (int k = 0; k < PHASE; ++k)
                if (dst[k] >= 0) lx[dst[k]] = v[k];
        }
    }
}

// One block of rows (blocked layout) by a workgroup of SPMV_BLOCK_THREADS; lx: LDS of the
// block's footprint (zero_slot + 1 entries).  Shared by the single-matrix kernel and the grouped
// one (group.hip), so both compute every row with the same operations in the same order.
template <class V, class Epi>
__device__ __forceinline__ void spmv_block(const int4 bi, const int4 *__restrict__ rows,
                                           const int2 *__restrict__ segs, const uint16_t *__restrict__ loc,
                                           const int32_t *__restrict__ JA, const V *__restrict__ vA,
                                           int32_t ncols, int32_t zero_slot, const V *__restrict__ x, Epi epi,
                                           V *lx) {
    constexpr int NW = SPMV_BLOCK_THREADS / WAVE;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & (WAVE - 1);
    if (bi.z >= 0) {
        int q = bi.x + wv;
        int4 ri = make_int4(0, 0, 0, 0);
        StagedChunk<V> ch;
        if (q < bi.y) {
            ri = rows[q];
            ch.load((ri.y & ~1) + 2 * lane, ri.z, loc, vA);
        }
        stage_footprint<V>(bi, segs, x, lx);
        if (threadIdx.x == 0) lx[zero_slot] = (V)0;
        __syncthreads();
        bool first = true;
        for (; q < bi.y; q += NW) {
            if (!first) ri = rows[q];
            V acc0 = 0, acc1 = 0;
            int base = (ri.y & ~1) + 2 * lane;
            if (first) {
                ch.use(base, ri.y, ri.z, lx, acc0, acc1);
                base += 2 * WAVE * SPMV_UNROLL;
                first = false;
            }
            for (; base < ri.z; base += 2 * WAVE * SPMV_UNROLL) {
                ch.load(base, ri.z, loc, vA);
                ch.use(base, ri.y, ri.z, lx, acc0, acc1);
            }
            const V acc = wave_sum_dpp(acc0 + acc1);
            if (lane == 0) epi(ri.x, acc);
        }
    } else {
        for (int q = bi.x + wv; q < bi.y; q += NW) {
            const int4 ri = rows[q];
            const V acc = row_direct<V>(ri.y, ri.z, lane, JA, vA, x, ncols);
            if (lane == 0) epi(ri.x, acc);
        }
    }
}

template <class V, class Src, class Epi>
__global__ void __launch_bounds__(SPMV_BLOCK_THREADS) k_spmv_blocked(
    const int4 *__restrict__ binfo, const int4 *__restrict__ rows, const int2 *__restrict__ segs,
    const uint16_t *__restrict__ loc, const int32_t *__restrict__ JA, const V *__restrict__ vA,
    int32_t ncols, int32_t zero_slot, Src src, Epi epi) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const V *__restrict__ x = src();
    if (!x) return;
    spmv_block<V, Epi>(binfo[blockIdx.x], rows, segs, loc, JA, vA, ncols, zero_slot, x, epi, (V *)smem_raw);
}

template <class V, class Src, class Epi>
inline void spmv_launch_raw(const int4 *rows, int32_t nrows, const int32_t *JA, const V *vA,
                            int32_t ncols, Src src, Epi epi, hipStream_t s) {
    if (nrows <= 0) return;
    const int waves = (nrows + SPMV_RPW - 1) / SPMV_RPW;
    hipLaunchKernelGGL((k_spmv_wave<V, Src, Epi>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                       s, rows, nrows, JA, vA, ncols, src, epi);
}

template <class V, class Src, class Epi>
inline void spmv_launch_blocked(const Blocked &B, const int32_t *JA, const V *vA, int32_t ncols,
                                Src src, Epi epi, hipStream_t s) {
    if (B.nblocks <= 0) return;
    const size_t shm = (size_t)B.lds_elems * sizeof(V);
    hipLaunchKernelGGL((k_spmv_blocked<V, Src, Epi>), dim3((unsigned)B.nblocks), dim3(SPMV_BLOCK_THREADS), shm, s,
                       (const int4 *)B.binfo.p, (const int4 *)B.rows.p, (const int2 *)B.segs.p,
                       B.loc.p, JA, vA, ncols, B.lds_elems - 1, src, epi);
}

template <class Src, class Epi>
inline void spmv_launch_src(const Csr &A, Src src, Epi epi) {
    if (A.blk.ok)
        spmv_launch_blocked<double, Src, Epi>(A.blk, A.JA.p, A.vA.p, A.n, src, epi, stream());
    else
        spmv_launch_raw<double, Src, Epi>((const int4 *)A.rows_nz.p, A.n_rows_nz, A.JA.p, A.vA.p,
                                          A.n, src, epi, stream());
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
