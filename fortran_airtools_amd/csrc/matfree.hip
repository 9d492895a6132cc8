// matfree.hip -- matrix-free fan-beam projector: A*x and its adjoint recomputed per call.
//
// No counterpart in the reference (its to-do "make matrix-free implementations", README.md:7);
// the geometry is the one of fanlineartomo (system_generation.f90:98-157): the same angle and ray
// tables, computed on the host, so that at small N the operator equals the stored CSR matrix up
// to FP32 rounding of the entries.
//
// Both products walk rays through the pixel grid (Siddon / incremental traversal): a ray is
// clipped to a box of pixels, then every step moves to the nearer of the next x-line and the next
// y-line crossing.  Crossing parameters come straight from the line index in FP64,
// t = ((L - n/2) - x0) * (1/a), never accumulated, so long rays do not drift, and a walk clipped
// to a sub-box visits exactly the segments of the full walk that lie in it.  Pixel values, segment
// lengths and sums are FP32.
//
//  forward  one thread per ray over the whole image; MF_BATCH segments are generated before
//           their gathers are issued, and the next MF_BATCH are walked while those are in
//           flight (the gathers hit the last-level cache, and latency bounds a one-load-at-a-time
//           walk).
//  adjoint  one workgroup per (TS x TS image tile, block of MF_AB angles): for each angle the
//           rays that can cross the tile form a contiguous range of each monotone run of
//           tan(omega) (found by binary search); their segments inside the tile are accumulated
//           into the tile in LDS as 64-bit fixed-point integers, and the tile is added to a
//           fixed-point image once with coalesced integer atomics.  The segment lengths are the
//           forward walk's, so the adjoint is the transpose of the forward up to the rounding of
//           the sums (the dot-product test checks it), and it is deterministic.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "fair_internal.hpp"

namespace fair {
constexpr int MF_MAX_RUNS = 4;
}

struct fair_mf {
    int32_t n = 0, p = 0, nth = 0;
    int64_t m = 0;
    fair::DBuf<double> ang;    // per angle {x0th, y0th, ath, bth}
    fair::DBuf<double> rayc;   // per ray {cos omega, sin omega}
    fair::DBuf<double> tanw;   // per ray tan omega
    int32_t nruns = 0;
    int32_t runs[2 * fair::MF_MAX_RUNS] = {};   // [begin, end) of the increasing runs of tan omega
    // adjoint workspace: 64-bit fixed-point image (kept zero between calls) and max |r| bits;
    // calls on one handle must not overlap
    fair::DBuf<unsigned long long> xacc;
    fair::DBuf<unsigned> rmax;
    int32_t ecap = 0;   // scale exponent bound of the fixed-point sums
    // rays within ~1e-13 of a grid line: their entries as the matrix build computes them
    // (a small CSR: ray ids, offsets, 0-based columns, FP32 lengths), applied by table
    fair::DBuf<int32_t> ex_rays, ex_offs, ex_cols;
    fair::DBuf<float> ex_vals;
    int32_t n_exact = 0;
    int64_t ex_segments = 0;    // table entries of positive length
    fair::DBuf<uint8_t> flag;   // per ray: 1 = in the table (skipped by the traversal kernels)
    // forward workspace: a copy of x, then x transposed (see k_mf_transpose); calls on one
    // handle must not overlap
    fair::DBuf<float> xx;
    bool mf_gather_buffer = true;   // gathers through a buffer resource (workspace below 4 GiB)
};

namespace fair {
namespace {

// segments per gather batch (forward)
constexpr int MF_BATCH = 12;
constexpr int MF_TS = 64;     // adjoint image tile side
constexpr int MF_AB = 64;     // adjoint angles per workgroup
// LDS row stride of the adjoint tile: an odd stride puts x-neighbours (rays of one angle that
// run mostly along y hit them in the same step) in different banks.
constexpr int MF_LDS_STRIDE = MF_TS + 1;

struct MfArgs {
    int32_t n, p, nth, nruns;
    const double *ang, *rayc, *tanw;
    const uint8_t *flag;
    int32_t runs[2 * MF_MAX_RUNS];
};

// Line of ray (i, j): origin (x0, y0) = rotated source, direction (a, b) = central direction
// rotated by omega_j (system_generation.f90:146-157).
__device__ __forceinline__ void ray_line(const MfArgs &g, int i, int j, double &x0, double &y0, double &a,
                                         double &b) {
    x0 = g.ang[4 * i];
    y0 = g.ang[4 * i + 1];
    const double ath = g.ang[4 * i + 2], bth = g.ang[4 * i + 3];
    const double co = g.rayc[2 * j], so = g.rayc[2 * j + 1];
    a = co * ath - so * bth;
    b = so * ath + co * bth;
}

// Incremental traversal of one line through the pixel box [bx0, bx1) x [by0, by1) (pixel
// indices: ix = floor(x + n/2), iy = floor(y + n/2)).  A line with a == 0 belongs to the column
// floor(x0 + n/2) (half-open pixels), likewise b == 0.
struct Walk {
    double x0, y0, h, inva, invb, len, t, t1, tx, ty;
    // Per axis: dx = L - n/2 for the next grid line L (exact: integers and halves), and the line's
    // t = (dx - qx) * kx, i.e. tline's ((L - n/2) - x0) * (1/a) with the same roundings.  A step adds
    // +-1 to dx; the other axis recomputes its unchanged t.  An axis the ray runs parallel to keeps
    // t = 1e300 (dx = 1e300, qx = 0, kx = 1, no steps), which compares like the infinity it stands for.
    double dx, dy, qx, qy, kx, ky;
    int sxhi, syhi;   // high words of the doubles +-1.0 (their low words are 0), or 0 for no steps
    int ix, iy, sx, sy, ixe, iye;
    // a caller's linear index of pixel (ix, iy), stepped with ix and iy (set by track)
    int lin, linx, liny;
    bool live;

    // lin = index of the current pixel; it changes by dx_lin when ix steps by sx, by dy_lin when iy
    // steps by sy
    __device__ __forceinline__ void track(int lin0, int dx_lin, int dy_lin) {
        lin = lin0;
        linx = dx_lin;
        liny = dy_lin;
    }

    // t of the grid line with index L: ((L - n/2) - x0) * (1/a).  L - n/2 is exact, so a line
    // through the ray's own start coordinate gets t = 0 exactly, as the reference's
    // (x - x0th) / a (system_generation.f90:160) does; a fused L * (1/a) + c form cancels
    // catastrophically for nearly axis-aligned rays (a ~ 1e-16 at theta = pi).
    __device__ __forceinline__ double tline(int L, double p0, double inv) const {
        return (((double)L - h) - p0) * inv;
    }
    __device__ __forceinline__ double tX(int L) const { return tline(L, x0, inva); }
    __device__ __forceinline__ double tY(int L) const { return tline(L, y0, invb); }

    __device__ __forceinline__ void init(double px0, double py0, double a, double b, int n, int bx0, int bx1,
                                         int by0, int by1) {
        live = false;
        h = 0.5 * (double)n;
        x0 = px0;
        y0 = py0;
        len = sqrt(a * a + b * b);
        if (!(len > 0.0)) return;
        inva = a != 0.0 ? 1.0 / a : 0.0;
        invb = b != 0.0 ? 1.0 / b : 0.0;
        double txe, txl, tye, tyl;
        if (a != 0.0) {
            txe = tX(a > 0.0 ? bx0 : bx1);
            txl = tX(a > 0.0 ? bx1 : bx0);
        } else {
            const double xc = x0 + h;
            if (!(xc >= (double)bx0 && xc < (double)bx1)) return;
            txe = -INFINITY;
            txl = INFINITY;
        }
        if (b != 0.0) {
            tye = tY(b > 0.0 ? by0 : by1);
            tyl = tY(b > 0.0 ? by1 : by0);
        } else {
            const double yc = y0 + h;
            if (!(yc >= (double)by0 && yc < (double)by1)) return;
            tye = -INFINITY;
            tyl = INFINITY;
        }
        const double t0 = fmax(txe, tye);
        t1 = fmin(txl, tyl);
        if (!(t1 > t0)) return;
        sx = a > 0.0 ? 1 : -1;
        sy = b > 0.0 ? 1 : -1;
        // entry pixel: exact on the axis whose boundary is crossed, floor on the other
        if (a == 0.0) ix = (int)floor(x0 + h);
        else if (txe >= tye) ix = a > 0.0 ? bx0 : bx1 - 1;
        else ix = min(max((int)floor(x0 + t0 * a + h), bx0), bx1 - 1);
        if (b == 0.0) iy = (int)floor(y0 + h);
        else if (tye >= txe) iy = b > 0.0 ? by0 : by1 - 1;
        else iy = min(max((int)floor(y0 + t0 * b + h), by0), by1 - 1);
        if (a != 0.0) {
            dx = (double)(ix + (sx > 0)) - h;
            qx = x0;
            kx = inva;
            sxhi = __double2hiint((double)sx);
        } else {
            dx = 1e300;
            qx = 0.0;
            kx = 1.0;
            sxhi = 0;
        }
        if (b != 0.0) {
            dy = (double)(iy + (sy > 0)) - h;
            qy = y0;
            ky = invb;
            syhi = __double2hiint((double)sy);
        } else {
            dy = 1e300;
            qy = 0.0;
            ky = 1.0;
            syhi = 0;
        }
        tx = (dx - qx) * kx;   // = tX(ix + (sx > 0)) for a != 0
        ty = (dy - qy) * ky;
        ixe = sx > 0 ? bx1 : bx0 - 1;
        iye = sy > 0 ? by1 : by0 - 1;
        t = t0;
        live = true;
    }

    // The next segment (pixel ix, iy; length seg >= 0).  Call only while live; every call moves
    // ix or iy one pixel towards the box edge or ends the walk.  Pieces of length <= 1e-10 count
    // as 0, as the matrix build drops them (system_generation.f90:196).
    __device__ __forceinline__ void next(int &oix, int &oiy, float &seg) {
        const double tn = fmin(fmin(tx, ty), t1);
        const double d = (tn - t) * len;
        seg = d > 1e-10 ? (float)d : 0.f;
        oix = ix;
        oiy = iy;
        t = tn;
        const bool stepx = tx <= ty;
        ix += stepx ? sx : 0;
        iy += stepx ? 0 : sy;
        live = tn < t1 && ix != ixe && iy != iye;
        dx += __hiloint2double(stepx ? sxhi : 0, 0);
        dy += __hiloint2double(stepx ? 0 : syhi, 0);
        tx = (dx - qx) * kx;
        ty = (dy - qy) * ky;
    }

    // next() for a tracked walk: the segment's pixel as its linear index
    __device__ __forceinline__ void next_lin(int &olin, float &seg) {
        const double tn = fmin(fmin(tx, ty), t1);
        const double d = (tn - t) * len;
        seg = d > 1e-10 ? (float)d : 0.f;
        olin = lin;
        t = tn;
        const bool stepx = tx <= ty;
        ix += stepx ? sx : 0;
        iy += stepx ? 0 : sy;
        lin += stepx ? linx : liny;
        live = tn < t1 && ix != ixe && iy != iye;
        dx += __hiloint2double(stepx ? sxhi : 0, 0);
        dy += __hiloint2double(stepx ? 0 : syhi, 0);
        tx = (dx - qx) * kx;
        ty = (dy - qy) * ky;
    }
};

// The forward's workspace xx: xx[i] = x[i] for i < n^2, then xT[ry * n + ix] = x[ix * n + ry]
// (ry = n - 1 - iy counts down y, as the columns of A do) at xx + n^2, so that one base address
// and a 32-bit element offset reach either copy.
__global__ void __launch_bounds__(256) k_mf_transpose(const float *__restrict__ x, float *__restrict__ xx, int n) {
    __shared__ float t[32][33];
    const int rb = blockIdx.x * 32, ib = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t nn = (int64_t)n * n;
    for (int k = ty; k < 32; k += 8) {
        const int ix = ib + k, ry = rb + tx;
        if (ix < n && ry < n) {
            const float v = x[(int64_t)ix * n + ry];
            t[k][tx] = v;
            xx[(int64_t)ix * n + ry] = v;
        }
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int ry = rb + k, ix = ib + tx;
        if (ix < n && ry < n) xx[nn + (int64_t)ry * n + ix] = t[tx][k];
    }
}

// One thread per ray.  Neighbouring lanes hold neighbouring rays of one angle, which at a given
// step sit one pixel apart across the ray: for rays running mostly along x that is the next
// column index of x, for rays running mostly along y the next row, n entries away.  Those lanes
// gather from the transposed copy instead, so that a gather instruction touches a few cache lines
// rather than one per lane.  The values and the order of the sums are the same either way.
//
// The walk is software-pipelined: the gathers of one batch of MF_BATCH segments are issued, the
// next batch is walked while they are in flight, and only then are they summed.  A gather is a
// 32-bit element offset into the workspace (BUF: through a buffer resource, no 64-bit address
// registers; otherwise added to the base pointer).  With 64-bit addresses the compiler reuses a
// few address registers across a batch's loads, and a VALU write to an address register waits
// until the load before has read it, which serialises the batch (measured: as slow as the walk
// without pipelining).
template <bool BUF>
__global__ void __launch_bounds__(256) k_mf_forward(MfArgs g, int64_t m, const float *__restrict__ xx,
                                                    const float *__restrict__ b, float *__restrict__ y) {
    const int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= m || g.flag[ray]) return;
    const int i = (int)(ray / g.p), j = (int)(ray - (int64_t)i * g.p);
    double x0, y0, a, bb;
    ray_line(g, i, j, x0, y0, a, bb);
    const int n = g.n;
    const bool ymajor = fabs(bb) > fabs(a);
    const uint32_t nn = (uint32_t)n * (uint32_t)n, base = ymajor ? nn : 0u;
    // BUF: the host checked that 2 n^2 floats fit the resource's 32-bit byte range
    const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(xx), (short)0,
                                                                         BUF ? (int)(8u * nn) : 0, 0x00020000);
    Walk w;
    w.init(x0, y0, a, bb, n, 0, n, 0, n);
    // gather offsets, in bytes for BUF (the loads take them as they are), in floats otherwise
    uint32_t off[MF_BATCH];
    float sg[MF_BATCH];
    // every lane walks the first batch: off[] = 0 where the walk is not live
    auto walk_batch = [&]() {
#pragma unroll
        for (int k = 0; k < MF_BATCH; ++k) {
            off[k] = 0;
            sg[k] = 0.f;
            if (w.live) {
                int ix, iy;
                w.next(ix, iy, sg[k]);
                const uint32_t e = base + (uint32_t)(ymajor ? (n - 1 - iy) * n + ix : ix * n + (n - 1 - iy));
                off[k] = BUF ? 4u * e : e;
            }
        }
    };
    walk_batch();
    float acc = 0.f;
    bool more = true;
    while (more) {
        float v[MF_BATCH], sc[MF_BATCH];
#pragma unroll
        for (int k = 0; k < MF_BATCH; ++k) {
            if (BUF) v[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, off[k], 0, 0));
            else v[k] = xx[off[k]];
            sc[k] = sg[k];
        }
        more = w.live;
        if (more) walk_batch();
#pragma unroll
        for (int k = 0; k < MF_BATCH; ++k) acc = fmaf(sc[k], v[k], acc);
    }
    y[ray] = b ? acc - b[ray] : acc;
}

// Segments of positive length per ray, summed (the work count of one product).
__global__ void __launch_bounds__(256) k_mf_count(MfArgs g, int64_t m, unsigned long long *out) {
    const int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c = 0;
    if (ray < m && !g.flag[ray]) {
        const int i = (int)(ray / g.p), j = (int)(ray - (int64_t)i * g.p);
        double x0, y0, a, bb;
        ray_line(g, i, j, x0, y0, a, bb);
        Walk w;
        w.init(x0, y0, a, bb, g.n, 0, g.n, 0, g.n);
        while (w.live) {
            int ix, iy;
            float sg;
            w.next(ix, iy, sg);
            c += sg > 0.f;
        }
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, c);
}

__device__ __forceinline__ int lower_bound_d(const double *v, int lo, int hi, double key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int upper_bound_d(const double *v, int lo, int hi, double key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Fixed-point accumulation of the adjoint (fix_scale / fix_round): LDS float atomics are slow on
// gfx950 (tens of times slower than integer ones), so every term r_i * seg is scaled by a power
// of two, rounded to an integer and summed in 64-bit integers; the rounding error of a term is at
// most 2^-(ecap+1) * max|r|, far below FP32 resolution, and the sums do not depend on order.
__global__ void __launch_bounds__(256) k_mf_absmax(const float *__restrict__ r, int64_t m, unsigned *maxbits) {
    unsigned v = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256)
        v = max(v, __float_as_uint(r[i]) & 0x7fffffffu);   // |r| bits order like the values
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, v);
}

__global__ void k_mf_reset(unsigned *maxbits) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *maxbits = 0u;
}

// x = acc / S (FP32), and acc is cleared for the next call.
__global__ void __launch_bounds__(256) k_mf_fixed_to_f32(unsigned long long *__restrict__ acc, int64_t nn,
                                                         const unsigned *maxbits, int ecap, float *__restrict__ x) {
    const double inv = 1.0 / fix_scale(maxbits, ecap);
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nn; i += (int64_t)gridDim.x * 256) {
        const long long v = (long long)acc[i];
        x[i] = (float)((double)v * inv);
        acc[i] = 0ull;
    }
}

// Workgroup (tile, angle block); blockIdx = angle block * tiles + tile.
__global__ void __launch_bounds__(256) k_mf_adjoint_tiles(MfArgs g, int32_t nty, int32_t ntiles,
                                                          const float *__restrict__ r, const unsigned *maxbits,
                                                          int32_t ecap, unsigned long long *__restrict__ xacc) {
    __shared__ unsigned long long acc[MF_TS * MF_LDS_STRIDE];
    __shared__ int e_lo[256], e_pref[257];
    const int n = g.n, tid = threadIdx.x;
    const int tile = blockIdx.x % ntiles, abk = blockIdx.x / ntiles;
    const int X0 = (tile / nty) * MF_TS, Y0 = (tile % nty) * MF_TS;
    const int X1 = min(X0 + MF_TS, n), Y1 = min(Y0 + MF_TS, n);
    const int i0 = abk * MF_AB, na = min(MF_AB, g.nth - i0);
    const double S = fix_scale(maxbits, ecap);
    for (int q = tid; q < MF_TS * MF_LDS_STRIDE; q += 256) acc[q] = 0ull;
    // candidate rays: entry e = (angle, run); host guarantees MF_AB * nruns <= 256
    const int E = na * g.nruns;
    int cnt = 0;
    if (tid < E) {
        const int i = i0 + tid / g.nruns, run = tid % g.nruns;
        const double sx0 = g.ang[4 * i], sy0 = g.ang[4 * i + 1], ux = g.ang[4 * i + 2], uy = g.ang[4 * i + 3];
        const double h = 0.5 * (double)n;
        double tmin = INFINITY, tmax = -INFINITY;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double cx = (double)((c & 1) ? X1 : X0) - h - sx0;
            const double cy = (double)((c & 2) ? Y1 : Y0) - h - sy0;
            const double tc = (ux * cy - uy * cx) / (ux * cx + uy * cy);
            tmin = fmin(tmin, tc);
            tmax = fmax(tmax, tc);
        }
        tmin -= 1e-9 * (1.0 + fabs(tmin));
        tmax += 1e-9 * (1.0 + fabs(tmax));
        const int rs = g.runs[2 * run], re = g.runs[2 * run + 1];
        const int lo = lower_bound_d(g.tanw, rs, re, tmin);
        const int hi = upper_bound_d(g.tanw, lo, re, tmax);
        e_lo[tid] = lo;
        cnt = hi - lo;
    }
    // exclusive scan of the counts (wavefront scans, then the four wavefront totals)
    {
        int v = cnt;
        const int lane = tid & 63;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(v, off, 64);
            if (lane >= off) v += u;
        }
        __shared__ int wsum[4];
        if (lane == 63) wsum[tid >> 6] = v;
        __syncthreads();
        int base = 0;
        for (int q = 0; q < (tid >> 6); ++q) base += wsum[q];
        e_pref[tid] = base + v - cnt;
        if (tid == 255) e_pref[256] = base + v;
    }
    __syncthreads();
    const int total = e_pref[E];
    for (int wi = tid; wi < total; wi += 256) {
        int lo = 0, hi = E;   // last e with e_pref[e] <= wi
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (e_pref[mid] <= wi) lo = mid;
            else hi = mid;
        }
        const int e = lo;
        const int i = i0 + e / g.nruns, j = e_lo[e] + (wi - e_pref[e]);
        const int64_t ray = (int64_t)i * g.p + j;
        const float rv = r[ray];
        if (rv == 0.f || g.flag[ray]) continue;
        const double rs = (double)rv * S;
        double x0, y0, a, bb;
        ray_line(g, i, j, x0, y0, a, bb);
        Walk w;
        w.init(x0, y0, a, bb, n, X0, X1, Y0, Y1);
        if (w.live) w.track((w.ix - X0) * MF_LDS_STRIDE + (w.iy - Y0), w.sx * MF_LDS_STRIDE, w.sy);
        while (w.live) {
            int q;
            float sg;
            w.next_lin(q, sg);
            atomicAdd(&acc[q], fix_round(rs * (double)sg));
        }
    }
    __syncthreads();
    for (int q = tid; q < MF_TS * MF_TS; q += 256) {
        const int lx = q / MF_TS, ly = q - lx * MF_TS;
        if (X0 + lx < X1 && Y0 + ly < Y1) {
            const unsigned long long v = acc[lx * MF_LDS_STRIDE + ly];
            if (v != 0ull) atomicAdd(&xacc[(int64_t)(X0 + lx) * n + (n - 1 - (Y0 + ly))], v);
        }
    }
}

// Sum-of-products reductions into one FP64 accumulator (block partials, one atomic per wavefront).
__global__ void k_zero_f64(double *a, int n) {
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0.0;
}

__global__ void __launch_bounds__(256) k_dot_f32(const float *a, const float *b, int64_t n, double *out) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        s += (double)a[i] * (double)b[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

// xn = box(x - step * g) with the reference's bound order (min with ub, then max with lb,
// AIRMs.f90:964-973); out += sum (xn - x)^2 in FP64.
__global__ void __launch_bounds__(256) k_mf_gd_step(const float *__restrict__ x, const float *__restrict__ g,
                                                    float *__restrict__ xn, int64_t n, double step, int upb,
                                                    float ub, int lwrb, float lb, double *out) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        float v = (float)((double)x[i] - step * (double)g[i]);
        if (upb) v = fmin_ref(v, ub);
        if (lwrb) v = fmax_ref(v, lb);
        xn[i] = v;
        const double d = (double)v - (double)x[i];
        s += d * d;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

// vn <- vn / nv; out += sum (vn - v)^2 (normest's stopping test, AIRMs.f90:1162-1163).
__global__ void __launch_bounds__(256) k_mf_normalize(float *__restrict__ vn, const float *__restrict__ v, int64_t n,
                                                      const double *nrm2, double *out) {
    const double nv = sqrt(*nrm2);
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float q = (float)((double)vn[i] / nv);
        vn[i] = q;
        const double d = (double)q - (double)v[i];
        s += d * d;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

__global__ void k_neg_f32(const float *b, float *r, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = -b[i];
}

constexpr int RED_GRID = 1024;

// Table rays: forward sums in emission order (FP32), adjoint adds fixed-point terms.
__global__ void __launch_bounds__(64) k_mf_table_forward(const int32_t *rays, const int32_t *offs,
                                                         const int32_t *cols, const float *vals, int32_t nr,
                                                         const float *x, const float *b, float *y) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nr) return;
    float acc = 0.f;
    for (int k = offs[q]; k < offs[q + 1]; ++k) acc = fmaf(vals[k], x[cols[k]], acc);
    const int32_t ray = rays[q];
    y[ray] = b ? acc - b[ray] : acc;
}

__global__ void __launch_bounds__(64) k_mf_table_adjoint(const int32_t *rays, const int32_t *offs,
                                                         const int32_t *cols, const float *vals, int32_t nr,
                                                         const float *r, const unsigned *maxbits, int32_t ecap,
                                                         unsigned long long *xacc) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nr) return;
    const float rv = r[rays[q]];
    if (rv == 0.f) return;
    const double rs = (double)rv * fix_scale(maxbits, ecap);
    for (int k = offs[q]; k < offs[q + 1]; ++k) atomicAdd(&xacc[cols[k]], fix_round(rs * (double)vals[k]));
}

MfArgs args(const fair_mf *h) {
    MfArgs a{h->n, h->p, h->nth, h->nruns, h->ang.p, h->rayc.p, h->tanw.p, h->flag.p, {}};
    for (int q = 0; q < 2 * MF_MAX_RUNS; ++q) a.runs[q] = h->runs[q];
    return a;
}

void mf_forward(fair_mf *h, const float *x, const float *b, float *y) {
    if (h->m > 0) {
        const unsigned nb = (unsigned)((h->n + 31) / 32);
        hipLaunchKernelGGL(k_mf_transpose, dim3(nb, nb), dim3(256), 0, stream(), x, h->xx.p, h->n);
        if (h->mf_gather_buffer)
            hipLaunchKernelGGL(k_mf_forward<true>, dim3(grid_for(h->m, 256)), dim3(256), 0, stream(), args(h), h->m,
                               h->xx.p, b, y);
        else
            hipLaunchKernelGGL(k_mf_forward<false>, dim3(grid_for(h->m, 256)), dim3(256), 0, stream(), args(h), h->m,
                               h->xx.p, b, y);
    }
    if (h->n_exact > 0)
        hipLaunchKernelGGL(k_mf_table_forward, dim3(grid_for(h->n_exact, 64)), dim3(64), 0, stream(), h->ex_rays.p,
                           h->ex_offs.p, h->ex_cols.p, h->ex_vals.p, h->n_exact, x, b, y);
    launch_check("matrix-free forward");
}

void mf_adjoint(const fair_mf *h, const float *r, float *x) {
    const int64_t n = (int64_t)h->n * h->n;
    const int nt = (h->n + MF_TS - 1) / MF_TS, nab = (h->nth + MF_AB - 1) / MF_AB;
    const int64_t blocks = (int64_t)nt * nt * nab;
    FAIR_CHECK(blocks < INT32_MAX, FAIR_ERR_OVERFLOW, "matrix-free adjoint grid too large");
    hipLaunchKernelGGL(k_mf_reset, dim3(1), dim3(64), 0, stream(), h->rmax.p);
    if (h->m > 0) {
        hipLaunchKernelGGL(k_mf_absmax, dim3(1024), dim3(256), 0, stream(), r, h->m, h->rmax.p);
        hipLaunchKernelGGL(k_mf_adjoint_tiles, dim3((unsigned)blocks), dim3(256), 0, stream(), args(h), nt, nt * nt,
                           r, (const unsigned *)h->rmax.p, h->ecap, h->xacc.p);
        if (h->n_exact > 0)
            hipLaunchKernelGGL(k_mf_table_adjoint, dim3(grid_for(h->n_exact, 64)), dim3(64), 0, stream(),
                               h->ex_rays.p, h->ex_offs.p, h->ex_cols.p, h->ex_vals.p, h->n_exact, r,
                               (const unsigned *)h->rmax.p, h->ecap, h->xacc.p);
    }
    hipLaunchKernelGGL(k_mf_fixed_to_f32, dim3(1024), dim3(256), 0, stream(), h->xacc.p, n,
                       (const unsigned *)h->rmax.p, h->ecap, x);
    launch_check("matrix-free adjoint");
}

double read_acc(const DBuf<double> &acc, int k) {
    double h = 0.0;
    FAIR_HIP(hipMemcpyAsync(&h, acc.p + k, sizeof(double), hipMemcpyDeviceToHost, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
    return h;
}

double dev_dot(const float *a, const float *b, int64_t n) {
    DBuf<double> acc(1);
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, stream(), acc.p, 1);
    hipLaunchKernelGGL(k_dot_f32, dim3(RED_GRID), dim3(256), 0, stream(), a, b, n, acc.p);
    launch_check("dot product");
    return read_acc(acc, 0);
}

void check_mf(const fair_mf *h) { FAIR_CHECK(h != nullptr, FAIR_ERR_ARG, "null matrix-free handle"); }

}  // namespace
}  // namespace fair

using namespace fair;

extern "C" int fair_mf_create(const fair_geom *g, const double *theta, fair_mf **out) {
    return guard([&] {
        FAIR_CHECK(g && theta && out, FAIR_ERR_ARG, "null argument");
        FAIR_CHECK(g->n >= 1 && g->p >= 1 && g->nth >= 1, FAIR_ERR_ARG, "geometry needs n, p, nth >= 1");
        FAIR_CHECK((int64_t)g->n * g->n <= INT32_MAX, FAIR_ERR_OVERFLOW, "n**2 exceeds int32");
        const int n = g->n, p = g->p, nth = g->nth;
        // the tables of fanlineartomo (system_generation.f90:98-157)
        const double R = g->r * (double)n, dw = g->dw * (double)n, sd = g->sd * (double)n;
        const double x0 = 0.0, y0 = R, dew = dw / (double)p, pi = 4.0 * atan(1.0);
        std::vector<double> omega(p), rayc(2 * (size_t)p), ang(4 * (size_t)nth);
        if (p % 2 == 1) {
            int hh = (p - 1) / 2;
            for (int j = 1; j <= hh; ++j) omega[j - 1] = atan(sd / ((double)(hh + 1 - j) * dew)) - .5;
            omega[hh] = 0.0;
            for (int k = 1; k <= hh; ++k) omega[hh + k] = .5 * pi - atan(sd / ((double)k * dew));
        } else {
            int hh = p / 2;
            for (int j = 1; j <= hh; ++j) omega[j - 1] = atan(sd / (((double)(hh + 1 - j) - .5) * dew)) - .5 * pi;
            for (int k = 1; k <= hh; ++k) omega[hh + k - 1] = .5 * pi - atan(sd / (((double)k - .5) * dew));
        }
        for (int j = 0; j < p; ++j) sincos(omega[j], &rayc[2 * j + 1], &rayc[2 * j]);
        for (int i = 0; i < nth; ++i) {
            double st, ct;
            sincos(theta[i], &st, &ct);
            const double x0th = ct * x0 - st * y0, y0th = st * x0 + ct * y0;
            ang[4 * i] = x0th;
            ang[4 * i + 1] = y0th;
            ang[4 * i + 2] = -(x0th / R);
            ang[4 * i + 3] = -(y0th / R);
        }
        // tan(omega) in increasing runs (one run for even p; the odd-p left half, whose offset is
        // .5 instead of pi/2 (:124), forms a second run)
        std::vector<double> tanw(p);
        std::vector<int32_t> runs{0};
        for (int j = 0; j < p; ++j) {
            tanw[j] = rayc[2 * j + 1] / rayc[2 * j];
            if (j > 0 && tanw[j] < tanw[j - 1]) {
                runs.push_back(j);
                runs.push_back(j);
            }
        }
        runs.push_back(p);
        const int nruns = (int)runs.size() / 2;
        FAIR_CHECK(nruns <= MF_MAX_RUNS && MF_AB * nruns <= 256, FAIR_ERR_ARG,
                   "ray directions form too many monotone runs for the matrix-free projector");
        for (int j = 0; j < p; ++j)
            FAIR_CHECK(rayc[2 * j] > 0.0, FAIR_ERR_ARG, "matrix-free projector needs |omega| < pi/2");
        // the tile selection measures ray angles at the source: it must lie outside the image
        FAIR_CHECK(R > 0.5 * (double)n * std::sqrt(2.0) + 1.0, FAIR_ERR_ARG,
                   "matrix-free projector needs the source outside the image (r*n > n/sqrt(2) + 1)");
        std::unique_ptr<fair_mf> h(new fair_mf());
        h->n = n;
        h->p = p;
        h->nth = nth;
        h->m = (int64_t)nth * p;
        h->nruns = nruns;
        for (int q = 0; q < 2 * nruns; ++q) h->runs[q] = runs[q];
        h->ang.alloc(ang.size());
        h->rayc.alloc(rayc.size());
        h->tanw.alloc(tanw.size());
        h->ang.upload(ang.data(), ang.size());
        h->rayc.upload(rayc.data(), rayc.size());
        h->tanw.upload(tanw.data(), tanw.size());
        // fixed-point range: one term below 2^49.5, all p*nth*sqrt(2)*M*S below 2^62
        const int ebits = (int)std::ceil(std::log2(1.5 * (double)p * (double)nth));
        h->ecap = std::min(49, 62 - ebits);
        FAIR_CHECK(h->ecap >= 24, FAIR_ERR_OVERFLOW, "too many rays for the fixed-point adjoint");
        const int64_t nn = (int64_t)n * n;
        h->xacc.alloc(nn);
        h->xx.alloc(2 * nn);
        // a buffer resource spans at most 4 GiB - 1 bytes; FAIR_MF_GATHER=pointer selects the
        // pointer path for any size (the results are the same)
        const char *gather = getenv("FAIR_MF_GATHER");
        h->mf_gather_buffer = 8 * nn < ((int64_t)1 << 32) && !(gather && strcmp(gather, "pointer") == 0);
        h->rmax.alloc(1);
        FAIR_HIP(hipMemsetAsync(h->xacc.p, 0, sizeof(unsigned long long) * nn, stream()));
        FAIR_HIP(hipMemsetAsync(h->rmax.p, 0, sizeof(unsigned), stream()));
        FAIR_HIP(hipStreamSynchronize(stream()));
        std::vector<int32_t> ex, eoff, ecol;
        std::vector<float> eval;
        mf_exact_rays(n, p, nth, h->ang.p, h->rayc.p, ex, eoff, ecol, eval);
        std::vector<uint8_t> flag((size_t)h->m, 0);
        for (int32_t r : ex) flag[r] = 1;
        h->n_exact = (int32_t)ex.size();
        for (float v : eval) h->ex_segments += v > 0.f;
        h->ex_rays.alloc(std::max<size_t>(ex.size(), 1));
        h->ex_offs.alloc(eoff.size());
        h->ex_cols.alloc(std::max<size_t>(ecol.size(), 1));
        h->ex_vals.alloc(std::max<size_t>(eval.size(), 1));
        h->ex_rays.upload(ex.data(), ex.size());
        h->ex_offs.upload(eoff.data(), eoff.size());
        h->ex_cols.upload(ecol.data(), ecol.size());
        h->ex_vals.upload(eval.data(), eval.size());
        h->flag.alloc((size_t)h->m);
        h->flag.upload(flag.data(), flag.size());
        FAIR_HIP(hipStreamSynchronize(stream()));
        *out = h.release();
    });
}

extern "C" int fair_mf_free(fair_mf *h) {
    return guard([&] {
        (void)hipStreamSynchronize(stream());
        delete h;
    });
}

extern "C" int fair_mf_forward_dev(fair_mf *h, const float *d_x, float *d_y) {
    return guard([&] {
        check_mf(h);
        mf_forward(h, d_x, nullptr, d_y);
    });
}

extern "C" int fair_mf_adjoint_dev(fair_mf *h, const float *d_y, float *d_x) {
    return guard([&] {
        check_mf(h);
        mf_adjoint(h, d_y, d_x);
    });
}

extern "C" int fair_mf_forward(fair_mf *h, const float *x, float *y) {
    return guard([&] {
        check_mf(h);
        const int64_t n = (int64_t)h->n * h->n;
        DBuf<float> dx(n), dy(h->m > 0 ? h->m : 1);
        dx.upload(x, n);
        mf_forward(h, dx.p, nullptr, dy.p);
        dy.download(y, h->m);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_mf_adjoint(fair_mf *h, const float *y, float *x) {
    return guard([&] {
        check_mf(h);
        const int64_t n = (int64_t)h->n * h->n;
        DBuf<float> dx(n), dy(h->m > 0 ? h->m : 1);
        dy.upload(y, h->m);
        mf_adjoint(h, dy.p, dx.p);
        dx.download(x, n);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_mf_segments(fair_mf *h, int64_t *count) {
    return guard([&] {
        check_mf(h);
        FAIR_CHECK(count, FAIR_ERR_ARG, "null argument");
        DBuf<unsigned long long> c(1);
        FAIR_HIP(hipMemsetAsync(c.p, 0, sizeof(unsigned long long), stream()));
        if (h->m > 0)
            hipLaunchKernelGGL(k_mf_count, dim3(grid_for(h->m, 256)), dim3(256), 0, stream(), args(h), h->m, c.p);
        launch_check("matrix-free segment count");
        unsigned long long v = 0;
        c.download(&v, 1);
        FAIR_HIP(hipStreamSynchronize(stream()));
        *count = (int64_t)v + h->ex_segments;
    });
}

extern "C" int fair_mf_residual_dev(fair_mf *h, const float *d_x, const float *d_b, float *d_r) {
    return guard([&] {
        check_mf(h);
        FAIR_CHECK(d_b, FAIR_ERR_ARG, "null right-hand side");
        mf_forward(h, d_x, d_b, d_r);
    });
}

// normest (AIRMs.f90:1129-1173) on the matrix-free operator: FP32 vectors, FP64 norms.
// v0 (host, n entries) replaces the reference's RANDOM_NUMBER start vector.
extern "C" int fair_mf_normest(fair_mf *h, const float *v0, double tol, int32_t maxit, double *sigma) {
    return guard([&] {
        check_mf(h);
        FAIR_CHECK(v0 && sigma, FAIR_ERR_ARG, "null argument");
        const int64_t n = (int64_t)h->n * h->n;
        DBuf<float> v(n), vn(n), w(h->m > 0 ? h->m : 1);
        DBuf<double> acc(2);
        v.upload(v0, n);
        for (int i = 1; i <= maxit; ++i) {
            mf_forward(h, v.p, nullptr, w.p);                  // w = A v
            mf_adjoint(h, w.p, vn.p);                          // vn = A^T w
            hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, stream(), acc.p, 2);
            hipLaunchKernelGGL(k_dot_f32, dim3(RED_GRID), dim3(256), 0, stream(), vn.p, vn.p, n, acc.p);
            hipLaunchKernelGGL(k_mf_normalize, dim3(RED_GRID), dim3(256), 0, stream(), vn.p, v.p, n, acc.p,
                               acc.p + 1);
            launch_check("matrix-free normest");
            if (std::sqrt(read_acc(acc, 1)) < tol) break;     // v keeps the previous vector
            std::swap(v.p, vn.p);
        }
        mf_forward(h, v.p, nullptr, w.p);
        const double ww = dev_dot(w.p, w.p, h->m), vv = dev_dot(v.p, v.p, n);
        *sigma = std::sqrt(ww / vv);
    });
}

// gradient_descent (AIRMs.f90:860-995) on 1/2 ||Ax - b||^2 with the matrix-free operator, FP32
// vectors and FP64 stopping norm.  The reference's control flow is kept: r starts as -b
// (whatever x0 is), xn = box(x - rho/sigma^2 A^T r), the loop exits BEFORE accepting xn when
// ||xn - x|| < tol, otherwise x = xn and r = A x - b.  d_x holds x0 on entry and the result on
// exit; *iters receives the number of accepted updates.  tol = 0 runs all maxiter iterations
// without reading the norm back.
extern "C" int fair_mf_gd_dev(fair_mf *h, const float *d_b, int32_t maxiter, double rho, double tol, double sigma,
                              const fair_opts *o, float *d_x, int32_t *iters) {
    return guard([&] {
        check_mf(h);
        FAIR_CHECK(d_b && d_x, FAIR_ERR_ARG, "null argument");
        FAIR_CHECK(tol >= 0.0, FAIR_ERR_ARG, "Must use positive tolerance.");
        FAIR_CHECK(sigma > 0.0, FAIR_ERR_ARG, "sigma must be positive");
        const int64_t n = (int64_t)h->n * h->n;
        DBuf<float> r(h->m > 0 ? h->m : 1), g(n), xw(n);
        DBuf<double> acc(1);
        const double step = rho / (sigma * sigma);
        const int upb = o ? o->upb : 0, lwrb = o ? o->lwrb : 0;
        const float ub = o ? (float)o->ubound : 0.f, lb = o ? (float)o->lbound : 1.f;
        float *x = d_x, *xn = xw.p;
        if (h->m > 0)
            hipLaunchKernelGGL(k_neg_f32, dim3(grid_for(h->m, 256)), dim3(256), 0, stream(), d_b, r.p, h->m);
        int accepted = 0;
        for (int it = 1; it <= maxiter; ++it) {
            mf_adjoint(h, r.p, g.p);
            hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, stream(), acc.p, 1);
            hipLaunchKernelGGL(k_mf_gd_step, dim3(RED_GRID), dim3(256), 0, stream(), x, g.p, xn, n, step, upb, ub,
                               lwrb, lb, acc.p);
            launch_check("matrix-free gradient step");
            if (tol > 0.0 && std::sqrt(read_acc(acc, 0)) < tol) break;
            std::swap(x, xn);
            ++accepted;
            if (it < maxiter) mf_forward(h, x, d_b, r.p);
        }
        if (x != d_x) FAIR_HIP(hipMemcpyAsync(d_x, x, sizeof(float) * n, hipMemcpyDeviceToDevice, stream()));
        FAIR_HIP(hipStreamSynchronize(stream()));
        if (iters) *iters = accepted;
    });
}

// Host-array form of fair_mf_gd_dev: b (m), x0 (n*n, may be NULL: zero start), xout (n*n).
extern "C" int fair_mf_gd(fair_mf *h, const float *b, int32_t maxiter, const float *x0, double rho, double tol,
                          double sigma, const fair_opts *o, float *xout, int32_t *iters) {
    return guard([&] {
        check_mf(h);
        FAIR_CHECK(b && xout, FAIR_ERR_ARG, "null argument");
        const int64_t n = (int64_t)h->n * h->n;
        DBuf<float> db(h->m > 0 ? h->m : 1), dx(n);
        db.upload(b, h->m);
        if (x0) dx.upload(x0, n);
        else FAIR_HIP(hipMemsetAsync(dx.p, 0, sizeof(float) * n, stream()));
        const int st = fair_mf_gd_dev(h, db.p, maxiter, rho, tol, sigma, o, dx.p, iters);
        if (st != FAIR_OK) throw Error{st, fair_last_error()};
        dx.download(xout, n);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}
