// raytrace.hip -- fan / parallel beam system-matrix build on the device.
//
// Replaces fanlineartomo (system_generation.f90:79-307) and paralleltomo (:309-492).  One
// thread per ray.  A ray's crossings with the n+1 lines x = const and the n+1 lines y = const
// form two monotone sequences in the ray parameter t, so the sorted intersection list that the
// reference builds with a quicksort (:172-180) is produced here by MERGING the two sequences
// on the fly (no per-ray arrays, no sort).  For distinct t values the merge yields exactly the
// quicksort's order.  Rays whose order the merge cannot vouch for -- an exact tie between an
// x-crossing and a y-crossing, or an axis-aligned direction where t is +-Inf/NaN and the
// reference's unstable sort decides the order -- are marked irregular and traced by a second
// kernel that runs the reference's Hoare quicksort step for step in per-ray global scratch.
//
// Pipeline: count kernel (per-ray entry counts, irregular list) -> exact kernel (counts of the
// irregular rays) -> exclusive scan (IA) -> fill kernel -> exact kernel in fill mode.
//
// Bit-exactness: host-side trig (cos, sin, atan from the C library, as the reference calls
// them) is uploaded as tables; the device uses only IEEE +, -, *, /, sqrt and floor, and this
// file is compiled with -ffp-contract=off so no product is fused into an add.
#include <algorithm>
#include <cmath>

#include "fair_internal.hpp"

namespace fair {

namespace {

constexpr int BEAM_FAN = 0, BEAM_PAR = 1;

struct RayArgs {
    int32_t n, p, nth;
    const double *ang;    // fan: per angle {x0th, y0th, ath, bth}; parallel: {cos, sin}
    const double *rayc;   // fan: per ray {cos(omega), sin(omega)}; parallel: {x0, y0}
};

struct RayGeom {
    double a, b, x0th, y0th;
    bool excl;
};

template <int BEAM>
__device__ __forceinline__ RayGeom ray_geom(const RayArgs &g, int64_t ray) {
    int i = (int)(ray / g.p), j = (int)(ray - (int64_t)i * g.p);
    RayGeom r;
    if (BEAM == BEAM_FAN) {
        // system_generation.f90:146-157
        double x0th = g.ang[4 * i], y0th = g.ang[4 * i + 1];
        double ath = g.ang[4 * i + 2], bth = g.ang[4 * i + 3];
        double co = g.rayc[2 * j], so = g.rayc[2 * j + 1];
        r.a = co * ath - so * bth;
        r.b = so * ath + co * bth;
        r.x0th = x0th;
        r.y0th = y0th;
        r.excl = false;
    } else {
        // system_generation.f90:348-353, exclusion :394 / :466-467
        double ct = g.ang[2 * i], st = g.ang[2 * i + 1];
        double x0 = g.rayc[2 * j], y0 = g.rayc[2 * j + 1];
        r.x0th = ct * x0 - st * y0;
        r.y0th = st * x0 + ct * y0;
        r.a = -st;
        r.b = ct;
        double half = .5 * (double)g.n;
        r.excl = (r.b == 0.0 && fabs(r.y0th - half) < 1e-15) ||
                 (r.a == 0.0 && fabs(r.x0th - half) < 1e-15);
    }
    return r;
}

// Grid line coordinate k (one-based, 1..n+1): k - n/2 - 1, system_generation.f90:108-111
// (the parallel-beam linspace of :331 gives the same exact values).
__device__ __forceinline__ double gridc(int k, double half) { return (double)k - half - 1.0; }

// Output of the fill pass: entry cnt of the ray gets the one-based column col and length tv.
struct Emit {
    int32_t *JA;
    double *vA;
    int cap;   // entries reserved for this ray by the count pass

    __device__ __forceinline__ void emit(int cnt, int col, double tv) {
        if (cnt < cap) {
            JA[cnt] = col - 1;
            vA[cnt] = tv;
        }
    }
};

// Filter + de-duplication + emission for one point in sorted order (:183-203 / :264-297).
struct PointSink {
    double lo, hi, nd;
    int n;
    bool have_prev = false;
    double px = 0, py = 0;
    int cnt = 0;
    bool bad_col = false;

    template <bool FILL, class E>
    __device__ __forceinline__ void push(double x, double y, bool excl, E &e) {
        if (!(x >= lo && x <= hi && y >= lo && y <= hi)) return;
        if (have_prev) {
            double dx = px - x, dy = py - y;
            double tv = sqrt(dx * dx + dy * dy);
            if (tv > 1e-10 && !excl) {
                if (FILL) {
                    double xm = .5 * ((px + x) + nd);
                    double ym = .5 * ((py + y) + nd);
                    // one-based; the reference stops only on 0 (:293).  Rays that run along the
                    // +n/2 grid line without being caught by the exclusion of :394 (direction
                    // not exactly axis-aligned) yield columns above n**2 in the reference too;
                    // they are kept bit-exact and skipped by every product kernel.
                    int col = (int)floor(xm) * n + (n - (int)floor(ym));
                    if (col == 0) bad_col = true;
                    e.emit(cnt, col, tv);
                }
                cnt++;
            }
        }
        have_prev = true;
        px = x;
        py = y;
    }
};

// Regular rays: merge of the x-crossing and y-crossing sequences.
// Returns -1 when the ray must take the exact path.
template <int BEAM, bool FILL, class E>
__device__ int merge_ray(const RayArgs &g, int64_t ray, E &e, bool *bad_col) {
    RayGeom r = ray_geom<BEAM>(g, ray);
    if (!(r.a != 0.0 && r.b != 0.0)) return -1;   // axis-aligned (or NaN) direction
    const int n = g.n;
    const double half = .5 * (double)n;
    PointSink s;
    s.lo = -half;
    s.hi = half;
    s.nd = (double)n;
    s.n = n;
    // one-based line indices walked in increasing t
    const int kx0 = r.a > 0 ? 1 : n + 1, dkx = r.a > 0 ? 1 : -1;
    const int ky0 = r.b > 0 ? 1 : n + 1, dky = r.b > 0 ? 1 : -1;
    int cx = 0, cy = 0;
    double tx = (gridc(kx0, half) - r.x0th) / r.a;
    double ty = (gridc(ky0, half) - r.y0th) / r.b;
    for (int step = 0; step < 2 * n + 2; ++step) {
        bool takex;
        if (cx > n) takex = false;
        else if (cy > n) takex = true;
        else {
            if (tx == ty) return -1;                  // tie: quicksort order is not defined
            takex = tx < ty;
        }
        if (takex) {
            double xg = gridc(kx0 + dkx * cx, half);
            double yx = r.b * tx + r.y0th;            // :164
            s.template push<FILL, E>(xg, yx, r.excl, e);
            ++cx;
            if (cx <= n) tx = (gridc(kx0 + dkx * cx, half) - r.x0th) / r.a;
        } else {
            double yg = gridc(ky0 + dky * cy, half);
            double xy = r.a * ty + r.x0th;            // :169
            s.template push<FILL, E>(xy, yg, r.excl, e);
            ++cy;
            if (cy <= n) ty = (gridc(ky0 + dky * cy, half) - r.y0th) / r.b;
        }
    }
    if (s.bad_col) *bad_col = true;
    return s.cnt;
}

template <int BEAM>
__global__ void __launch_bounds__(256) k_count(RayArgs g, int64_t m, int32_t *counts,
                                               uint8_t *irregular, int32_t *irr_list,
                                               int32_t *irr_n) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= m) return;
    bool bad = false;
    Emit none{nullptr, nullptr, 0};
    int c = merge_ray<BEAM, false>(g, ray, none, &bad);
    if (c < 0) {
        irregular[ray] = 1;
        counts[ray] = 0;
        int slot = atomicAdd(irr_n, 1);
        irr_list[slot] = (int32_t)ray;
    } else {
        irregular[ray] = 0;
        counts[ray] = c;
    }
}

template <int BEAM>
__global__ void __launch_bounds__(256) k_fill(RayArgs g, int64_t m, const int32_t *IA,
                                              const uint8_t *irregular, int32_t *JA, double *vA,
                                              int32_t *err) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= m || irregular[ray]) return;
    int32_t off = IA[ray];
    bool bad = false;
    Emit e{JA + off, vA + off, IA[ray + 1] - off};
    int c = merge_ray<BEAM, true>(g, ray, e, &bad);
    if (bad) atomicOr(err, 1);
    if (c != IA[ray + 1] - off) atomicOr(err, 2);
}

// The reference's algorithm for one ray; base: 36 * (2n+2) bytes of scratch.  Returns the
// number of entries (columns of any value) the ray produces.
template <int BEAM, bool FILL, class E>
__device__ int exact_ray(const RayArgs &g, int64_t ray, char *base, E &e, bool *bad_col) {
    const int n = g.n, nn = 2 * n + 2;
    double *t = (double *)base - 1;
    double *xx = (double *)base + nn - 1;
    double *yy = (double *)base + 2 * nn - 1;
    int32_t *ind = (int32_t *)((double *)base + 3 * nn) - 1;
    int32_t *stk = (int32_t *)((double *)base + 3 * nn) + nn;
    RayGeom r = ray_geom<BEAM>(g, ray);
    const double half = .5 * (double)n;
    for (int k = 1; k <= n + 1; ++k) {
        double xg = gridc(k, half);
        double tx = (xg - r.x0th) / r.a;
        double yx = r.b * tx + r.y0th;
        double ty = (xg - r.y0th) / r.b;
        double xy = r.a * ty + r.x0th;
        t[k] = tx;
        t[n + 1 + k] = ty;
        xx[k] = xg;
        xx[n + 1 + k] = xy;
        yy[k] = yx;
        yy[n + 1 + k] = xg;
    }
    for (int k = 1; k <= nn; ++k) ind[k] = k;
    // quicksort(t, ind) of system_generation.f90:37-77 with an explicit range stack; the
    // sub-ranges are disjoint, so the processing order does not change the result.
    int sp = 0;
    stk[sp++] = 1;
    stk[sp++] = nn;
    while (sp > 0) {
        int l = stk[--sp], f = stk[--sp];
        double x = t[(f + l) / 2];
        int i = f, j = l;
        for (;;) {
            while (i < l && t[i] < x) ++i;
            while (j > f && x < t[j]) --j;
            if (i >= j) break;
            double tt = t[i]; t[i] = t[j]; t[j] = tt;
            int32_t mm = ind[i]; ind[i] = ind[j]; ind[j] = mm;
            ++i;
            --j;
        }
        if (j + 1 < l) { stk[sp++] = j + 1; stk[sp++] = l; }
        if (f < i - 1) { stk[sp++] = f; stk[sp++] = i - 1; }
    }
    PointSink s;
    s.lo = -half;
    s.hi = half;
    s.nd = (double)n;
    s.n = n;
    for (int k = 1; k <= nn; ++k) s.template push<FILL, E>(xx[ind[k]], yy[ind[k]], r.excl, e);
    if (s.bad_col) *bad_col = true;
    return s.cnt;
}

// The same algorithm with its sort in LDS, one 64-lane wavefront per ray.  The lanes compute the
// 2n+2 crossing parameters; the reference's quicksort of (t, ind) (system_generation.f90:37-77:
// Hoare partition around the middle element, both sub-ranges sorted in turn) then runs with each
// partition spread over the wavefront.  Its inputs (the x-line and the y-line parameters, two
// monotone runs) split very unevenly, so the partitions, not the ranges, carry the work.
//
// A Hoare partition of [f, l] around x swaps, while they have not crossed, the k-th "left stopper"
// (a position the left scan stops at: !(t < x), or l) with the k-th "right stopper" from the right
// (!(x < t), or f).  Every position a scan reaches before the two meet still holds its original
// value, and a swapped position holds a stopper for the scan that meets it next, so: with L and R
// the stopper positions of the original range (L increasing, R decreasing), K the first k with
// L[k] >= R[k], the partition swaps (L[k], R[k]) for k < K and ends with i = min(L[K], R[K-1]),
// j = max(R[K], L[K-1]) (i = L[0], j = R[0] when K = 0), exactly as the sequential loop does, NaN
// included (a NaN compares false, so it stops both scans).  Here L and R come from wave prefix
// sums, K from a wave vote, and the K swaps are independent.  The pending ranges are disjoint, so
// the order they are taken in does not change the result.
//
// The point rule follows in sorted order (system_generation.f90:162-203): each point is
// recomputed from its index as the reference's arrays hold it, the points inside the box are
// compacted, and each lane takes one pair of consecutive kept points, emitting the pair's segment
// at its prefix position among the accepted pairs.  LDS per crossing: t (8 B), the range stack
// (8 B), ind (4 B), the stopper lists (2 x 4 B); after the sort the kept points overwrite t and
// the stack.
constexpr int EXACT_LDS_PER_CROSSING = 28;

__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __popcll(mask & ((1ull << (threadIdx.x & 63)) - 1ull));
}

template <int BEAM, bool FILL, class E>
__device__ int exact_ray_lds(const RayArgs &g, int64_t ray, unsigned char *lds, E &e, bool *bad_col) {
    const int n = g.n, nn = 2 * n + 2;
    const int lane = (int)threadIdx.x;
    double *t = (double *)lds - 1;                              // t[1..nn]
    int32_t *stk = (int32_t *)(lds + 8 * (size_t)nn);          // pending ranges, 2 entries each
    int32_t *ind = (int32_t *)(lds + 16 * (size_t)nn) - 1;     // ind[1..nn]
    int32_t *Lp = (int32_t *)(lds + 20 * (size_t)nn);          // left stoppers, increasing
    int32_t *Rp = (int32_t *)(lds + 24 * (size_t)nn);          // right stoppers, increasing
    const RayGeom r = ray_geom<BEAM>(g, ray);
    const double half = .5 * (double)n;
    for (int k = 1 + lane; k <= n + 1; k += WAVE) {
        const double xg = gridc(k, half);
        t[k] = (xg - r.x0th) / r.a;
        t[n + 1 + k] = (xg - r.y0th) / r.b;
        ind[k] = k;
        ind[n + 1 + k] = n + 1 + k;
    }
    if (lane == 0) {
        stk[0] = 1;
        stk[1] = nn;
    }
    __syncthreads();
    // pending ranges are disjoint with at least two entries each: at most nn / 2 of them
    int sp = 1;
    while (sp > 0) {
        --sp;
        const int f = stk[2 * sp], l = stk[2 * sp + 1];
        const double x = t[(f + l) / 2];
        int nL = 0, nR = 0;
        for (int p0 = f; p0 <= l; p0 += WAVE) {
            const int p = p0 + lane;
            const bool in = p <= l;
            const double v = in ? t[p] : 0.0;
            const bool ls = in && (!(v < x) || p == l);
            const bool rs = in && (!(x < v) || p == f);
            const uint64_t bl = __ballot(ls), br = __ballot(rs);
            if (ls) Lp[nL + lanes_below(bl)] = p;
            if (rs) Rp[nR + lanes_below(br)] = p;
            nL += __popcll(bl);
            nR += __popcll(br);
        }
        __syncthreads();
        // L holds l and R holds f, so some k < min(nL, nR) has L[k] >= R[k] (R[k] = Rp[nR-1-k])
        const int kmax = min(nL, nR);
        int K = kmax - 1;   // (the bound; the vote finds it at the latest there)
        for (int k0 = 0; k0 < kmax; k0 += WAVE) {
            const int k = k0 + lane;
            const bool c = k < kmax && Lp[k] >= Rp[nR - 1 - k];
            const uint64_t b = __ballot(c);
            if (b) {
                K = k0 + __ffsll((long long)b) - 1;
                break;
            }
        }
        for (int k = lane; k < K; k += WAVE) {
            const int a = Lp[k], c = Rp[nR - 1 - k];
            const double tt = t[a];
            t[a] = t[c];
            t[c] = tt;
            const int32_t mm = ind[a];
            ind[a] = ind[c];
            ind[c] = mm;
        }
        const int i = K > 0 ? min(Lp[K], Rp[nR - K]) : Lp[0];
        const int j = K > 0 ? max(Rp[nR - 1 - K], Lp[K - 1]) : Rp[nR - 1];
        __syncthreads();   // swaps and stopper reads done before the next partition
        if (j + 1 < l) {
            if (lane == 0) {
                stk[2 * sp] = j + 1;
                stk[2 * sp + 1] = l;
            }
            ++sp;
        }
        if (f < i - 1) {
            if (lane == 0) {
                stk[2 * sp] = f;
                stk[2 * sp + 1] = i - 1;
            }
            ++sp;
        }
        __syncthreads();
    }
    // the points in sorted order; those inside the box (:187-193), compacted, over t and the stack
    double *kx = (double *)lds, *ky = (double *)(lds + 8 * (size_t)nn);
    const double lo = -half, hi = half;
    int nk = 0;
    for (int k0 = 1; k0 <= nn; k0 += WAVE) {
        const int k = k0 + lane;
        double px = 0.0, py = 0.0;
        bool keep = false;
        if (k <= nn) {
            const int q = ind[k];
            if (q <= n + 1) {                       // x-line crossing (:162-164)
                const double xg = gridc(q, half);
                px = xg;
                py = r.b * ((xg - r.x0th) / r.a) + r.y0th;
            } else {                                // y-line crossing (:167-169)
                const double yg = gridc(q - (n + 1), half);
                px = r.a * ((yg - r.y0th) / r.b) + r.x0th;
                py = yg;
            }
            keep = px >= lo && px <= hi && py >= lo && py <= hi;
        }
        const uint64_t b = __ballot(keep);
        __syncthreads();   // ind of this chunk read before kx / ky (over t, the stack) are written
        if (keep) {
            const int o = nk + lanes_below(b);
            kx[o] = px;
            ky[o] = py;
        }
        nk += __popcll(b);
    }
    __syncthreads();
    // consecutive kept points (previous, current): the segment test and the midpoint column rule
    // (:196-203, :279-297)
    const double nd = (double)n;
    int cnt = 0;
    bool bad = false;
    for (int q0 = 1; q0 < nk; q0 += WAVE) {
        const int q = q0 + lane;
        bool val = false;
        double tv = 0.0;
        int col = 0;
        if (q < nk) {
            const double ax = kx[q - 1], ay = ky[q - 1], bx = kx[q], by = ky[q];
            const double dx = ax - bx, dy = ay - by;
            tv = sqrt(dx * dx + dy * dy);
            val = tv > 1e-10 && !r.excl;
            if (FILL && val) {
                const double xm = .5 * ((ax + bx) + nd);
                const double ym = .5 * ((ay + by) + nd);
                col = (int)floor(xm) * n + (n - (int)floor(ym));
                if (col == 0) bad = true;
            }
        }
        const uint64_t b = __ballot(val);
        if (FILL && val) e.emit(cnt + lanes_below(b), col, tv);
        cnt += __popcll(b);
    }
    if (__ballot(bad)) *bad_col = true;
    return cnt;
}

template <int BEAM, bool FILL>
__global__ void __launch_bounds__(64) k_exact_lds(RayArgs g, const int32_t *irr_list, int32_t *counts,
                                                  const int32_t *IA, int32_t *JA, double *vA, int32_t *err) {
    extern __shared__ __align__(16) unsigned char exact_lds[];
    const int64_t ray = irr_list[blockIdx.x];
    Emit e{nullptr, nullptr, 0};
    if (FILL) {
        const int32_t off = IA[ray];
        e = Emit{JA + off, vA + off, IA[ray + 1] - off};
    }
    bool bad = false;
    const int c = exact_ray_lds<BEAM, FILL, Emit>(g, ray, exact_lds, e, &bad);
    if (threadIdx.x != 0) return;
    if (FILL) {
        if (bad) atomicOr(err, 1);
        if (c != IA[ray + 1] - IA[ray]) atomicOr(err, 2);
    } else {
        counts[ray] = c;
    }
}

// Irregular rays: the reference algorithm verbatim (t arrays, Hoare quicksort with middle
// pivot, permutation, filter, de-duplication), one thread per ray, arrays in global scratch.
template <int BEAM, bool FILL>
__global__ void k_exact(RayArgs g, const int32_t *irr_list, int32_t nirr, char *scratch,
                        size_t per_ray, int32_t *counts, const int32_t *IA, int32_t *JA,
                        double *vA, int32_t *err) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nirr) return;
    const int64_t ray = irr_list[q];
    Emit e{nullptr, nullptr, 0};
    if (FILL) {
        int32_t off = IA[ray];
        e = Emit{JA + off, vA + off, IA[ray + 1] - off};
    }
    bool bad = false;
    const int c = exact_ray<BEAM, FILL, Emit>(g, ray, scratch + per_ray * (size_t)q, e, &bad);
    if (FILL) {
        if (bad) atomicOr(err, 1);
        if (c != IA[ray + 1] - IA[ray]) atomicOr(err, 2);
    } else {
        counts[ray] = c;
    }
}

// ---- matrix-free projector: rays that need the reference's own point rule --------------------
// A ray within ~1e-13 of a grid line (direction nearly, but not exactly, axis-aligned, such as
// the central ray at theta = pi) has segment midpoints that the reference's floor(xm) rule
// (system_generation.f90:285-293) assigns by rounding, and crossings that tie exactly; an exact
// traversal reproduces neither.  The matrix-free projector keeps the entries of such rays, as
// the matrix build computes them, in a small table.
constexpr double NEAR_AXIS = 1e-10;

// In-range entries of one ray: 0-based column and FP32 length.
struct MfEntries {
    int32_t *cols;
    float *vals;
    int32_t nn, k;
    __device__ __forceinline__ void emit(int, int col, double tv) {
        const int c = col - 1;
        if ((uint32_t)c < (uint32_t)nn) {
            if (cols) {
                cols[k] = c;
                vals[k] = (float)tv;
            }
            ++k;
        }
    }
};

__global__ void __launch_bounds__(256) k_mf_near_axis(RayArgs g, int64_t m, int32_t *list, int32_t *cnt,
                                                      int32_t cap) {
    const int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= m) return;
    const RayGeom r = ray_geom<BEAM_FAN>(g, ray);
    const double fa = fabs(r.a), fb = fabs(r.b);
    if (!(r.a != 0.0 && r.b != 0.0 && (fa < NEAR_AXIS * fb || fb < NEAR_AXIS * fa))) return;
    const int q = atomicAdd(cnt, 1);
    if (q < cap) list[q] = (int32_t)ray;
}

// FILL = false: per listed ray the number of in-range entries; true: write them at offs[q].
template <bool FILL>
__global__ void k_mf_exact_entries(RayArgs g, const int32_t *list, int32_t nlist, char *scratch, size_t per_ray,
                                   int32_t *counts, const int32_t *offs, int32_t *cols, float *vals) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nlist) return;
    MfEntries e{FILL ? cols + offs[q] : nullptr, FILL ? vals + offs[q] : nullptr, g.n * g.n, 0};
    bool bad = false;
    exact_ray<BEAM_FAN, true, MfEntries>(g, list[q], scratch + per_ray * (size_t)q, e, &bad);
    if (!FILL) counts[q] = e.k;
}

// ---- exclusive scan of int32 counts into int32 offsets (with int64 overflow check) ---------
constexpr int SCAN_BLOCK = 1024;

__global__ void __launch_bounds__(256) k_block_sums(const int32_t *c, int64_t m, int64_t *part) {
    __shared__ int64_t sh[256];
    int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK;
    int64_t s = 0;
    for (int k = threadIdx.x; k < SCAN_BLOCK; k += 256) {
        int64_t idx = base + k;
        if (idx < m) s += c[idx];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Single-block exclusive scan of the block sums; part[nb] receives the total.
__global__ void __launch_bounds__(1024) k_scan_parts(int64_t *part, int64_t nb) {
    __shared__ int64_t sh[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        int64_t idx = base + threadIdx.x;
        int64_t v = idx < nb ? part[idx] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int64_t add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (idx < nb) part[idx] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[nb] = carry;
}

__global__ void __launch_bounds__(256) k_scan_apply(const int32_t *c, int64_t m,
                                                    const int64_t *part, int32_t *out) {
    __shared__ int64_t sh[SCAN_BLOCK];
    int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK;
    for (int k = threadIdx.x; k < SCAN_BLOCK; k += 256) {
        int64_t idx = base + k;
        sh[k] = idx < m ? c[idx] : 0;
    }
    __syncthreads();
    // inclusive scan over 1024 entries with 256 threads: serial per thread chunk of 4
    __shared__ int64_t tsum[256];
    int64_t loc[4], acc = 0;
    for (int q = 0; q < 4; ++q) { acc += sh[threadIdx.x * 4 + q]; loc[q] = acc; }
    tsum[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int64_t add = threadIdx.x >= (unsigned)off ? tsum[threadIdx.x - off] : 0;
        __syncthreads();
        tsum[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t pre = part[blockIdx.x] + (threadIdx.x ? tsum[threadIdx.x - 1] : 0);
    for (int q = 0; q < 4; ++q) {
        int64_t idx = base + threadIdx.x * 4 + q;
        if (idx < m) out[idx] = (int32_t)(pre + loc[q] - sh[threadIdx.x * 4 + q]);
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[m] = (int32_t)part[gridDim.x];
}

}  // namespace

// Exclusive scan of m int32 counts into out[0..m] (out[m] = total); returns the total and
// raises FAIR_ERR_OVERFLOW if it exceeds the reference's int32 range.
int64_t exclusive_scan_i32(const int32_t *d_counts, int64_t m, int32_t *d_out) {
    int64_t nb = (m + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb < 1) nb = 1;
    DBuf<int64_t> part(nb + 1);
    hipLaunchKernelGGL(k_block_sums, dim3((unsigned)nb), dim3(256), 0, stream(), d_counts, m, part.p);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, stream(), part.p, nb);
    int64_t total = 0;
    FAIR_HIP(hipMemcpyAsync(&total, part.p + nb, sizeof(int64_t), hipMemcpyDeviceToHost, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
    FAIR_CHECK(total <= INT32_MAX, FAIR_ERR_OVERFLOW,
               "nnz " + std::to_string(total) + " exceeds the int32 index range");
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, stream(), d_counts, m, part.p, d_out);
    launch_check("exclusive_scan");
    return total;
}

// Matrix-free projector support (fan beam tables as in trace<BEAM_FAN>): the near-axis rays
// and their entries by the matrix build's rule.
void mf_exact_rays(int32_t n, int32_t p, int32_t nth, const double *d_ang, const double *d_rayc,
                   std::vector<int32_t> &rays, std::vector<int32_t> &offs, std::vector<int32_t> &cols,
                   std::vector<float> &vals) {
    const RayArgs g{n, p, nth, d_ang, d_rayc};
    const int64_t m = (int64_t)nth * p;
    // at most the central rays of the angles close to multiples of pi/2
    const int32_t cap = 4 * nth + 64;
    DBuf<int32_t> list(cap), cnt(1);
    FAIR_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t), stream()));
    if (m > 0)
        hipLaunchKernelGGL(k_mf_near_axis, dim3(grid_for(m, 256)), dim3(256), 0, stream(), g, m, list.p, cnt.p, cap);
    launch_check("near-axis ray list");
    int32_t hc = 0;
    cnt.download(&hc, 1);
    FAIR_HIP(hipStreamSynchronize(stream()));
    FAIR_CHECK(hc <= cap, FAIR_ERR_INTERNAL, "too many near-axis rays");
    rays.assign(hc, 0);
    list.download(rays.data(), hc);
    FAIR_HIP(hipStreamSynchronize(stream()));
    std::sort(rays.begin(), rays.end());
    offs.assign(1, 0);
    cols.clear();
    vals.clear();
    if (hc == 0) return;
    list.upload(rays.data(), hc);
    const size_t per_ray = 36 * (size_t)(2 * n + 2);
    DBuf<char> scratch(per_ray * hc);
    DBuf<int32_t> dcnt(hc), doff(hc);
    hipLaunchKernelGGL(k_mf_exact_entries<false>, dim3(grid_for(hc, 64)), dim3(64), 0, stream(), g, list.p, hc,
                       scratch.p, per_ray, dcnt.p, (const int32_t *)nullptr, (int32_t *)nullptr, (float *)nullptr);
    launch_check("near-axis entry count");
    std::vector<int32_t> c(hc);
    dcnt.download(c.data(), hc);
    FAIR_HIP(hipStreamSynchronize(stream()));
    for (int32_t q = 0; q < hc; ++q) offs.push_back(offs.back() + c[q]);
    const int32_t tot = offs.back();
    if (tot == 0) return;
    doff.upload(offs.data(), hc);
    DBuf<int32_t> dcols(tot);
    DBuf<float> dvals(tot);
    hipLaunchKernelGGL(k_mf_exact_entries<true>, dim3(grid_for(hc, 64)), dim3(64), 0, stream(), g, list.p, hc,
                       scratch.p, per_ray, dcnt.p, (const int32_t *)doff.p, dcols.p, dvals.p);
    launch_check("near-axis entries");
    cols.assign(tot, 0);
    vals.assign(tot, 0.f);
    dcols.download(cols.data(), tot);
    dvals.download(vals.data(), tot);
    FAIR_HIP(hipStreamSynchronize(stream()));
}

namespace {

template <int BEAM>
void trace(const fair_geom &geo, const std::vector<double> &ang, const std::vector<double> &rayc,
           Csr &A) {
    const int64_t m = (int64_t)geo.nth * geo.p;
    FAIR_CHECK(m <= INT32_MAX, FAIR_ERR_OVERFLOW, "nth*p exceeds the int32 row range");
    FAIR_CHECK((int64_t)geo.n * geo.n <= INT32_MAX, FAIR_ERR_OVERFLOW, "n**2 exceeds int32");
    DBuf<double> d_ang(ang.size()), d_rayc(rayc.size());
    d_ang.upload(ang.data(), ang.size());
    d_rayc.upload(rayc.data(), rayc.size());
    RayArgs g{geo.n, geo.p, geo.nth, d_ang.p, d_rayc.p};

    DBuf<int32_t> counts(m > 0 ? m : 1), irr_list(m > 0 ? m : 1), irr_n(1), err(1);
    DBuf<uint8_t> irregular(m > 0 ? m : 1);
    FAIR_HIP(hipMemsetAsync(irr_n.p, 0, sizeof(int32_t), stream()));
    FAIR_HIP(hipMemsetAsync(err.p, 0, sizeof(int32_t), stream()));
    if (m > 0)
        hipLaunchKernelGGL(k_count<BEAM>, dim3(grid_for(m, 256)), dim3(256), 0, stream(), g, m,
                           counts.p, irregular.p, irr_list.p, irr_n.p);
    launch_check("ray count");
    int32_t nirr = 0;
    irr_n.download(&nirr, 1);
    FAIR_HIP(hipStreamSynchronize(stream()));

    const int nn = 2 * geo.n + 2;
    // irregular rays: the sort in LDS when its arrays fit, else in global scratch
    const size_t lds_bytes = (size_t)EXACT_LDS_PER_CROSSING * nn;
    const bool in_lds = lds_bytes <= (size_t)lds_per_block();
    const size_t per_ray = ((size_t)3 * nn * sizeof(double) + (size_t)3 * nn * sizeof(int32_t) + 255) & ~(size_t)255;
    DBuf<char> scratch(nirr > 0 && !in_lds ? per_ray * (size_t)nirr : 1);
    if (nirr > 0 && in_lds)
        hipLaunchKernelGGL((k_exact_lds<BEAM, false>), dim3((unsigned)nirr), dim3(64), lds_bytes, stream(), g,
                           irr_list.p, counts.p, nullptr, nullptr, nullptr, err.p);
    else if (nirr > 0)
        hipLaunchKernelGGL((k_exact<BEAM, false>), dim3(grid_for(nirr, 64)), dim3(64), 0, stream(),
                           g, irr_list.p, nirr, scratch.p, per_ray, counts.p, nullptr, nullptr,
                           nullptr, err.p);
    launch_check("ray count (exact)");

    A.m = (int32_t)m;
    A.n = geo.n * geo.n;
    A.IA.alloc(m + 1);
    A.nnz = exclusive_scan_i32(counts.p, m, A.IA.p);
    alloc_entries(A, A.nnz);
    if (m > 0)
        hipLaunchKernelGGL(k_fill<BEAM>, dim3(grid_for(m, 256)), dim3(256), 0, stream(), g, m,
                           A.IA.p, irregular.p, A.JA.p, A.vA.p, err.p);
    if (nirr > 0 && in_lds)
        hipLaunchKernelGGL((k_exact_lds<BEAM, true>), dim3((unsigned)nirr), dim3(64), lds_bytes, stream(), g,
                           irr_list.p, nullptr, A.IA.p, A.JA.p, A.vA.p, err.p);
    else if (nirr > 0)
        hipLaunchKernelGGL((k_exact<BEAM, true>), dim3(grid_for(nirr, 64)), dim3(64), 0, stream(),
                           g, irr_list.p, nirr, scratch.p, per_ray, nullptr, A.IA.p, A.JA.p,
                           A.vA.p, err.p);
    launch_check("ray fill");
    int32_t herr = 0;
    err.download(&herr, 1);
    FAIR_HIP(hipStreamSynchronize(stream()));
    FAIR_CHECK(!(herr & 1), FAIR_ERR_INTERNAL, "++ Something went wrong");
    FAIR_CHECK(!(herr & 2), FAIR_ERR_INTERNAL, "Didn't hit all elements... somehow");
    A.n_rows_nz = -1;
    build_row_list(A);
}

void check_geom(const fair_geom *g, const double *theta) {
    FAIR_CHECK(g != nullptr && theta != nullptr, FAIR_ERR_ARG, "null geometry or theta");
    FAIR_CHECK(g->n >= 1 && g->p >= 1 && g->nth >= 1, FAIR_ERR_ARG,
               "geometry needs n, p, nth >= 1");
}

}  // namespace
}  // namespace fair

using namespace fair;

extern "C" int fair_fanlineartomo(const fair_geom *g, const double *theta, fair_csr **out) {
    return guard([&] {
        check_geom(g, theta);
        FAIR_CHECK(out != nullptr, FAIR_ERR_ARG, "null output handle");
        const int n = g->n, p = g->p, nth = g->nth;
        // system_generation.f90:98-130, evaluated on the host in the reference's order
        const double R = g->r * (double)n, dw = g->dw * (double)n, sd = g->sd * (double)n;
        const double x0 = 0.0, y0 = R, dew = dw / (double)p;
        const double pi = 4.0 * atan(1.0);
        std::vector<double> omega(p);
        if (p % 2 == 1) {
            int h = (p - 1) / 2;
            for (int j = 1; j <= h; ++j) omega[j - 1] = atan(sd / ((double)(h + 1 - j) * dew)) - .5;
            omega[h] = 0.0;
            for (int k = 1; k <= h; ++k) omega[h + k] = .5 * pi - atan(sd / ((double)k * dew));
        } else {
            int h = p / 2;
            for (int j = 1; j <= h; ++j)
                omega[j - 1] = atan(sd / (((double)(h + 1 - j) - .5) * dew)) - .5 * pi;
            for (int k = 1; k <= h; ++k)
                omega[h + k - 1] = .5 * pi - atan(sd / (((double)k - .5) * dew));
        }
        std::vector<double> rayc(2 * (size_t)p), ang(4 * (size_t)nth);
        for (int j = 0; j < p; ++j) {
            double s, c;
            sincos(omega[j], &s, &c);
            rayc[2 * j] = c;
            rayc[2 * j + 1] = s;
        }
        for (int i = 0; i < nth; ++i) {
            double st, ct;
            sincos(theta[i], &st, &ct);
            double x0th = ct * x0 - st * y0;
            double y0th = st * x0 + ct * y0;
            ang[4 * i] = x0th;
            ang[4 * i + 1] = y0th;
            ang[4 * i + 2] = -(x0th / R);
            ang[4 * i + 3] = -(y0th / R);
        }
        auto *h = new fair_csr();
        try {
            trace<BEAM_FAN>(*g, ang, rayc, h->A);
        } catch (...) {
            delete h;
            throw;
        }
        h->g_n = n;
        h->g_p = p;
        h->g_nth = nth;
        *out = h;
    });
}

extern "C" int fair_paralleltomo(const fair_geom *g, const double *theta, fair_csr **out) {
    return guard([&] {
        check_geom(g, theta);
        FAIR_CHECK(out != nullptr, FAIR_ERR_ARG, "null output handle");
        const int n = g->n, p = g->p, nth = g->nth;
        (void)n;
        // system_generation.f90:328-330
        const double d = (double)(p - 1);
        std::vector<double> x0(p);
        fair_linspace(-.5 * d, .5 * d, p, x0.data());
        std::vector<double> rayc(2 * (size_t)p), ang(2 * (size_t)nth);
        for (int j = 0; j < p; ++j) {
            rayc[2 * j] = x0[j];
            rayc[2 * j + 1] = 0.0 * x0[j];
        }
        for (int i = 0; i < nth; ++i) {
            double st, ct;
            sincos(theta[i], &st, &ct);
            ang[2 * i] = ct;
            ang[2 * i + 1] = st;
        }
        auto *h = new fair_csr();
        try {
            trace<BEAM_PAR>(*g, ang, rayc, h->A);
        } catch (...) {
            delete h;
            throw;
        }
        h->g_n = g->n;
        h->g_p = p;
        h->g_nth = nth;
        *out = h;
    });
}

extern "C" int fair_linspace(double a, double b, int32_t n, double *x) {
    return guard([&] {
        FAIR_CHECK(x != nullptr || n <= 0, FAIR_ERR_ARG, "null output");
        double c = (b - a) / (double)(n - 1);
        double d = a - c;
        for (int i = 1; i <= n; ++i) x[i - 1] = c * (double)i + d;
    });
}
