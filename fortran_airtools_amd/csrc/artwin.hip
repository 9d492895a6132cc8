// artwin.hip -- Kaczmarz row sequences without a level schedule (random orders, row-norm
// sampling, a caller's sequence), walked a window of W steps at a time by one workgroup.
//
// basic_sweep (AIRMs.f90:813-858) is a chain: each row step reads the pixels the earlier steps
// wrote.  A one-wavefront sweep pays a global-memory round trip and the whole row's instruction
// stream per step.  Here the work of a window that does not depend on the order is done side by
// side: every step's entries and x values are loaded at the window's start, and only the values
// of pixels that two or more steps of the window touch ("shared" pixels) pass from step to step,
// through LDS.  A step waits only for the earlier steps of its window that share a pixel with it, so
// steps that do not conflict run at the same time.
//
// A pre-pass kernel classifies each entry of each step (per window, from a bit mask of the steps
// touching each pixel) into a 16-bit code, idx | P << 14 | L << 15, where idx is the LDS index
// the entry's value is read from once the step's wait is over (PUB: the number of shared slots):
//   idx >= PUB   the pixel is touched by no other step of the window: idx is the entry's own
//                parking place (wavefront, register slot, lane), where the step puts its
//                window-start value before the wait; it stores its result to x;
//   idx <  PUB   shared: idx is the pixel's LDS slot.  P = an earlier step of the window touched
//                it (the slot then holds that step's value), else the step itself puts the
//                window-start value there; L = no later step touches it (store the result to x);
// and records per step the mask of the earlier steps of its window that share a pixel with it.
// The executor then computes every row exactly as k_art_sweep does (entries of lane q:
// beg + q + 64 t in increasing t, products and sums rounded separately, the DPP wavefront sum,
// coeff = relax*(b - dot)/nrA, x + coeff*a then max(lb), min(ub)), on the same values in the
// same order, so its iterate equals k_art_sweep's bit for bit.  Columns outside the image (the
// parallel-beam tracer emits some above N^2 when an angle is exactly pi) read zero and are not
// stored, as in k_art_sweep.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "fair_internal.hpp"

namespace fair {
namespace {

constexpr int WIN_IDX_MASK = 0x3FFF;
constexpr int WIN_HASPRED = 1 << 14;
constexpr int WIN_LAST = 1 << 15;
constexpr int WIN_PRE_WAVES = 8;   // pre-pass workgroup: 512 threads

// Window sizes: W steps, NW wavefronts of SPW steps each (wavefront v: steps v, v + NW, ...).
// Rows up to 64 R entries.  The executor's LDS holds PUB = W * 64 R / 2 doubles, one per shared
// pixel (at most half of a window's entries), and a parking place per wavefront, register slot and
// lane: 96 KB for both shapes.
template <int R> struct WinShape;
template <> struct WinShape<8> { static constexpr int NW = 8, SPW = 4; };
template <> struct WinShape<16> { static constexpr int NW = 8, SPW = 1; };

template <int R> constexpr int win_steps() { return WinShape<R>::NW * WinShape<R>::SPW; }
template <int R> constexpr int win_pub() { return win_steps<R>() * WAVE * R / 2; }
// the executor's LDS: the shared slots, then the parking places (NW wavefronts x R slots x 64 lanes)
template <int R> constexpr int win_lds() { return win_pub<R>() + WinShape<R>::NW * R * WAVE; }
static_assert(win_lds<8>() <= WIN_IDX_MASK + 1 && win_lds<16>() <= WIN_IDX_MASK + 1, "LDS index field");
// the parking place of lane `lane`, register slot u, of the wavefront that runs window step i
template <int R>
__device__ __forceinline__ int32_t win_park(int i, int u, int lane) {
    return win_pub<R>() + ((i % WinShape<R>::NW) * R + u) * WAVE + lane;
}

struct WinChains {
    int64_t nsteps[2];
    double *x[2];
    const int4 *desc[2];        // per step {row (zero-based), IA(row), IA(row+1), dep}
    const double2 *bw[2];       // per step {b(row), nrA(row)}
    const int32_t *cslot[2];    // per step: LDS slots of its first 64 shared entries
    const double *cval[2];      // per step: their matrix values
    const uint16_t *codes[2];   // 64 R per step, lane-major (the R codes of a lane are adjacent)
};

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- pre-pass: entry codes and step descriptors of one chain -------------------------------------
// Each workgroup takes windows blockIdx.x, + gridDim.x, ... and owns the tables mask / slotof
// (n entries each) at offset blockIdx.x * n; mask is all zero between windows (every window
// clears the entries it set).  Within a window:
//   A  every entry ORs its step's bit into mask[pixel];
//   B  every entry reads the mask; the first step of a shared pixel takes an LDS slot number
//      and publishes it in slotof[pixel]; each step's dependency mask is the OR of the earlier
//      bits of its pixels;
//   C  every entry writes its code; each step lists its shared entries (LDS slot and matrix
//      value, the first 64 in slot order) and writes its descriptor {number of shared entries,
//      entry range, dependency mask}, b and nrA;
//   D  every entry clears mask[pixel].
// The tables are private to the workgroup and every access to them is performed in the L2 (atomics,
// and loads and stores at device scope, which bypass the CU's L1): each wavefront waits until its
// own accesses are done before the barrier that ends a phase, and no cache write-back is needed
// (a device-wide fence would write back the L2 in every phase, slowing a sweep running beside).
template <int R>
__global__ void __launch_bounds__(WIN_PRE_WAVES * WAVE) k_win_codes(const int32_t *__restrict__ rows,
                                                                    int64_t nsteps,
                                                                    const int32_t *__restrict__ IA,
                                                                    const int32_t *__restrict__ JA,
                                                                    const double *__restrict__ vA,
                                                                    const double *__restrict__ b,
                                                                    const double *__restrict__ nrA,
                                                                    int32_t ncols, uint32_t *mask_all,
                                                                    int32_t *slot_all, uint16_t *codes,
                                                                    int4 *desc, double2 *bw, int32_t *cslot,
                                                                    double *cval) {
    constexpr int W = win_steps<R>();
    constexpr int SPW = W / WIN_PRE_WAVES;
    static_assert(W % WIN_PRE_WAVES == 0 && W <= 32, "window shape");
    __shared__ int32_t nslot;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    uint32_t *mask = mask_all + (size_t)blockIdx.x * (size_t)ncols;
    int32_t *slotof = slot_all + (size_t)blockIdx.x * (size_t)ncols;
    if (threadIdx.x == 0) nslot = 0;
    __syncthreads();
    const int64_t nwin = (nsteps + W - 1) / W;
    for (int64_t k = blockIdx.x; k < nwin; k += gridDim.x) {
        int32_t jc[SPW][R];
        uint32_t mk[SPW][R];
        int32_t rq[SPW], bq[SPW], eq[SPW], dq[SPW];
#pragma unroll
        for (int q = 0; q < SPW; ++q) {
            const int i = wv + WIN_PRE_WAVES * q;
            const int64_t s = k * W + i;
            const bool ok = s < nsteps;
            const int32_t r = ok ? rows[s] - 1 : 0;
            const int32_t beg = IA[r], end = ok ? IA[r + 1] : beg;
            rq[q] = r;
            bq[q] = beg;
            eq[q] = end;
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int32_t e = beg + lane + WAVE * u;
                const int32_t j = e < end ? JA[e] : -1;
                jc[q][u] = (uint32_t)j < (uint32_t)ncols ? j : -1;
                if (jc[q][u] >= 0)
                    (void)__hip_atomic_fetch_or(&mask[jc[q][u]], 1u << i, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's table accesses are done
        __syncthreads();
#pragma unroll
        for (int q = 0; q < SPW; ++q) {
            const int i = wv + WIN_PRE_WAVES * q;
            uint32_t dm = 0;
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int32_t j = jc[q][u];
                const uint32_t M = j >= 0 ? ld_agent(&mask[j]) : 0u;
                mk[q][u] = M;
                if (__popc(M) >= 2) {
                    if ((M & (0u - M)) == (1u << i)) {
                        const int32_t sl = atomicAdd(&nslot, 1);
                        __hip_atomic_store(&slotof[j], sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    const uint32_t below = M & ((1u << i) - 1u);
                    dm |= below;
                }
            }
#pragma unroll
            for (int off = WAVE / 2; off > 0; off >>= 1) dm |= (uint32_t)__shfl_xor((int)dm, off, WAVE);
            dq[q] = (int32_t)dm;
        }
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's table accesses are done
        __syncthreads();
#pragma unroll
        for (int q = 0; q < SPW; ++q) {
            const int i = wv + WIN_PRE_WAVES * q;
            const int64_t s = k * W + i;
            if (s >= nsteps) continue;
            int32_t nsh = 0;   // shared entries listed so far (wavefront-uniform)
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int32_t j = jc[q][u];
                const uint32_t M = mk[q][u];
                uint16_t c = (uint16_t)win_park<R>(i, u, lane);
                int32_t sl = 0;
                const bool sh = __popc(M) >= 2;
                if (sh) {
                    sl = __hip_atomic_load(&slotof[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    c = (uint16_t)((sl & WIN_IDX_MASK) | ((M & ((1u << i) - 1u)) ? WIN_HASPRED : 0) |
                                   ((M >> i) == 1u ? WIN_LAST : 0));
                }
                codes[s * (WAVE * R) + lane * R + u] = c;
                const uint64_t bal = __ballot(sh);
                const int32_t pos =
                    nsh + (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                if (sh && pos < WAVE) {
                    cslot[s * WAVE + pos] = sl & WIN_IDX_MASK;
                    cval[s * WAVE + pos] = vA[bq[q] + lane + WAVE * u];
                }
                nsh += __popcll(bal);
            }
            if (lane == 0) {
                desc[s] = make_int4(nsh, bq[q], eq[q], dq[q]);
                bw[s] = make_double2(b[rq[q]], nrA[rq[q]]);
            }
        }
        __syncthreads();   // every slot number is read; nslot is no longer used by this window
#pragma unroll
        for (int q = 0; q < SPW; ++q)
#pragma unroll
            for (int u = 0; u < R; ++u)
                if (jc[q][u] >= 0)
                    __hip_atomic_store(&mask[jc[q][u]], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (threadIdx.x == 0) nslot = 0;
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's table accesses are done
        __syncthreads();
    }
}

// wave_sum_dpp's sum with DPP moves that need no zero-filled destination, the four row sums
// joined in the vector unit: row_bcast:15 adds row 0 to row 1 and row 2 to row 3, row_bcast:31
// then adds row 1 to row 3, so lane 63 holds (r3 + r2) + (r1 + r0), which is wave_sum_dpp's
// (r0 + r1) + (r2 + r3) bit for bit (each add's operands are the same two values).
template <int CTRL>
__device__ __forceinline__ double win_dpp(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double win_wave_sum(double v) {
    v += win_dpp<0xB1>(v);
    v += win_dpp<0x4E>(v);
    v += win_dpp<0x141>(v);
    v += win_dpp<0x140>(v);
    v += win_dpp<0x142>(v);   // row_bcast:15 (row 0 reads out of range: 0)
    v += win_dpp<0x143>(v);   // row_bcast:31 (rows 0 and 1 read out of range: 0)
    return lane_value(v, 63);
}

// The value a padded slot gathers (its product must be +0 whatever x holds).
__device__ double g_win_zero = 0.0;

template <int R> __device__ __forceinline__ bool win_shared(int32_t c) { return (c & WIN_IDX_MASK) < win_pub<R>(); }

// ---- executor ----------------------------------------------------------------------------------
// One workgroup per chain.  Each wavefront holds the entries and codes of its SPW steps of the
// current window, and the descriptors of those steps in this window and the next (one lane per
// step); right after its turn a step loads its next-window entries, so at a window's start only
// the x values are loaded (the barrier that ends a window orders its stores before them).
// A turn: wait until done[l] == window for every l <= dep (LDS flags, one per step of the window);
// replace the values of the pixels an earlier step of the window wrote by theirs from LDS; compute
// the row; write the new values of its shared pixels to LDS and mark itself done; then store the
// unshared pixels (and the shared ones no later step touches) to x.  Every slot is computed as
// k_art_sweep computes it (slots past the row hold zero products and are not stored), without
// branches, so the compiler can overlap the slots.  While a step computes its row its wavefront
// runs at a raised issue priority, ahead of the wavefronts that poll.
// Waits are bounded: after spin_limit polls a wavefront sets the stop flag and *err and stops
// computing, the other wavefronts stop at their next poll, and every wavefront still passes the
// same barriers, so the launch ends normally; a bound of 0 fails the first wait (failure-path
// tests).
template <int R>
__global__ void __launch_bounds__(WinShape<R>::NW *WAVE) k_art_window(WinChains ch, const int32_t *__restrict__ JA,
                                                                     const double *__restrict__ vA, int32_t pad,
                                                                     int32_t ncols, double relax, int lwrb, double lb, int upb,
                                                                     double ub, int32_t spin_limit, int32_t *err) {
    constexpr int NW = WinShape<R>::NW, SPW = WinShape<R>::SPW, W = NW * SPW;
    // pub: a double per shared pixel; after it, per wavefront and register slot, the 64 lanes'
    // window-start values of the step being waited for (the patch reads take every slot from LDS)
    __shared__ double pub[win_lds<R>()];
    __shared__ int32_t done[W + 1];   // done[W]: the stop flag (0 while running)
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t nsteps = ch.nsteps[blockIdx.x];
    double *x = ch.x[blockIdx.x];
    const int4 *__restrict__ desc = ch.desc[blockIdx.x];
    const double2 *__restrict__ bw = ch.bw[blockIdx.x];
    const uint16_t *__restrict__ codes = ch.codes[blockIdx.x];
    const int32_t *__restrict__ cslot = ch.cslot[blockIdx.x];
    const double *__restrict__ cval = ch.cval[blockIdx.x];
    if (threadIdx.x <= W) done[threadIdx.x] = threadIdx.x < W ? -1 : 0;
    __syncthreads();
    const int64_t nwin = (nsteps + W - 1) / W;
    if (nwin == 0) return;   // the whole workgroup leaves together
    // The entries of a step are held in register slot q % P from two turns of the wavefront before
    // its own (P = 2, when SPW is even; else one window ahead, P = SPW): the registers then do not
    // grow with the window.
    constexpr int P = SPW % 2 == 0 ? 2 : SPW;
    int32_t jc[P][R];
    uint32_t cw[P][R / 2];   // the lane's codes, two per word
    double vc[P][R];
    // x at the pixels of the next step, or of the next two (by parity of q) where the registers
    // allow (three steps or fewer per wavefront)
    constexpr bool TWO = SPW <= 3;
    double xva[R], xvb[R];
    int32_t ccs[P];            // lane t: LDS slot and matrix value of the step's shared entry t
    double cva[P];
    // Descriptors of the wavefront's steps, lane q holding step q: {beg, end, dep} and {b, nrA} of
    // this window (cur) and of the next (nxt), loaded a window ahead by one vector load each.
    int4 cur_d, nxt_d;
    double2 cur_b, nxt_b;
    bool bailed = false;
    // the box bounds live in vector registers: a select then needs no scalar-to-vector move
    double lbv = lb, ubv = ub;
    asm volatile("" : "+v"(lbv), "+v"(ubv));
    auto code = [&](int q, int u) -> int32_t {
        return (int32_t)((cw[q % P][u >> 1] >> ((u & 1) * 16)) & 0xFFFFu);
    };
    // lanes 0..SPW-1: the descriptors of the wavefront's steps of window k (a step past the end: no
    // entries, no wait)
    auto load_descs = [&](int64_t k, int4 &d, double2 &bv) {
        const int64_t s = k * W + wv + NW * (lane < SPW ? lane : 0);
        const bool ok = s < nsteps;
        const int4 t = desc[ok ? s : 0];
        d = ok ? t : make_int4(t.x, t.y, t.y, 0);
        bv = bw[ok ? s : 0];
    };
    // entries and codes of step q of window k, entry range [beg, end).  Slots past the row load
    // the zero spare entry `pad` after the matrix's entries (column 0, value 0): no load result is
    // selected afterwards, so the loads stay in flight together; a slot's validity is tested
    // where it matters (lane + 64 u < the row's length).
    auto load_entries = [&](int q, int64_t k, int32_t beg, int32_t end) {   // step q of window k
        const int64_t s = k * W + wv + NW * q;
        const int64_t sc = s < nsteps ? s : 0;
        // with 16 slots per lane the slots past the row's end are not loaded (they would all read the
        // spare entry; measured faster at N=512); with 8, where rows fill most slots, the test costs
        // more than it saves
#pragma unroll
        for (int g = 0; g < R / 4; ++g) {
            if (R == 8 || g == 0 || WAVE * 4 * g < end - beg) {
#pragma unroll
                for (int u = 4 * g; u < 4 * g + 4; ++u) {
                    const int32_t e = beg + lane + WAVE * u;
                    const int32_t ee = e < end ? e : pad;
                    jc[q % P][u] = JA[ee];
                    vc[q % P][u] = vA[ee];
                }
            } else {
#pragma unroll
                for (int u = 4 * g; u < 4 * g + 4; ++u) {
                    jc[q % P][u] = 0;
                    vc[q % P][u] = 0.0;
                }
            }
        }
        ccs[q % P] = cslot[sc * WAVE + lane];
        cva[q % P] = cval[sc * WAVE + lane];
        const uint4 *cp = reinterpret_cast<const uint4 *>(codes + sc * (WAVE * R) + lane * R);
#pragma unroll
        for (int h = 0; h < R / 8; ++h) {
            const uint4 c4 = cp[h];
            cw[q % P][4 * h] = c4.x;
            cw[q % P][4 * h + 1] = c4.y;
            cw[q % P][4 * h + 2] = c4.z;
            cw[q % P][4 * h + 3] = c4.w;
        }
    };
    load_descs(0, cur_d, cur_b);
#pragma unroll
    for (int q = 0; q < P; ++q)
        load_entries(q, 0, __builtin_amdgcn_readlane(cur_d.y, q), __builtin_amdgcn_readlane(cur_d.z, q));
    if (nwin > 1) load_descs(1, nxt_d, nxt_b);
    for (int64_t k = 0; k < nwin; ++k) {
        const int32_t kw = (int32_t)k;
        // x of a step's pixels is loaded one or two turns of its wavefront ahead (the first at the
        // window's start): no store of this window can reach them first (an unshared pixel has no
        // other writer; a shared one is stored by its last occurrence, which waits for this step,
        // or read from LDS)
        auto gather = [&](int q, double (&dst)[R]) {
            const int32_t len = __builtin_amdgcn_readlane(cur_d.z, q) - __builtin_amdgcn_readlane(cur_d.y, q);
#pragma unroll
            for (int u = 0; u < R; ++u)
                dst[u] = *(lane + WAVE * u < len && (uint32_t)jc[q % P][u] < (uint32_t)ncols ? x + jc[q % P][u] : &g_win_zero);
        };
        gather(0, xva);
        if (TWO && SPW > 1) gather(1, xvb);
#pragma unroll
        for (int q = 0; q < SPW; ++q) {
            const int i = wv + NW * q;
            if (k * W + i >= nsteps) break;
            double(&xv)[R] = (TWO && (q & 1)) ? xvb : xva;
            const uint32_t dm = (uint32_t)__builtin_amdgcn_readlane(cur_d.w, q);   // steps to wait for
            const int32_t nsh = __builtin_amdgcn_readlane(cur_d.x, q);
            const bool compact = nsh <= WAVE;
            // a first occurrence of a shared pixel puts the window-start value in the pixel's LDS
            // slot (a later occurrence reads it there once this step is done)
            if (!bailed)
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    const int32_t c = code(q, u);
                    if (win_shared<R>(c) && !(c & WIN_HASPRED)) pub[c & WIN_IDX_MASK] = xv[u];
                }
            const int32_t len = __builtin_amdgcn_readlane(cur_d.z, q) - __builtin_amdgcn_readlane(cur_d.y, q);
            // register slots holding entries of the row (wave-uniform); the slots past them hold
            // zero products only, which the dot product leaves out
            const int nsl = (len + WAVE - 1) / WAVE;
            // The step's turn on its row values xw: the lane sums (k_art_sweep's order; the slots
            // past nsl hold +0 products, which change no sum, and are left out), the wavefront sum,
            // coeff, the shared pixels published and the step marked done, then the stores.
            auto turn = [&](double(&xw)[R]) {
                // the current values of the listed shared entries (lane t: entry t)
                const double cv = pub[lane < nsh ? ccs[q % P] : 0];
                __builtin_amdgcn_s_setprio(2);
                double dot = 0.0;
#pragma unroll
                for (int g = 0; g < R / 4; ++g) {
                    if (g > 0 && 4 * g >= nsl) break;
#pragma unroll
                    for (int u = 4 * g; u < 4 * g + 4; ++u) dot += xw[u] * vc[q % P][u];
                }
                dot = win_wave_sum(dot);
                const double coeff = (relax * (lane_value(cur_b.x, q) - dot)) / lane_value(cur_b.y, q);
                if (compact) {
                    // only the shared pixels are needed before the step is marked done
                    if (lane < nsh) {
                        double v = cv + coeff * cva[q % P];
                        if (lwrb) v = fmax_ref(v, lbv);
                        if (upb) v = fmin_ref(v, ubv);
                        pub[ccs[q % P]] = v;
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < R; ++u) {
                        if (u >= nsl) break;
                        double v = xw[u] + coeff * vc[q % P][u];
                        if (lwrb) v = fmax_ref(v, lbv);
                        if (upb) v = fmin_ref(v, ubv);
                        xw[u] = v;
                        const int32_t c = code(q, u);
                        if (win_shared<R>(c)) pub[c & WIN_IDX_MASK] = v;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                if (lane == 0) __hip_atomic_store(&done[i], kw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __builtin_amdgcn_s_setprio(0);
                // slots past the row (u >= nsl) hold no entry in any lane: skipped by a uniform test;
                // columns outside the image take no store
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    if (u >= nsl) break;
                    double v = xw[u];
                    if (compact) {
                        v = v + coeff * vc[q % P][u];
                        if (lwrb) v = fmax_ref(v, lbv);
                        if (upb) v = fmin_ref(v, ubv);
                    }
                    const int32_t c = code(q, u);
                    if (lane + WAVE * u < len && (uint32_t)jc[q % P][u] < (uint32_t)ncols &&
                        (!win_shared<R>(c) || (c & WIN_LAST)))
                        x[jc[q % P][u]] = v;
                }
            };
            bool waited = false;
            if (dm != 0 && !bailed) {
                // every slot's window-start value goes to its parking place, so that after the wait
                // each slot reads its value from LDS at the index its code holds, with no select
#pragma unroll
                for (int g = 0; g < R / 4; ++g) {
                    if (g > 0 && 4 * g >= nsl) break;
#pragma unroll
                    for (int u = 4 * g; u < 4 * g + 4; ++u) pub[win_park<R>(wv, u, lane)] = xv[u];
                }
                // one LDS access per poll: lane l < W of the mask reads step l's flag, the other
                // lanes the stop flag (lane 63 among them)
                const bool mine = lane < W && ((dm >> (lane & 31)) & 1u);
                const int32_t want = mine ? kw : 0;
                if (spin_limit <= 0) {
                    __hip_atomic_store(&done[W], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (lane == 0) atomicExch(err, 1);
                    bailed = true;
                } else {
                    int32_t spins = 0;
                    while (true) {
                        const int32_t v = __hip_atomic_load(&done[mine ? lane : W], __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (__all(v == want)) break;
                        if (__builtin_amdgcn_readlane(v, WAVE - 1) != 0) {   // stopped elsewhere
                            bailed = true;
                            break;
                        }
                        if (++spins >= spin_limit) {
                            __hip_atomic_store(&done[W], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (lane == 0) atomicExch(err, 1);
                            bailed = true;
                            break;
                        }
                    }
                }
                __builtin_amdgcn_s_setprio(2);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                waited = !bailed;
            }
            if (waited) {
                // every slot read back from LDS, the reads issued back to back
#pragma unroll
                for (int g = 0; g < R / 4; ++g) {
                    if (g > 0 && 4 * g >= nsl) break;
#pragma unroll
                    for (int u = 4 * g; u < 4 * g + 4; ++u) xv[u] = pub[code(q, u) & WIN_IDX_MASK];
                }
            }
            if (!bailed) turn(xv);
            if (TWO) {
                if (q + 2 < SPW && k * W + i + 2 * NW < nsteps) gather(q + 2, xv);
            } else if (q + 1 < SPW && k * W + i + NW < nsteps) {
                gather(q + 1, xva);
            }
            // the entries of the wavefront's step P turns ahead, into the slot this step held
            // (loading the next window's past the barrier instead delays the window's first steps:
            // measured slower)
            if (q + P < SPW)
                load_entries(q + P, k, __builtin_amdgcn_readlane(cur_d.y, q + P),
                             __builtin_amdgcn_readlane(cur_d.z, q + P));
            else if (k + 1 < nwin)
                load_entries(q + P - SPW, k + 1, __builtin_amdgcn_readlane(nxt_d.y, q + P - SPW),
                             __builtin_amdgcn_readlane(nxt_d.z, q + P - SPW));
        }
        __syncthreads();
        cur_d = nxt_d;
        cur_b = nxt_b;
        if (k + 2 < nwin) load_descs(k + 2, nxt_d, nxt_b);
    }
}

__global__ void k_zero_u32(uint32_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- the window sweep's pixel numbering in 4 x 4 tiles ---------------------------------------
// Pixel j = c N + r of an N x N image (c, r < N) becomes ((c/4) (NP/4) + r/4) 16 + (c%4) 4 + r%4,
// NP = N rounded up to a multiple of 4: the 16 pixels of a tile are consecutive (two 128-byte
// lines), so a ray's consecutive pixels, in whatever direction it runs, share lines: at N = 256 a
// row step's gathers touch 66 lines instead of 135 in the image's own order.  The sweep runs on
// the renumbered columns and a renumbered copy of x (copied exactly, in and out), so its iterate is
// the same, bit for bit.
__host__ __device__ __forceinline__ int32_t win_tiled(int32_t j, int32_t N, int32_t NP) {
    const int32_t c = j / N, r = j - c * N;
    return ((c >> 2) * (NP >> 2) + (r >> 2)) * 16 + (c & 3) * 4 + (r & 3);
}

// columns of the matrix's entries (and its spare entries, column 0) renumbered; columns outside
// the image become INT32_MAX (outside the renumbered image too)
__global__ void k_tile_cols(const int32_t *__restrict__ JA, int32_t *__restrict__ JAt, int64_t total, int32_t n,
                            int32_t N, int32_t NP) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int32_t j = JA[e];
    JAt[e] = (uint32_t)j < (uint32_t)n ? win_tiled(j, N, NP) : INT32_MAX;
}

// xt[t] = x at the pixel renumbered t (0 for the padding pixels), for t < NP^2
__global__ void k_x_to_tiles(const double *__restrict__ x, double *__restrict__ xt, int32_t N, int32_t NP) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)NP * NP) return;
    const int32_t tile = (int32_t)(t >> 4), w = (int32_t)(t & 15);
    const int32_t c = (tile / (NP >> 2)) * 4 + (w >> 2), r = (tile % (NP >> 2)) * 4 + (w & 3);
    xt[t] = c < N && r < N ? x[(int64_t)c * N + r] : 0.0;
}

__global__ void k_x_from_tiles(const double *__restrict__ xt, double *__restrict__ x, int32_t N, int32_t NP) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int64_t)N * N) return;
    x[j] = xt[win_tiled((int32_t)j, N, NP)];
}

}  // namespace

bool art_window_usable(Csr &A, int64_t nsteps) {
    if (!env_int("FAIR_ART_WINDOW", 1) || A.n <= 0 || nsteps < 2) return false;
    if (A.n_rows_nz < 0) build_row_list(A);
    return A.max_row <= 16 * WAVE;
}

namespace {

int win_R(const Csr &A) { return A.max_row <= 8 * WAVE ? 8 : 16; }

}  // namespace

// The pre-pass of one sequence (one-based rows on the device) into buffer set `slot` (0 or 1), on
// stream s.  (Run beside the previous sweep on a second stream it slowed that sweep more than it
// saved: uniform randomized Kaczmarz at N = 256, 20 sweeps, 0.42 s against 0.38 s in sequence.)
void art_window_prepare(Csr &A, const int32_t *cols, int32_t ncols, int slot, const int32_t *rows, int64_t nsteps,
                        const double *b, const double *nrA, hipStream_t s) {
    if (A.n_rows_nz < 0) build_row_list(A);
    FAIR_CHECK(A.max_row <= 16 * WAVE, FAIR_ERR_ARG, "row too long for the window sweep");
    FAIR_CHECK(slot == 0 || slot == 1, FAIR_ERR_INTERNAL, "window buffer set");
    const int R = win_R(A);
    ArtWindowBufs &wb = A.win;
    const int64_t n = ncols;
    // pre-pass tables: 8 bytes per pixel per workgroup, at most 128 workgroups (more tables crowd
    // the matrix out of the caches the sweep then reads it through) and 256 MB
    const int G = (int)std::max<int64_t>(8, std::min<int64_t>(128, ((int64_t)256 << 20) / (8 * n)));
    if (wb.G != G || wb.n != n) {
        wb.mask.alloc((size_t)G * (size_t)n);
        wb.slotof.alloc((size_t)G * (size_t)n);
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for((int64_t)G * n, 256)), dim3(256), 0, s, wb.mask.p,
                           (int64_t)G * n);
        launch_check("window tables");
        wb.G = G;
        wb.n = n;
    }
    const size_t need = (size_t)nsteps * (size_t)(WAVE * R);
    if (wb.codes[slot].n < need) wb.codes[slot].alloc(need);
    if (wb.desc[slot].n < 4 * (size_t)nsteps) wb.desc[slot].alloc(4 * (size_t)nsteps);
    if (wb.bw[slot].n < 2 * (size_t)nsteps) wb.bw[slot].alloc(2 * (size_t)nsteps);
    if (wb.cslot[slot].n < (size_t)nsteps * WAVE) wb.cslot[slot].alloc((size_t)nsteps * WAVE);
    if (wb.cval[slot].n < (size_t)nsteps * WAVE) wb.cval[slot].alloc((size_t)nsteps * WAVE);
    wb.nsteps[slot] = nsteps;
    if (nsteps <= 0) return;
    const int W = R == 8 ? win_steps<8>() : win_steps<16>();
    const int64_t nwin = (nsteps + W - 1) / W;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(G, nwin));
    if (R == 8)
        hipLaunchKernelGGL(k_win_codes<8>, dim3(grid), dim3(WIN_PRE_WAVES * WAVE), 0, s, rows, nsteps, A.IA.p,
                           cols, A.vA.p, b, nrA, ncols, wb.mask.p, wb.slotof.p, wb.codes[slot].p,
                           (int4 *)wb.desc[slot].p, (double2 *)wb.bw[slot].p, wb.cslot[slot].p, wb.cval[slot].p);
    else
        hipLaunchKernelGGL(k_win_codes<16>, dim3(grid), dim3(WIN_PRE_WAVES * WAVE), 0, s, rows, nsteps, A.IA.p,
                           cols, A.vA.p, b, nrA, ncols, wb.mask.p, wb.slotof.p, wb.codes[slot].p,
                           (int4 *)wb.desc[slot].p, (double2 *)wb.bw[slot].p, wb.cslot[slot].p, wb.cval[slot].p);
    launch_check("window codes");
}

// The executor over `chains` chains (chain c: buffer set slots[c], prepared earlier in stream order),
// iterate x[c] updated in place.  Waits for the sweep and reports a stop.
void art_window_run(Csr &A, const int32_t *cols, int32_t ncols, int chains, const int slots[2], double *const x[2],
                    double relax, int lwrb, double lb, int upb, double ub, bool check) {
    const int R = win_R(A);
    ArtWindowBufs &wb = A.win;
    FAIR_CHECK(wb.err.n >= 1, FAIR_ERR_INTERNAL, "window flag not cleared");
    WinChains ch{};
    for (int c = 0; c < 2; ++c) {
        const bool on = c < chains;
        const int sl = on ? slots[c] : 0;
        ch.nsteps[c] = on ? wb.nsteps[sl] : 0;
        ch.x[c] = on ? x[c] : nullptr;
        ch.codes[c] = on ? wb.codes[sl].p : nullptr;
        ch.desc[c] = on ? (const int4 *)wb.desc[sl].p : nullptr;
        ch.bw[c] = on ? (const double2 *)wb.bw[sl].p : nullptr;
        ch.cslot[c] = on ? wb.cslot[sl].p : nullptr;
        ch.cval[c] = on ? wb.cval[sl].p : nullptr;
    }
    const int32_t spin = std::max(0, env_int("FAIR_ART_WINDOW_SPIN", 1 << 22));
    // the first spare entry after the matrix's entries (JA = 0, vA = 0, see alloc_entries)
    const int32_t pad = (int32_t)A.nnz;
    if (R == 8)
        hipLaunchKernelGGL(k_art_window<8>, dim3(chains), dim3(WinShape<8>::NW * WAVE), 0, stream(), ch, cols,
                           A.vA.p, pad, ncols, relax, lwrb, lb, upb, ub, spin, wb.err.p);
    else
        hipLaunchKernelGGL(k_art_window<16>, dim3(chains), dim3(WinShape<16>::NW * WAVE), 0, stream(), ch, cols,
                           A.vA.p, pad, ncols, relax, lwrb, lb, upb, ub, spin, wb.err.p);
    launch_check("window sweep");
    if (env_int("FAIR_DEBUG", 0))
        fprintf(stderr, "[fair] window sweep: %lld + %lld steps, windows of %d, %d entries per lane\n",
                (long long)ch.nsteps[0], (long long)ch.nsteps[1], R == 8 ? win_steps<8>() : win_steps<16>(), R);
    if (check) art_window_check(A);
}

// The stop flag of the window sweeps since the last art_window_clear zeroed it, on stream order;
// synchronises and fails the call when a step waited beyond its bound.
void art_window_clear(Csr &A) {
    ArtWindowBufs &wb = A.win;
    if (wb.err.n == 0) wb.err.alloc(1);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream(), (uint32_t *)wb.err.p, (int64_t)1);
    launch_check("window flag");
}

void art_window_check(Csr &A) {
    ArtWindowBufs &wb = A.win;
    if (wb.err.n == 0) return;
    int32_t herr = 0;
    FAIR_HIP(hipMemcpyAsync(&herr, wb.err.p, sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
    FAIR_CHECK(herr == 0, FAIR_ERR_HIP, "Kaczmarz window sweep: a step waited beyond its bound");
}

namespace {

// The side N of the square image whose pixels the matrix's columns are (n = N^2, N >= 16), for the
// tiled numbering; 0 when the columns do not form one or FAIR_ART_TILED=0.
int32_t win_tile_side(const Csr &A) {
    if (!env_int("FAIR_ART_TILED", 1) || A.n < 256) return 0;
    int64_t N = (int64_t)std::sqrt((double)A.n);
    while (N * N > A.n) --N;
    while ((N + 1) * (N + 1) <= A.n) ++N;
    return N * N == A.n ? (int32_t)N : 0;
}

}  // namespace

// The window sweep over up to two chains (twin_art's x and y), each with its own sequence.  For a
// square image it runs on the tiled numbering: the renumbered columns are made once per matrix, the
// iterates are copied into it and back around the sweep.
void launch_art_window(Csr &A, int chains, const int32_t *const rows[2], const int64_t nsteps[2],
                       double *const x[2], const double *b, const double *nrA, double relax, int lwrb, double lb,
                       int upb, double ub, bool deferred) {
    const int slots[2] = {0, 1};
    if (!deferred) art_window_clear(A);
    const int32_t N = win_tile_side(A);
    if (N == 0) {
        for (int c = 0; c < chains; ++c) art_window_prepare(A, A.JA.p, A.n, c, rows[c], nsteps[c], b, nrA, stream());
        art_window_run(A, A.JA.p, A.n, chains, slots, x, relax, lwrb, lb, upb, ub, !deferred);
        return;
    }
    const int32_t NP = (N + 3) & ~3;
    const int64_t tn = (int64_t)NP * NP;
    FAIR_CHECK(tn < INT32_MAX, FAIR_ERR_ARG, "image too large for the window sweep");
    if (A.t_side != N) {
        const int64_t total = A.nnz + CSR_PAD;   // the spare entries too (column 0)
        A.JAt.alloc((size_t)total);
        hipLaunchKernelGGL(k_tile_cols, dim3(grid_for(total, 256)), dim3(256), 0, stream(), A.JA.p, A.JAt.p, total,
                           A.n, N, NP);
        launch_check("window tile columns");
        A.t_side = N;
    }
    ArtWindowBufs &wb = A.win;
    double *xt[2] = {nullptr, nullptr};
    for (int c = 0; c < chains; ++c) {
        if (wb.xt[c].n < (size_t)tn) wb.xt[c].alloc((size_t)tn);
        xt[c] = wb.xt[c].p;
        hipLaunchKernelGGL(k_x_to_tiles, dim3(grid_for(tn, 256)), dim3(256), 0, stream(), x[c], xt[c], N, NP);
        launch_check("window tiles in");
    }
    for (int c = 0; c < chains; ++c)
        art_window_prepare(A, A.JAt.p, (int32_t)tn, c, rows[c], nsteps[c], b, nrA, stream());
    art_window_run(A, A.JAt.p, (int32_t)tn, chains, slots, xt, relax, lwrb, lb, upb, ub, false);
    for (int c = 0; c < chains; ++c) {
        hipLaunchKernelGGL(k_x_from_tiles, dim3(grid_for((int64_t)N * N, 256)), dim3(256), 0, stream(), xt[c], x[c],
                           N, NP);
        launch_check("window tiles out");
    }
    if (!deferred) art_window_check(A);
}

}  // namespace fair
