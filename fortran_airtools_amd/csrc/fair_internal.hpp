// fair_internal.hpp -- shared pieces of the gfx950 library: status handling, device buffers,
// the device CSR object, wavefront reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fair_hip.h"

namespace fair {

constexpr int WAVE = 64;  // CDNA wavefront width

// ---- status ------------------------------------------------------------------------------
void set_error(const std::string &msg);

struct Error {
    int code;
    std::string msg;
};

#define FAIR_HIP(expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::fair::Error{FAIR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define FAIR_CHECK(cond, code, msg)                      \
    do {                                                 \
        if (!(cond)) throw ::fair::Error{(code), (msg)}; \
    } while (0)

// Runs f() and converts exceptions into status codes.
template <class F>
int guard(F &&f) {
    try {
        f();
        return FAIR_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return FAIR_ERR_HIP;
    }
}

hipStream_t stream();
void launch_check(const char *what);

// ---- device memory -------------------------------------------------------------------------
// Every device allocation of the library passes through here first.  FAIR_FAIL_ALLOC=N makes the
// N-th allocation of the process (counted from 1) fail as an out-of-memory hipMalloc would,
// before hipMalloc is called, so that failure paths can be tested, without a GPU too.
void before_device_alloc(size_t bytes);

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) {
            before_device_alloc(count * sizeof(T));
            FAIR_HIP(hipMalloc(&p, count * sizeof(T)));
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count) {
        if (count) FAIR_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
    }
    void download(T *h, size_t count) const {
        if (count) FAIR_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Host array whose elements are not initialised on allocation (a device copy fills them).  Large
// arrays are 2 MB aligned and marked for transparent huge pages: a device-to-host copy into fresh
// 4 KB pages spends most of its time on page faults (455 MB: 110 ms at 4 KB pages).
void *host_alloc_huge(size_t bytes);
void host_free_huge(void *p);

template <class T>
struct HostArray {
    struct Free {
        void operator()(T *q) const { host_free_huge(q); }
    };
    std::unique_ptr<T, Free> p;
    size_t n = 0;
    void alloc(size_t count) {
        p.reset(static_cast<T *>(host_alloc_huge(count * sizeof(T))));
        n = count;
    }
    void release() {
        p.reset();
        n = 0;
    }
    bool empty() const { return n == 0; }
    T *data() { return p.get(); }
};

// ---- device CSR ----------------------------------------------------------------------------
// Row offsets IA (m+1, int32, zero-based), column indices JA (int32, ZERO-based on the device;
// the host boundary is one-based), values vA (double).
// Spare entries after JA / vA: the SpMV's aligned pair loads may touch one entry past a row.
constexpr int CSR_PAD = 512;
// Longest contiguous column run one thread stages into LDS (blocked SpMV).
constexpr int SEG_CHUNK = 16;

// Layout of the LDS-staged SpMV (blocked.hip): non-empty rows grouped into blocks whose
// column footprint (per bucket of W consecutive columns, the [min,max] range touched) is
// loaded into LDS with coalesced segment loads; entries then address LDS through a 16-bit
// local index instead of gathering from global memory.
struct Blocked {
    bool ok = false;
    int32_t nblocks = 0;
    int32_t lds_elems = 0;      // footprint capacity per block (entries; the last slot holds 0)
    int32_t n_direct = 0;       // blocks whose footprint does not fit (global gathers)
    DBuf<int32_t> binfo;        // int4 per block {row_begin, row_end, seg_begin, seg_end}
    DBuf<int32_t> rows;         // int4 row descriptors {row, start, end, 0}, grouped by block
    DBuf<int32_t> segs;         // int2 {global column start, length << 16 | LDS offset}
    DBuf<uint16_t> loc;         // per entry: LDS slot of its column (parallel to JA)
};

// Level schedule of one Kaczmarz row sequence (methods.hip): step s gets level
// 1 + max(level of the last earlier step touching one of its pixels), so the steps of one level
// touch disjoint pixels and every pixel sees its updates in sequence order.  rows holds the
// steps sorted by (level, position), off the level offsets (nlev + 1).  Depends on the sparsity
// pattern only, which a Csr never changes after it is built.
// Per-call arguments of a captured level graph, read by its kernels from device memory.
struct ArtParams {
    double *x;
    const double *b, *nrA;
    double relax, lb, ub;
    int lwrb, upb;
};

struct LevelSched {
    std::vector<int32_t> seq;   // the one-based row sequence it was made for
    std::vector<int32_t> hoff;  // host copy of off
    std::vector<int32_t> hmaxlen;  // longest row of each level
    DBuf<int32_t> rows, off;
    DBuf<int32_t> desc;         // int4 per step in level order: {row (zero-based), IA(row), IA(row+1), 0}
    int32_t nlev = 0;
    // one kernel per level, captured once into a graph; its arguments come from params
    hipGraphExec_t graph = nullptr;
    DBuf<ArtParams> params;
    LevelSched() = default;
    LevelSched(const LevelSched &) = delete;
    LevelSched &operator=(const LevelSched &) = delete;
    ~LevelSched() {
        if (graph) (void)hipGraphExecDestroy(graph);
    }
};

// Buffers of the Kaczmarz window sweep (artwin.hip): per-step entry codes and descriptors of up
// to two chains, the pre-pass tables (G workgroups x n pixels each) and the error flag.
struct ArtWindowBufs {
    DBuf<uint16_t> codes[2];
    DBuf<int32_t> desc[2];     // int4 per step
    DBuf<double> bw[2];        // double2 per step
    DBuf<int32_t> cslot[2];    // 64 per step: the step's shared entries (LDS slot, value)
    DBuf<double> cval[2];
    DBuf<uint32_t> mask;
    DBuf<int32_t> slotof;
    DBuf<int32_t> err;
    DBuf<double> xt[2];        // the chains' iterates in 4 x 4 pixel tiles (see Csr::JAt)
    int64_t nsteps[2] = {0, 0};   // steps each buffer set was prepared for
    int G = 0;
    int64_t n = 0;
};

struct Csr {
    int32_t m = 0, n = 0;
    int64_t nnz = 0;
    DBuf<int32_t> IA, JA;
    DBuf<double> vA;
    DBuf<float> vA32;        // FP32 copy of the values (precision variants), made on demand
    DBuf<int32_t> rows_nz;   // compacted list of non-empty rows, int4 {row, start, end, 0}
    int32_t n_rows_nz = -1;  // -1: list not built
    DBuf<int32_t> rows_empty;  // ids of the empty rows
    int32_t n_rows_empty = 0;
    int32_t max_row = 0;     // longest row
    Blocked blk;             // LDS-staged layout (built for ray-traced matrices)
    DBuf<int32_t> spmm_rows; // rows_nz in the batched SpMM's tile order spmm_code (int4 each)
    int32_t spmm_code = -1;
    // Tiled layout of the batched SpMM (rows grouped in tiles; each tile's union of columns,
    // sorted, is staged through LDS in chunks): built for tile shape stile_code
    struct {
        int32_t code = -1, nblocks = 0;
        DBuf<int32_t> binfo;   // int4 per tile {row_begin, nrows, union_begin, union_len}
        DBuf<int32_t> rows;    // int4 per row {row, start, end, 0}
        DBuf<int32_t> uni;     // the tiles' union columns
        DBuf<uint16_t> slot;   // per entry: its column's position in its tile's union
    } stile;
    // Kaczmarz level schedules, least recently used first; shared so that a caller holding one
    // keeps it (and its graph and device buffers) alive when a later plan evicts it
    std::vector<std::shared_ptr<LevelSched>> lev;
    HostArray<int32_t> hIA, hJA;                    // host copy of the pattern for level planning
                                                    // (kept up to 64 M entries)
    ArtWindowBufs win;
    // The window sweep's column indices with the pixels of an N x N image renumbered in 4 x 4 tiles
    // (a tile's 16 pixels are consecutive, so a ray's consecutive pixels share cache lines); built on
    // first use for a square image, kept with the pattern it was made from
    DBuf<int32_t> JAt;
    int32_t t_side = 0;      // N of JAt (0: not built)
};

}  // namespace fair

struct fair_csr {
    fair::Csr A;        // the matrix as the caller sees it
    fair::Csr T;        // CSR of A^T, built on demand
    bool have_T = false;
    // The caller's rows are to be seen sorted by column (the reference's double transpose);
    // the device keeps its own entry order and sorts on download.
    bool rows_sorted = false;
    // Geometry of a ray-traced matrix (image side, rays per angle, angles); 0 when unknown.
    int32_t g_n = 0, g_p = 0, g_nth = 0;
    bool transposed = false;   // fair_transpose swapped the roles of A and T
    // Some row holds a column twice (possible only in a caller-made matrix; a transpose keeps
    // it).  Products sum such entries as the reference does; the Kaczmarz row updates, which
    // update every entry of a row from the same x, refuse such a matrix.
    bool dup_cols = false;
    bool blk_tuned = false;    // staged layouts chosen by fair_autotune_spmv (possibly none)
    // SIRT weights cached across calls (index: FAIR_CIMMINO / FAIR_SART) and loop workspace;
    // dropped whenever the matrix changes.
    fair::DBuf<double> M[2], D;
    bool w_ok[2] = {false, false};
    fair::DBuf<double> ws_r, ws_x;
    fair::DBuf<float> ws_r32, ws_x32, ws_rb, ws_xb;   // FP32 / batched workspaces
    bool graphs_used = false;
    // batched SpMM row orders (A, A^T): 0 = stored order, else a sinogram / image tile shape
    int spmm_tile[2] = {3, 6};
    // Forgets the weights and FP32 copies after the matrix changed, and releases the captured
    // batched-SIRT graphs, which hold pointers to them (defined in sparse.hip).
    void drop_cache();
};

namespace fair {

// ---- kernels shared across translation units (defined in sparse.hip) -----------------------
void build_row_list(Csr &A);
void alloc_entries(Csr &A, int64_t nnz);          // JA / vA with CSR_PAD zeroed spare entries
void build_transpose(const Csr &A, Csr &T);      // T = CSR(A^T), rows sorted by column
void sort_rows(Csr &A);                           // sort each row's entries by column
// Blocked layouts: rays grouped by DA x DJ sinogram tiles (columns bucketed by image column),
// or pixels grouped into TX x TY image tiles (columns bucketed by angle).
void build_blocked_rays(Csr &A, int32_t img_n, int32_t p, int32_t DA, int32_t DJ, int32_t lds_elems);
void build_blocked_tiles(Csr &T, int32_t img_n, int32_t p, int32_t nth, int32_t TX, int32_t TY,
                         int32_t lds_elems);
void ensure_blocked(fair_csr *h);
bool blocked_pending(const fair_csr *h);   // ensure_blocked would still build a layout
double *sirt_scratch_addr();               // fair_sirt_dev_noalloc's array on the current device
// SIRT weights of sirt (AIRMs.f90:1070-1083) cached in the handle (M[method], D for SART) and the
// device CSR of A^T; no LDS-staged layout, no workspaces.
void sirt_weights(fair_csr *h, int method);
void drop_graphs(const fair_csr *h);
// Refuses a matrix with a column held twice by one row (the Kaczmarz row updates).
inline void check_unique_columns(const fair_csr *h) {
    FAIR_CHECK(!h->dup_cols, FAIR_ERR_ARG,
               "a row holds the same column twice: the Kaczmarz row updates need unique columns per row");
}
void ensure_T(fair_csr *h);
void invalidate_T(fair_csr *h);

// Integer from the environment (dflt when unset or empty).
inline int env_int(const char *name, int dflt) {
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

// Kaczmarz window sweep (artwin.hip): the row steps of up to two chains (one-based rows on the
// device), each iterate updated in place, bit-identical to the one-wavefront sweep.  Synchronous.
// With deferred, the call neither clears nor reads the stop flag and does not synchronise: the
// caller clears it before a run of sweeps (art_window_clear) and checks it after (art_window_check).
bool art_window_usable(Csr &A, int64_t nsteps);   // FAIR_ART_WINDOW, rows up to 1024 entries
void launch_art_window(Csr &A, int chains, const int32_t *const rows[2], const int64_t nsteps[2],
                       double *const x[2], const double *b, const double *nrA, double relax, int lwrb, double lb,
                       int upb, double ub, bool deferred = false);
void art_window_clear(Csr &A);
void art_window_check(Csr &A);
// Its two halves: the pre-pass into buffer set `slot` (0 or 1) on stream s, and the sweep over
// prepared sets.  cols / ncols: the column indices the sweep runs on (A.JA and A.n, or the tiled
// numbering of launch_art_window) and the length of the iterates.
void art_window_prepare(Csr &A, const int32_t *cols, int32_t ncols, int slot, const int32_t *rows, int64_t nsteps,
                        const double *b, const double *nrA, hipStream_t s);
void art_window_run(Csr &A, const int32_t *cols, int32_t ncols, int chains, const int slots[2], double *const x[2],
                    double relax, int lwrb, double lb, int upb, double ub, bool check);

// y[r] = sum_k vA[k] * x[JA[k]] over every row; empty rows give 0.
void spmv(const Csr &A, const double *x, double *y);
void fill(double *y, int64_t n, double v);
int64_t exclusive_scan_i32(const int32_t *d_counts, int64_t m, int32_t *d_out);
int lds_per_block();
int lds_per_cu();   // LDS bytes of one compute unit
void row_norms(const Csr &A, int p, double *out);  // (sum |a|^p)^(1/p), pow-based

// Matrix-free projector support in raytrace.hip: the rays whose entries must come from the
// matrix build's own point rule (near-axis directions), with their in-range entries as a small
// CSR table (offs: nrays + 1 offsets; cols 0-based; vals FP32 lengths).
void mf_exact_rays(int32_t n, int32_t p, int32_t nth, const double *d_ang, const double *d_rayc,
                   std::vector<int32_t> &rays, std::vector<int32_t> &offs, std::vector<int32_t> &cols,
                   std::vector<float> &vals);

// ---- fixed-point sums (matrix-free adjoint) --------------------------------------------------
// Terms r * seg are scaled by S = 2^(ecap - E), where max |r| < 2^E is read from the bit
// pattern of max |r| (*maxbits), rounded to integers and summed in 64-bit two's complement:
// exact, so independent of the order of the adds.  ecap keeps one term below 2^51 and the
// whole sum below 2^62.
__device__ __forceinline__ double fix_scale(const unsigned *maxbits, int ecap) {
    const unsigned mb = *maxbits;
    const int E = (int)((mb >> 23) & 0xffu) - 126;   // M < 2^E (M normal or subnormal)
    return ldexp(1.0, ecap - E);
}

// round(v) as a two's-complement 64-bit integer, |v| < 2^51: adding 1.5 * 2^52 leaves the
// rounded integer in the low mantissa bits.
__device__ __forceinline__ unsigned long long fix_round(double v) {
    const double magic = 6755399441055744.0;
    return (unsigned long long)(__double_as_longlong(v + magic) - __double_as_longlong(magic));
}

// ---- wavefront reduction --------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// The same sum through DPP lane moves in the vector ALU instead of LDS permutes: within each row
// of 16 lanes, lane i adds lane i^1, then i^2, then the mirror within 8 and within 16 (every lane
// of the row then holds the row sum, bit-identical across the row since the adds pair the same
// operands), and the four row sums are read as scalars and added as (r0 + r1) + (r2 + r3).  All
// 64 lanes must be active.  The summation order differs from wave_sum's.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double lane_value(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <class T>
__device__ __forceinline__ T wave_sum_dpp(T v) {
    v += dpp_move<0xB1>(v);    // quad_perm [1,0,3,2]: lane i^1
    v += dpp_move<0x4E>(v);    // quad_perm [2,3,0,1]: lane i^2
    v += dpp_move<0x141>(v);   // row_half_mirror: the other quad of the 8
    v += dpp_move<0x140>(v);   // row_mirror: the other 8 of the 16
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// Fortran MIN / MAX on reals as the reference compiler evaluates them:
// min(a,b) = a<b ? a : b, max(a,b) = a>b ? a : b (a NaN `a` yields b).
__host__ __device__ __forceinline__ double fmin_ref(double a, double b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ double fmax_ref(double a, double b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ float fmin_ref(float a, float b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ float fmax_ref(float a, float b) { return a > b ? a : b; }

inline unsigned grid_for(int64_t items, int per_block) {
    int64_t g = (items + per_block - 1) / per_block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace fair
