// sparse.hip -- device CSR operations behind the sparse_matrices call surface.
//
// Replaces sparse_matrices.f90: MMRS / csr_matrowvec (A*x, :252-298, 349-368), MMCS
// (A^T*y, :300-347, 370-393), reorder / transposition (:395-411, 492-570), compute_row_norms
// (:740-780), add_mult_row (:795-822) and scalar_mult (:824-831).
//
// Layout in HBM: IA int32[m+1], JA int32[nnz] (zero-based), vA double[nnz]; a compacted list
// of non-empty rows as int4 {row, start, end, 0} so that a wavefront finds its row extent in
// one load (for odd detector counts the fan geometry leaves about half the rows empty).
// A^T*y is a gather over a device-built CSR of A^T instead of the reference's racy scatter:
// deterministic, and every thread owns its output.
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>

#include "fair_internal.hpp"
#include "spmv.hpp"

namespace fair {

int64_t exclusive_scan_i32(const int32_t *d_counts, int64_t m, int32_t *d_out);

// ---- runtime -------------------------------------------------------------------------------
namespace {
// (error text is per host thread; the stream is process-wide)
thread_local std::string g_last_error;
hipStream_t g_stream = nullptr;
int g_lds_per_block = 65536;
int g_lds_per_cu = 65536;
}  // namespace

void set_error(const std::string &msg) { g_last_error = msg; }

void before_device_alloc(size_t bytes) {
    static std::atomic<long long> count{0};
    const long long k = ++count;
    const char *v = getenv("FAIR_FAIL_ALLOC");
    if (v && *v && atoll(v) == k)
        throw Error{FAIR_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + " bytes): out of memory (allocation " +
                                      std::to_string(k) + " failed by FAIR_FAIL_ALLOC)"};
}
void *host_alloc_huge(size_t bytes) {
    constexpr size_t HUGE_PAGE = size_t(2) << 20;
    void *q = nullptr;
    if (bytes >= HUGE_PAGE) {
        const size_t rounded = (bytes + HUGE_PAGE - 1) / HUGE_PAGE * HUGE_PAGE;
        if (posix_memalign(&q, HUGE_PAGE, rounded) != 0) q = nullptr;
        if (q) (void)madvise(q, rounded, MADV_HUGEPAGE);   // advice only: ignored where unsupported
    } else {
        q = malloc(bytes > 0 ? bytes : 1);
    }
    if (!q) throw std::bad_alloc();
    return q;
}

void host_free_huge(void *p) { free(p); }

hipStream_t stream() { return g_stream; }
int lds_per_block() { return g_lds_per_block; }
int lds_per_cu() { return g_lds_per_cu; }

void launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error{FAIR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

// ---- small kernels --------------------------------------------------------------------------
namespace {

__global__ void k_row_list(const int32_t *IA, int32_t m, int4 *info, int32_t *cnt, int32_t *maxlen) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int32_t s = IA[r], e = IA[r + 1];
    if (e > s) {
        // order-preserving compaction is done by the scan-based kernel below; here count only
        atomicMax(maxlen, e - s);
    }
    (void)info;
    (void)cnt;
}

__global__ void k_nonempty_flags(const int32_t *IA, int32_t m, int32_t *flag) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) flag[r] = IA[r + 1] > IA[r] ? 1 : 0;
}

__global__ void k_compact_rows(const int32_t *IA, int32_t m, const int32_t *pos, int4 *info,
                               int32_t *empty) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int32_t s = IA[r], e = IA[r + 1];
    if (e > s) info[pos[r]] = make_int4(r, s, e, 0);
    else empty[r - pos[r]] = r;
}

__global__ void k_fill_f64(double *y, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

__global__ void k_add_i32(const int32_t *a, int32_t *b, int64_t n, int32_t d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i] + d;
}

__global__ void k_scale(double *v, int64_t n, double c) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = c * v[i];
}

// Column histogram; indices outside [0, ncols) are dropped from the transpose.
__global__ void k_col_hist(const int32_t *JA, int64_t nnz, int32_t ncols, int32_t *cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) {
        int32_t c = JA[i];
        if ((uint32_t)c < (uint32_t)ncols) atomicAdd(&cnt[c], 1);
    }
}

// Scatter of the counting-sort transpose: one wavefront per row; positions within a target
// row come from atomics (arbitrary order) and are put in ascending source-row order by the
// segmented sort that follows.
__global__ void __launch_bounds__(256) k_transpose_scatter(const int4 *rows, int32_t nrows,
                                                           const int32_t *JA, const double *vA,
                                                           int32_t ncols, int32_t *cursor,
                                                           int32_t *JB, double *vB) {
    int w = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    int lane = threadIdx.x & (WAVE - 1);
    if (w >= nrows) return;
    int4 ri = rows[w];
    for (int k = ri.y + lane; k < ri.z; k += WAVE) {
        int32_t c = JA[k];
        if ((uint32_t)c >= (uint32_t)ncols) continue;
        int32_t pos = atomicAdd(&cursor[c], 1);
        JB[pos] = ri.x;
        vB[pos] = vA[k];
    }
}

// Bitonic sort of one segment (keys unique within the segment) held at key/val with
// capacity P (power of two, padded with INT32_MAX keys).
__device__ void bitonic(int32_t *key, double *val, int P) {
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    bool up = (i & k) == 0;
                    int32_t a = key[i], b = key[ixj];
                    if ((a > b) == up) {
                        key[i] = b;
                        key[ixj] = a;
                        double t = val[i];
                        val[i] = val[ixj];
                        val[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ int pow2ceil(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Segmented sort in LDS: one block per row of length <= cap.
__global__ void __launch_bounds__(256) k_segsort_lds(const int4 *rows, int32_t nrows, int32_t *JA,
                                                     double *vA, int cap) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *sval = (double *)smem;
    int32_t *skey = (int32_t *)(sval + cap);
    int4 ri = rows[blockIdx.x];
    int len = ri.z - ri.y;
    if (len < 2 || len > cap) return;
    int P = pow2ceil(len);
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        if (i < len) {
            skey[i] = JA[ri.y + i];
            sval[i] = vA[ri.y + i];
        } else {
            skey[i] = INT32_MAX;
            sval[i] = 0.0;
        }
    }
    __syncthreads();
    bitonic(skey, sval, P);
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        JA[ri.y + i] = skey[i];
        vA[ri.y + i] = sval[i];
    }
}

// Long rows: the same network on global scratch (one block per long row).
__global__ void __launch_bounds__(256) k_segsort_global(const int4 *rows, const int32_t *ids,
                                                        int32_t nlong, int32_t *JA, double *vA,
                                                        int32_t *kscr, double *vscr, int64_t cap) {
    int q = blockIdx.x;
    if (q >= nlong) return;
    int4 ri = rows[ids[q]];
    int len = ri.z - ri.y;
    int P = pow2ceil(len);
    int32_t *key = kscr + (int64_t)q * cap;
    double *val = vscr + (int64_t)q * cap;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        key[i] = i < len ? JA[ri.y + i] : INT32_MAX;
        val[i] = i < len ? vA[ri.y + i] : 0.0;
    }
    __syncthreads();
    bitonic(key, val, P);
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        JA[ri.y + i] = key[i];
        vA[ri.y + i] = val[i];
    }
}

__global__ void k_long_rows(const int4 *rows, int32_t nrows, int cap, int32_t *ids, int32_t *n) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nrows) return;
    int4 ri = rows[w];
    if (ri.z - ri.y > cap) ids[atomicAdd(n, 1)] = w;
}

// compiler-rt's __powidf2 (the reference's abs(v)**p with integer p): repeated squaring.
__device__ __forceinline__ double powi_ref(double a, int b) {
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return r;
}

__global__ void __launch_bounds__(256) k_row_norms(const int4 *rows, int32_t nrows,
                                                   const double *vA, int p, double *out) {
    int w = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    int lane = threadIdx.x & (WAVE - 1);
    if (w >= nrows) return;
    int4 ri = rows[w];
    double s = 0.0;
    for (int k = ri.y + lane; k < ri.z; k += WAVE) s += powi_ref(fabs(vA[k]), p);
    s = wave_sum(s);
    if (lane == 0) out[ri.x] = p == 1 ? s : (p == 2 ? sqrt(s) : pow(s, 1.0 / (double)p));
}

__global__ void __launch_bounds__(64) k_row_dot(const int32_t *IA, const int32_t *JA,
                                                const double *vA, int32_t k, const double *x,
                                                int32_t ncols, double *f) {
    int lane = threadIdx.x;
    double s = 0.0;
    for (int q = IA[k] + lane; q < IA[k + 1]; q += WAVE)
        if ((uint32_t)JA[q] < (uint32_t)ncols) s += x[JA[q]] * vA[q];
    s = wave_sum(s);
    if (lane == 0) *f = s;
}

// add_mult_row (sparse_matrices.f90:795-822): column indices are unique within a row, so
// every lane owns its entries; max(lb) then min(ub) on the touched entries only.
__global__ void __launch_bounds__(64) k_add_row(const int32_t *IA, const int32_t *JA,
                                                const double *vA, int32_t k, double c, double *w,
                                                int32_t ncols, int lwrb, double lb, int upb,
                                                double ub) {
    for (int q = IA[k] + threadIdx.x; q < IA[k + 1]; q += WAVE) {
        int j = JA[q];
        if ((uint32_t)j >= (uint32_t)ncols) continue;
        double v = w[j] + c * vA[q];
        if (lwrb) v = fmax_ref(v, lb);
        if (upb) v = fmin_ref(v, ub);
        w[j] = v;
    }
}

}  // namespace

void fill(double *y, int64_t n, double v) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256)), dim3(256), 0, stream(), y, n, v);
}

void alloc_entries(Csr &A, int64_t nnz) {
    A.JA.alloc((size_t)nnz + CSR_PAD);
    A.vA.alloc((size_t)nnz + CSR_PAD);
    FAIR_HIP(hipMemsetAsync(A.JA.p + nnz, 0, sizeof(int32_t) * CSR_PAD, stream()));
    FAIR_HIP(hipMemsetAsync(A.vA.p + nnz, 0, sizeof(double) * CSR_PAD, stream()));
}

void build_row_list(Csr &A) {
    int32_t m = A.m;
    DBuf<int32_t> flag(m > 0 ? m : 1), pos((size_t)m + 1), mx(1);
    FAIR_HIP(hipMemsetAsync(mx.p, 0, sizeof(int32_t), stream()));
    if (m > 0) {
        hipLaunchKernelGGL(k_nonempty_flags, dim3(grid_for(m, 256)), dim3(256), 0, stream(), A.IA.p, m, flag.p);
        hipLaunchKernelGGL(k_row_list, dim3(grid_for(m, 256)), dim3(256), 0, stream(), A.IA.p, m,
                           (int4 *)nullptr, (int32_t *)nullptr, mx.p);
    }
    int64_t cnt = exclusive_scan_i32(flag.p, m, pos.p);
    A.rows_nz.alloc(4 * (size_t)(cnt > 0 ? cnt : 1));
    A.rows_empty.alloc(m - cnt > 0 ? m - cnt : 1);
    if (m > 0)
        hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(m, 256)), dim3(256), 0, stream(), A.IA.p, m,
                           pos.p, (int4 *)A.rows_nz.p, A.rows_empty.p);
    launch_check("row list");
    int32_t hmx = 0;
    mx.download(&hmx, 1);
    FAIR_HIP(hipStreamSynchronize(stream()));
    A.n_rows_nz = (int32_t)cnt;
    A.n_rows_empty = (int32_t)(m - cnt);
    A.max_row = hmx;
    A.spmm_rows.release();   // the batched SpMM's reordered copy of the list
    A.spmm_code = -1;
    A.stile.code = -1;       // and its tiled layout
    A.stile.nblocks = 0;
}

// Sort the entries of every listed row by column index (values follow).
static void segsort_arrays(const int4 *rows, int32_t nrows, int32_t max_row, int32_t *JA, double *vA) {
    if (nrows <= 0 || max_row < 2) return;
    int cap = 1;
    while (cap * 2 * (int)(sizeof(double) + sizeof(int32_t)) <= lds_per_block() && cap < 8192) cap *= 2;
    size_t shm = (size_t)cap * (sizeof(double) + sizeof(int32_t));
    hipLaunchKernelGGL(k_segsort_lds, dim3(nrows), dim3(256), shm, stream(), rows, nrows, JA, vA, cap);
    launch_check("segmented sort");
    if (max_row > cap) {
        DBuf<int32_t> ids(nrows), nl(1);
        FAIR_HIP(hipMemsetAsync(nl.p, 0, sizeof(int32_t), stream()));
        hipLaunchKernelGGL(k_long_rows, dim3(grid_for(nrows, 256)), dim3(256), 0, stream(), rows, nrows, cap,
                           ids.p, nl.p);
        int32_t nlong = 0;
        nl.download(&nlong, 1);
        FAIR_HIP(hipStreamSynchronize(stream()));
        int64_t P = 1;
        while (P < max_row) P <<= 1;
        // bound the scratch: process long rows in batches
        const int32_t batch = (int32_t)std::max<int64_t>(1, std::min<int64_t>(nlong, (int64_t)(1 << 28) / P));
        DBuf<int32_t> ks((size_t)batch * P);
        DBuf<double> vs((size_t)batch * P);
        for (int32_t b0 = 0; b0 < nlong; b0 += batch) {
            int32_t nb = std::min(batch, nlong - b0);
            hipLaunchKernelGGL(k_segsort_global, dim3(nb), dim3(256), 0, stream(), rows, ids.p + b0, nb, JA, vA,
                               ks.p, vs.p, P);
        }
        launch_check("segmented sort (long rows)");
        FAIR_HIP(hipStreamSynchronize(stream()));
    }
}

static void segsort_rows(Csr &A) {
    if (A.n_rows_nz < 0) build_row_list(A);
    segsort_arrays((const int4 *)A.rows_nz.p, A.n_rows_nz, A.max_row, A.JA.p, A.vA.p);
}

void sort_rows(Csr &A) { segsort_rows(A); }

void build_transpose(const Csr &A, Csr &T) {
    FAIR_CHECK(A.n_rows_nz >= 0, FAIR_ERR_INTERNAL, "row list missing");
    T.m = A.n;
    T.n = A.m;
    T.nnz = A.nnz;
    DBuf<int32_t> cnt(A.n > 0 ? A.n : 1);
    FAIR_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), stream()));
    if (A.nnz > 0)
        hipLaunchKernelGGL(k_col_hist, dim3(grid_for(A.nnz, 256)), dim3(256), 0, stream(), A.JA.p, A.nnz,
                           A.n, cnt.p);
    T.IA.alloc((size_t)T.m + 1);
    T.nnz = exclusive_scan_i32(cnt.p, T.m, T.IA.p);
    alloc_entries(T, T.nnz);
    if (T.m > 0) FAIR_HIP(hipMemcpyAsync(cnt.p, T.IA.p, sizeof(int32_t) * T.m, hipMemcpyDeviceToDevice, stream()));
    if (A.n_rows_nz > 0)
        hipLaunchKernelGGL(k_transpose_scatter, dim3(grid_for((int64_t)A.n_rows_nz * WAVE, 256)), dim3(256), 0,
                           stream(), (const int4 *)A.rows_nz.p, A.n_rows_nz, A.JA.p, A.vA.p, A.n,
                           cnt.p, T.JA.p, T.vA.p);
    launch_check("transpose scatter");
    T.n_rows_nz = -1;
    build_row_list(T);
    segsort_rows(T);
}

void ensure_T(fair_csr *h) {
    if (h->have_T) return;
    if (h->A.n_rows_nz < 0) build_row_list(h->A);
    build_transpose(h->A, h->T);
    h->have_T = true;
}

void invalidate_T(fair_csr *h) {
    if (h->have_T) {
        h->T = Csr();
        h->have_T = false;
        h->blk_tuned = false;
    }
}

void spmv(const Csr &A, const double *x, double *y) {
    fill(y, A.m, 0.0);
    spmv_launch(A, x, EpiStore<double>{y});
}

void row_norms(const Csr &A, int p, double *out) {
    fill(out, A.m, 0.0);
    if (A.n_rows_nz > 0)
        hipLaunchKernelGGL(k_row_norms, dim3(grid_for((int64_t)A.n_rows_nz * WAVE, 256)), dim3(256), 0,
                           stream(), (const int4 *)A.rows_nz.p, A.n_rows_nz, A.vA.p, p, out);
    launch_check("row norms");
}

}  // namespace fair

using namespace fair;

namespace {

void copy_in(double *d, const double *h, size_t n) {
    if (n) FAIR_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, stream()));
}
void copy_out(double *h, const double *d, size_t n) {
    if (n) FAIR_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, stream()));
    FAIR_HIP(hipStreamSynchronize(stream()));
}

void check_handle(const fair_csr *A) { FAIR_CHECK(A != nullptr, FAIR_ERR_ARG, "null matrix handle"); }

}  // namespace

extern "C" const char *fair_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *fair_version(void) { return "fair-mi355x 0.1 (gfx950)"; }

extern "C" int fair_init(int device) {
    return guard([&] {
        int count = 0;
        FAIR_HIP(hipGetDeviceCount(&count));
        FAIR_CHECK(device >= 0 && device < count, FAIR_ERR_ARG, "no such device");
        FAIR_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        FAIR_HIP(hipGetDeviceProperties(&prop, device));
        g_lds_per_block = (int)prop.sharedMemPerBlock;
        g_lds_per_cu = prop.maxSharedMemoryPerMultiProcessor > 0 ? (int)prop.maxSharedMemoryPerMultiProcessor
                                                                 : g_lds_per_block;
        sirt_scratch_addr();   // the no-allocation SIRT's array on this device
    });
}

extern "C" int fair_set_stream(void *s) {
    return guard([&] { g_stream = (hipStream_t)s; });
}

extern "C" int fair_synchronize(void) {
    return guard([&] { FAIR_HIP(hipStreamSynchronize(stream())); });
}

extern "C" int fair_csr_create(int32_t m, int32_t n, int64_t nnz, const int32_t *IA,
                               const int32_t *JA, const double *vA, fair_csr **out) {
    return guard([&] {
        FAIR_CHECK(out && IA && (nnz == 0 || (JA && vA)), FAIR_ERR_ARG, "null argument");
        FAIR_CHECK(m >= 0 && n >= 0 && nnz >= 0 && nnz <= INT32_MAX, FAIR_ERR_ARG, "bad sizes");
        FAIR_CHECK(IA[0] == 0 && IA[m] == nnz, FAIR_ERR_ARG, "IA(0) must be 0 and IA(m) = nnz");
        for (int32_t r = 0; r < m; ++r)
            FAIR_CHECK(IA[r + 1] >= IA[r], FAIR_ERR_ARG, "IA must be non-decreasing");
        for (int64_t k = 0; k < nnz; ++k)
            FAIR_CHECK(JA[k] >= 1 && JA[k] <= n, FAIR_ERR_ARG, "column index out of range");
        // a column held twice by one row: mark[c] is the last row (one-based) that held c
        bool dup = false;
        if (nnz > 0) {
            std::vector<int32_t> mark((size_t)n + 1, 0);
            for (int32_t r = 0; r < m && !dup; ++r)
                for (int32_t k = IA[r]; k < IA[r + 1]; ++k) {
                    if (mark[JA[k]] == r + 1) {
                        dup = true;
                        break;
                    }
                    mark[JA[k]] = r + 1;
                }
        }
        auto *h = new fair_csr();
        h->dup_cols = dup;
        try {
            Csr &A = h->A;
            A.m = m;
            A.n = n;
            A.nnz = nnz;
            A.IA.alloc((size_t)m + 1);
            A.IA.upload(IA, (size_t)m + 1);
            alloc_entries(A, nnz);
            if (nnz > 0) {
                DBuf<int32_t> tmp(nnz);
                tmp.upload(JA, nnz);
                hipLaunchKernelGGL(k_add_i32, dim3(grid_for(nnz, 256)), dim3(256), 0, stream(), tmp.p,
                                   A.JA.p, nnz, -1);
                A.vA.upload(vA, nnz);
                FAIR_HIP(hipStreamSynchronize(stream()));
            }
            build_row_list(A);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

extern "C" int fair_csr_info(const fair_csr *A, int32_t *m, int32_t *n, int64_t *nnz) {
    return guard([&] {
        check_handle(A);
        if (m) *m = A->A.m;
        if (n) *n = A->A.n;
        if (nnz) *nnz = A->A.nnz;
    });
}

extern "C" int fair_csr_download(const fair_csr *h, int32_t *IA, int32_t *JA, double *vA) {
    return guard([&] {
        check_handle(h);
        const Csr &A = h->A;
        if (IA) A.IA.download(IA, (size_t)A.m + 1);
        if ((JA || vA) && A.nnz > 0) {
            // the device keeps its own entry order; the caller sees the reference's
            DBuf<int32_t> tj(A.nnz);
            DBuf<double> tv;
            const double *vsrc = A.vA.p;
            hipLaunchKernelGGL(k_add_i32, dim3(grid_for(A.nnz, 256)), dim3(256), 0, stream(), A.JA.p,
                               tj.p, A.nnz, 1);
            if (h->rows_sorted && A.max_row >= 2) {
                tv.alloc(A.nnz);
                FAIR_HIP(hipMemcpyAsync(tv.p, A.vA.p, sizeof(double) * A.nnz, hipMemcpyDeviceToDevice, stream()));
                segsort_arrays((const int4 *)A.rows_nz.p, A.n_rows_nz, A.max_row, tj.p, tv.p);
                vsrc = tv.p;
            }
            if (JA) tj.download(JA, A.nnz);
            if (vA) FAIR_HIP(hipMemcpyAsync(vA, vsrc, sizeof(double) * A.nnz, hipMemcpyDeviceToHost, stream()));
        }
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_csr_device_arrays(const fair_csr *h, const int32_t **IA, const int32_t **JA0,
                                      const double **vA) {
    return guard([&] {
        check_handle(h);
        if (IA) *IA = h->A.IA.p;
        if (JA0) *JA0 = h->A.JA.p;
        if (vA) *vA = h->A.vA.p;
    });
}

void fair_csr::drop_cache() {
    w_ok[0] = w_ok[1] = false;
    A.vA32.release();
    T.vA32.release();
    if (graphs_used) fair::drop_graphs(this);
}

extern "C" int fair_csr_free(fair_csr *h) {
    return guard([&] {
        if (h) {
            (void)hipStreamSynchronize(stream());
            if (h->graphs_used) drop_graphs(h);
            delete h;
        }
    });
}

extern "C" int fair_matvecmul_dev(fair_csr *h, const double *d_v, double *d_w) {
    return guard([&] {
        check_handle(h);
        ensure_blocked(h);
        spmv(h->A, d_v, d_w);
    });
}

extern "C" int fair_spmv_rows_dev(fair_csr *h, const double *d_v, double *d_w) {
    return guard([&] {
        check_handle(h);
        ensure_blocked(h);
        spmv_launch(h->A, d_v, EpiStore<double>{d_w});
    });
}

extern "C" int fair_transvecmul_dev(fair_csr *h, const double *d_v, double *d_w) {
    return guard([&] {
        check_handle(h);
        ensure_blocked(h);
        ensure_T(h);
        spmv(h->T, d_v, d_w);
    });
}

extern "C" int fair_matvecmul(fair_csr *h, const double *v, int32_t nv, double *w, int32_t nw) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(nv == h->A.n, FAIR_ERR_ARG, "Vector input size incorrect.");
        FAIR_CHECK(nw == h->A.m, FAIR_ERR_ARG, "Output size incorrect.");
        ensure_blocked(h);
        DBuf<double> dv(nv > 0 ? nv : 1), dw(nw > 0 ? nw : 1);
        copy_in(dv.p, v, nv);
        spmv(h->A, dv.p, dw.p);
        copy_out(w, dw.p, nw);
    });
}

extern "C" int fair_transvecmul(fair_csr *h, const double *v, int32_t nv, double *w, int32_t nw) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(nv == h->A.m, FAIR_ERR_ARG, "Vector input size incorrect.");
        FAIR_CHECK(nw == h->A.n, FAIR_ERR_ARG, "Output size incorrect.");
        ensure_blocked(h);
        ensure_T(h);
        DBuf<double> dv(nv > 0 ? nv : 1), dw(nw > 0 ? nw : 1);
        copy_in(dv.p, v, nv);
        spmv(h->T, dv.p, dw.p);
        copy_out(w, dw.p, nw);
    });
}

extern "C" int fair_rowvecmult(fair_csr *h, int32_t k, const double *v, int32_t nv, double *f) {
    return guard([&] {
        check_handle(h);
        check_unique_columns(h);
        FAIR_CHECK(k >= 1 && k <= h->A.m, FAIR_ERR_ARG, "row index out of range");
        FAIR_CHECK(nv == h->A.n, FAIR_ERR_ARG, "Vector input size incorrect.");
        DBuf<double> dv(nv > 0 ? nv : 1), df(1);
        copy_in(dv.p, v, nv);
        hipLaunchKernelGGL(k_row_dot, dim3(1), dim3(64), 0, stream(), h->A.IA.p, h->A.JA.p, h->A.vA.p,
                           k - 1, dv.p, h->A.n, df.p);
        launch_check("rowvecmult");
        copy_out(f, df.p, 1);
    });
}

extern "C" int fair_row_norms(fair_csr *h, int32_t p, double *nrA, int32_t nm) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(p >= 1, FAIR_ERR_ARG, "p must be a positive integer");
        FAIR_CHECK(nm == h->A.m, FAIR_ERR_ARG, "Output size incorrect.");
        DBuf<double> d(nm > 0 ? nm : 1);
        row_norms(h->A, p, d.p);
        copy_out(nrA, d.p, nm);
    });
}

extern "C" int fair_column_norms(fair_csr *h, int32_t p, double *ncA, int32_t nn) {
    return guard([&] {
        check_handle(h);
        FAIR_CHECK(p >= 1, FAIR_ERR_ARG, "p must be a positive integer");
        FAIR_CHECK(nn == h->A.n, FAIR_ERR_ARG, "Output size incorrect.");
        ensure_T(h);
        DBuf<double> d(nn > 0 ? nn : 1);
        row_norms(h->T, p, d.p);
        copy_out(ncA, d.p, nn);
    });
}

extern "C" int fair_transpose(fair_csr *h) {
    return guard([&] {
        check_handle(h);
        ensure_T(h);
        // the transposed matrix becomes the caller's matrix (its rows are in counting-sort
        // order); transposing back shows the old A with rows sorted by column, which the
        // reference's double transpose leaves -- downloads sort, the device order stays
        Csr old = std::move(h->A);
        h->A = std::move(h->T);
        h->T = std::move(old);
        h->have_T = true;
        h->rows_sorted = true;
        h->transposed = !h->transposed;
        h->drop_cache();
    });
}

extern "C" int fair_multiply_by_scalar(fair_csr *h, double c) {
    return guard([&] {
        check_handle(h);
        if (h->A.nnz > 0)
            hipLaunchKernelGGL(k_scale, dim3(grid_for(h->A.nnz, 256)), dim3(256), 0, stream(), h->A.vA.p,
                               h->A.nnz, c);
        if (h->have_T && h->T.nnz > 0)
            hipLaunchKernelGGL(k_scale, dim3(grid_for(h->T.nnz, 256)), dim3(256), 0, stream(), h->T.vA.p,
                               h->T.nnz, c);
        launch_check("multiply_by_scalar");
        h->drop_cache();
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}

extern "C" int fair_add_scalar_mult_row(fair_csr *h, int32_t k, double c, double *w, int32_t nw,
                                        const double *lwrb, const double *upb) {
    return guard([&] {
        check_handle(h);
        check_unique_columns(h);
        FAIR_CHECK(k >= 1 && k <= h->A.m, FAIR_ERR_ARG, "row index out of range");
        FAIR_CHECK(nw == h->A.n, FAIR_ERR_ARG, "Vector input size incorrect.");
        DBuf<double> dw(nw > 0 ? nw : 1);
        copy_in(dw.p, w, nw);
        hipLaunchKernelGGL(k_add_row, dim3(1), dim3(64), 0, stream(), h->A.IA.p, h->A.JA.p, h->A.vA.p,
                           k - 1, c, dw.p, h->A.n, lwrb != nullptr, lwrb ? *lwrb : 0.0,
                           upb != nullptr, upb ? *upb : 0.0);
        launch_check("add_scalar_mult_row");
        copy_out(w, dw.p, nw);
    });
}
