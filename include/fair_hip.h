/*
 * fair_hip.h -- C ABI of the MI355X (gfx950) FAIR tools reconstruction path.
 *
 * One process, one GPU.  Every entry point returns a status (FAIR_OK = 0); the text of the
 * last error of the calling thread is available from fair_last_error().  Host arrays follow
 * the reference's conventions: IA holds m+1 zero-based row offsets, JA holds one-based
 * column numbers, vectors are contiguous double (or float for the _f32 variants).
 * Functions named *_dev take device pointers (inputs already resident in HBM) and run on the
 * stream set by fair_set_stream (default: the null stream); all others take host arrays and
 * return after the result has been copied back.
 *
 * Each declaration cites the reference procedure it replaces (file:line in FAIR tools,
 * BartvLith/fortran_AIRtools).
 */
#ifndef FAIR_HIP_H
#define FAIR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FAIR_OK = 0,
    FAIR_ERR_ARG = 1,       /* bad argument or size mismatch ("Vector input size incorrect.") */
    FAIR_ERR_HIP = 2,       /* HIP runtime failure (allocation, launch, copy) */
    FAIR_ERR_OVERFLOW = 3,  /* nnz or an offset does not fit the reference's int32 */
    FAIR_ERR_INTERNAL = 4,  /* consistency check failed ("++ Something went wrong") */
    FAIR_ERR_METHOD = 5     /* unknown method name / invalid method arguments */
};

/* Geometry record, bind(C) mirror of `type(data)` (system_generation.f90:6-12) without the
 * allocatable arrays; theta (radians) is passed separately. */
typedef struct fair_geom {
    int32_t n, p, nth;
    double r, dw, sd;
} fair_geom;

/* Method options, bind(C) mirror of `type(opts)` (AIRMs.f90:6-13). */
typedef struct fair_opts {
    double relaxpar, ubound, lbound, damp;
    int32_t verbose, upb, lwrb, rowprobability;
} fair_opts;

typedef struct fair_csr fair_csr; /* opaque device CSR matrix (+ cached CSR of its transpose) */

/* ---- runtime ------------------------------------------------------------------------ */
const char *fair_last_error(void);
/* Select the device and prepare the library (no counterpart: the reference has no device). */
int fair_init(int device);
/* Stream for all subsequent launches (hipStream_t passed as void*; NULL = null stream). */
int fair_set_stream(void *stream);
int fair_synchronize(void);
/* Library version string. */
const char *fair_version(void);

/* ---- system_generation -------------------------------------------------------------- */
/* fanlineartomo, system_generation.f90:79-307 (count -> scan -> fill on the device). */
int fair_fanlineartomo(const fair_geom *g, const double *theta, fair_csr **A);
/* paralleltomo, system_generation.f90:309-492. */
int fair_paralleltomo(const fair_geom *g, const double *theta, fair_csr **A);
/* linspace, system_generation.f90:17-35 (host). */
int fair_linspace(double a, double b, int32_t n, double *x);

/* ---- sparse_matrices: construction and transfer (type(sparse_matrix), :33-176) -------- */
/* Upload a host CSR (the components m, n, nnz, vA, IA(0:m), JA of sparse_matrices.f90:47-49). */
int fair_csr_create(int32_t m, int32_t n, int64_t nnz, const int32_t *IA, const int32_t *JA,
                    const double *vA, fair_csr **A);
int fair_csr_info(const fair_csr *A, int32_t *m, int32_t *n, int64_t *nnz);
/* Download the components (JA one-based). */
int fair_csr_download(const fair_csr *A, int32_t *IA, int32_t *JA, double *vA);
/* Device addresses of IA (m+1, int32), zero-based JA (int32) and vA (double). */
int fair_csr_device_arrays(const fair_csr *A, const int32_t **IA, const int32_t **JA0,
                           const double **vA);
int fair_csr_free(fair_csr *A);

/* ---- matrix dumps (host only) --------------------------------------------------------------
 * The Fortran unformatted records of store_matrix (system_generation.f90:595-625): m, n, nnz, vA,
 * JA, IA (nia = m+1 for CSR, n+1 for CSC), readable by read_csr_matrix / read_csc_matrix
 * (:627-707) and written by the reference readable here. */
int fair_matrix_store(const char *path, int32_t m, int32_t n, int64_t nnz, const double *vA,
                      const int32_t *JA, const int32_t *IA, int32_t nia);
int fair_matrix_read_header(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int32_t *nia);
int fair_matrix_read(const char *path, int64_t nnz, int32_t nia, double *vA, int32_t *JA, int32_t *IA);

/* ---- measured data -------------------------------------------------------------------------
 * The file mode of main.f90: the angle cull of main.f90:229-245 (keep angles 1, 1+skip, ... of
 * the (p, nth) transmission data T; skip = 0 keeps all) and b = -log(T) (main.f90:270), on the
 * device.  b holds ceil(nth/skip)*p values (nth*p for skip = 0). */
int fair_absorbance(const double *T, int64_t p, int64_t nth, int64_t skip, double *b);
int fair_absorbance_dev(const double *T, int64_t p, int64_t nth, int64_t skip, double *b);

/* ---- sparse_matrices: operations ---------------------------------------------------- */
/* matvecmul -> sparsematmul/MMRS, sparse_matrices.f90:272-298, 349-368. */
int fair_matvecmul(fair_csr *A, const double *v, int32_t nv, double *w, int32_t nw);
/* transvecmul -> sparsetransmul/MMCS, sparse_matrices.f90:300-347, 370-393 (gather over the
 * device-built CSR of A^T: deterministic, no scatter race). */
int fair_transvecmul(fair_csr *A, const double *v, int32_t nv, double *w, int32_t nw);
/* rowvecmult -> csr_matrowvec, sparse_matrices.f90:252-270 (k one-based). */
int fair_rowvecmult(fair_csr *A, int32_t k, const double *v, int32_t nv, double *f);
/* row_norms -> compute_row_norms, sparse_matrices.f90:740-780 (p >= 1). */
int fair_row_norms(fair_csr *A, int32_t p, double *nrA, int32_t nm);
/* column_norms -> compute_column_norms, sparse_matrices.f90:782-793. */
int fair_column_norms(fair_csr *A, int32_t p, double *ncA, int32_t nn);
/* transpose -> transposition/reorder, sparse_matrices.f90:395-411, 492-570 (in place; the
 * result's rows are sorted by column, as the counting sort leaves them). */
int fair_transpose(fair_csr *A);
/* multiply_by_scalar -> scalar_mult, sparse_matrices.f90:824-831. */
int fair_multiply_by_scalar(fair_csr *A, double c);
/* add_scalar_mult_row -> add_mult_row, sparse_matrices.f90:795-822 (k one-based; lwrb/upb
 * NULL when absent). */
int fair_add_scalar_mult_row(fair_csr *A, int32_t k, double c, double *w, int32_t nw,
                             const double *lwrb, const double *upb);

/* Block shapes of the LDS-staged SpMV for A and A^T of a ray-traced matrix, picked by timing
 * (no counterpart in the reference, whose products are OpenMP loops over rows,
 * sparse_matrices.f90:349-393).  choice (may be NULL) receives {DA, DJ, LDS_A, TX, TY, LDS_T}. */
int fair_autotune_spmv(fair_csr *A, int32_t *choice);

/* Device-pointer forms of matvecmul / transvecmul (x, y resident in HBM). */
int fair_matvecmul_dev(fair_csr *A, const double *d_v, double *d_w);
int fair_transvecmul_dev(fair_csr *A, const double *d_v, double *d_w);
/* MMRS (sparse_matrices.f90:349-368) restricted to the non-empty rows: one kernel launch; the
 * entries of w belonging to empty rows are left untouched (callers zero them once). */
int fair_spmv_rows_dev(fair_csr *A, const double *d_v, double *d_w);

/* ---- AIRMs ----------------------------------------------------------------------------- */
enum { FAIR_CIMMINO = 0, FAIR_SART = 1 };
enum { FAIR_KACZMARZ = 0, FAIR_SYMKACZMARZ = 1, FAIR_RANDKACZMARZ = 2 };

/* sirt, AIRMs.f90:997-1127: maxits = maxval(Kin) iterations of
 * x <- box(x + relax * D^-1 A^T (M^-1 (b - A x))).  x0 may be NULL (zero start).  As in the
 * reference, SART leaves A's rows sorted by column. */
int fair_sirt(fair_csr *A, int32_t method, const double *b, int32_t maxits, const double *x0,
              const fair_opts *opts, double *xout);
/* Device-resident form: b (m), x (n, in: start, out: result) in HBM.  Asynchronous on the
 * library stream once the weights exist. */
int fair_sirt_dev(fair_csr *A, int32_t method, const double *d_b, int32_t maxits, double *d_x,
                  const fair_opts *opts);
/* fair_sirt_dev without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The residual and the second iterate live in a library-static device
 * array of 4 Mi doubles (32 MiB) per GPU, held until the process ends: m + n (the residual
 * rounded up to 32 entries) must fit it, else FAIR_ERR_ARG.  A must already hold what sirt sets
 * up for this method (fair_sirt_prepare), else FAIR_ERR_ARG.  Calls on one GPU share the array:
 * calls that overlap on one GPU race.  Same kernels and results as fair_sirt_dev. */
int fair_sirt_dev_noalloc(fair_csr *A, int32_t method, const double *d_b, int32_t maxits, double *d_x,
                          const fair_opts *opts);
/* Compute and cache the SIRT weights of sirt (AIRMs.f90:1070-1083) for A (SART also re-sorts
 * A's rows as the reference's double transpose does), so later fair_sirt_dev calls only
 * iterate. */
int fair_sirt_prepare(fair_csr *A, int32_t method);
/* Precision variant of sirt (AIRMs.f90:997-1127): FP32 matrix values, iterate and accumulation;
 * the weights M, D (row / column sums or norms) stay FP64 and scale each row / column in FP64. */
int fair_sirt_f32(fair_csr *A, int32_t method, const float *b, int32_t maxits, const float *x0,
                  const fair_opts *opts, float *xout);
int fair_sirt_f32_dev(fair_csr *A, int32_t method, const float *d_b, int32_t maxits, float *d_x,
                      const fair_opts *opts);
/* Batched sirt: k right-hand sides sharing A (k slices of one geometry), FP32 as above.
 * d_B: m x k, d_X: n x k (start in, result out), row-major.  use_graph != 0 captures the whole
 * run into a HIP graph on first use and replays it (the graph is keyed by A, d_B, d_X, method,
 * k, maxits; the library synchronises before returning).  No counterpart in the reference:
 * its to-do "blocking support" (README.md:6). */
int fair_sirt_batch_dev(fair_csr *A, int32_t method, const float *d_B, int32_t k, int32_t maxits,
                        float *d_X, const fair_opts *opts, int32_t use_graph);
/* Autotuner of the batched SpMM: times the row orders (rays in sinogram tiles for A*X, pixels in
 * image tiles for A^T*R; codes 0-11, 0 = stored order), also with two rows per wavefront in passes of
 * 128 right-hand sides (code + 1000) and four rows in passes of 64 (code + 2000), and the LDS-staged
 * tiled kernel (codes 100 + 10 c + i: chunks of 64 (c = 0) or 32 (c = 1) staged rows; pixel tiles
 * i = 4x4, 8x2, 2x8, 16x1, 1x16, 8x4, 4x8, 8x8, 16x4 for A^T*R, ray tiles i = 8x8, 16x4, 4x16, 8x4,
 * 4x8, 4x4 angles x detectors for A*X) on this matrix with k right-hand sides; pairs the fastest as
 * whole SART iterations, then tries at most b = 2..6 workgroups resident per compute unit for the row
 * kernels (code + 10000 b), and keeps the fastest in the handle (order_out[2] receives the codes, may
 * be NULL).  All choices give the same results. */
int fair_autotune_batch(fair_csr *A, int32_t k, int32_t *order_out);
/* Host-array form of fair_sirt_batch_dev (B: m x k, X0 may be NULL, Xout: n x k). */
int fair_sirt_batch(fair_csr *A, int32_t method, const float *B, int32_t k, int32_t maxits,
                    const float *X0, const fair_opts *opts, int32_t use_graph, float *Xout);
/* Per right-hand side ||X_j - Xref_j|| / ||Xref_j|| with FP64 sums (X FP32, Xref FP64, n x k). */
int fair_batch_relerr_dev(const float *d_X, const double *d_Xref, int32_t n, int32_t k, double *err);
/* normest, AIRMs.f90:1129-1173, started from v0 (the reference draws it with random_number). */
int fair_normest(fair_csr *A, const double *v0, double tol, int32_t maxit, double *sigma);
/* gradient_descent, AIRMs.f90:860-995.  sigma is normest's result (before squaring); iters
 * receives the number of accepted updates. */
int fair_gradient_descent(fair_csr *A, const double *b, int32_t maxiter, const double *x0,
                          double rho, double tol, double sigma, const fair_opts *opts,
                          double *xout, int32_t *iters);
/* art, AIRMs.f90:59-369 without xref (fair_art_oracle has it).  order may be NULL (natural order).
 * uniforms: the random_number draws in the reference's order (see fair_art_uniform_count);
 * may be NULL for the non-random methods.  Xout: nk x n, row k = iterate at Kin[k]. */
int fair_art(fair_csr *A, int32_t method, const double *b, int32_t nk, const int32_t *Kin,
             const double *x0, const int32_t *order, int32_t nord, const fair_opts *opts,
             const double *uniforms, int64_t nu, double *Xout);
/* art with xref, AIRMs.f90:322, 330-333: stops when ||x - xref|| grows over an iteration and
 * puts the iterate before that iteration in row nk-1 of Xout (rows not reached are zero).
 * *exit_iter (may be NULL): the iteration that grew the distance, 0 if none did (iterations
 * counted as Kin/2 for symkaczmarz); *consumed (may be NULL): the draws used.  nu may be below
 * fair_art_uniform_count's figure: the draws are checked as each iteration takes them. */
int fair_art_oracle(fair_csr *A, int32_t method, const double *b, int32_t nk, const int32_t *Kin,
                    const double *x0, const double *xref, const int32_t *order, int32_t nord,
                    const fair_opts *opts, const double *uniforms, int64_t nu, double *Xout,
                    int32_t *exit_iter, int64_t *consumed);
/* Number of random_number draws art consumes for these arguments. */
int fair_art_uniform_count(fair_csr *A, int32_t method, int32_t nk, const int32_t *Kin,
                           const int32_t *order, int32_t nord, const fair_opts *opts,
                           int64_t *count);
/* twin_art, AIRMs.f90:371-811: stoprule 0 = 'errorgauge', 1 = 'mutualstep'; x0 / y0 / order
 * may be NULL.  Xstore / Ystore: nk x n (row k = iterate at Kin[k]), EG: maxval(Kin) entries,
 * *final_it as the reference returns it.  uniforms: the random_number draws in the reference's
 * order (at most fair_twin_art_uniform_bound of them are read; *consumed, may be NULL, receives
 * the number used). */
int fair_twin_art(fair_csr *A, int32_t method, int32_t stoprule, const double *b, int32_t nk,
                  const int32_t *Kin, const double *x0, const double *y0, const int32_t *order, int32_t nord,
                  int32_t slack, const fair_opts *opts, const double *uniforms, int64_t nu, double *xout,
                  int32_t *final_it, double *Xstore, double *Ystore, double *EG, int64_t *consumed);
int fair_twin_art_uniform_bound(fair_csr *A, int32_t method, int32_t nk, const int32_t *Kin,
                                const int32_t *order, int32_t nord, const fair_opts *opts, int64_t *count);
/* basic_sweep, AIRMs.f90:813-858: the row steps over an explicit one-based row sequence (the
 * caller makes the row choice: ord, or the row-probability draws).  A sequence with at least 4
 * steps per level of its level schedule (steps of a level touch disjoint pixels) runs level by
 * level; the schedule is cached in the matrix, so repeating one sequence plans it once.  Other
 * sequences run in the window sweep (one workgroup; steps that share no pixel run side by side).
 * The iterate is the same bit for bit either way.
 * nrA: the rows' (damped squared) norms (m); x (n) updated in place.  Host arrays; rows outside
 * 1..m are rejected. */
int fair_art_steps(fair_csr *A, const int32_t *rows, int64_t nsteps, const double *b, const double *nrA,
                   const fair_opts *opts, double *x);
/* The same with device arrays (d_x in HBM); the rows must lie in 1..m.  Returns when the sweep is
 * done (the window sweep reads its stop flag back); the one-wavefront sweep (FAIR_ART_WINDOW=0, or
 * rows over 1,024 entries) leaves it running on the library stream. */
int fair_art_steps_dev(fair_csr *A, const int32_t *d_rows, int64_t nsteps, const double *d_b,
                       const double *d_nrA, const fair_opts *opts, double *d_x);
/* The level schedule the sweeps use, on host arrays (no device call).  No counterpart in the
 * reference: it reorders the row steps of basic_sweep's loop (AIRMs.f90:813-858) without changing
 * their result.  Step s gets level 1 + max(level of the last earlier step touching one of its
 * columns), so the steps of one level touch disjoint columns and every column sees its steps in
 * sequence order.  rows_out (nsteps): the sequence's rows sorted by (level, position);
 * off_out (nsteps + 1 capacity): level offsets, *nlev + 1 of them.  IA zero-based offsets, JA
 * one-based columns; columns outside 1..n are ignored. */
int fair_level_schedule(int32_t m, int32_t n, const int32_t *IA, const int32_t *JA, const int32_t *seq,
                        int64_t nsteps, int32_t *rows_out, int32_t *off_out, int32_t *nlev);

/* ---- groups of independent SIRT problems (uneven batch) ------------------------------------
 * No counterpart in the reference (its to-do "blocking support", README.md:6).  Problems with
 * their own matrices (different sizes) are iterated together: each sirt iteration
 * (AIRMs.f90:1089-1110) of every problem is two persistent launches over all problems' rows, and
 * every problem's iterate is bit-identical to its own fair_sirt_dev run.  The matrices must
 * outlive the group and stay unchanged. */
typedef struct fair_group fair_group;
int fair_group_create(int32_t nprob, fair_csr *const *As, int32_t method, fair_group **G);
int fair_group_free(fair_group *G);
/* maxits iterations of every problem: d_b[p] (m_p), d_x[p] (n_p; start in, result out) in HBM. */
int fair_group_sirt_dev(fair_group *G, const double *const *d_b, int32_t maxits, double *const *d_x,
                        const fair_opts *opts);
/* Picks the grid by timing: wg_per_cu workgroups per CU striding over the work items, or 0 for
 * one workgroup per item; wg_per_cu (may be NULL) receives the choice. */
int fair_group_autotune(fair_group *G, int32_t *wg_per_cu);

/* ---- matrix-free fan-beam projector (FP32) ------------------------------------------------
 * The reference's to-do "make matrix-free implementations" (README.md:7): the operator of
 * fanlineartomo (system_generation.f90:79-307) with the same angle and ray tables, its
 * intersections recomputed on every product instead of stored, so that problems whose CSR
 * would exceed int32 nnz (or HBM) can be solved.  Image n*n (column-major pixel order of the
 * reference), sinogram nth*p (angle-major). */
typedef struct fair_mf fair_mf;
int fair_mf_create(const fair_geom *g, const double *theta, fair_mf **P);
int fair_mf_free(fair_mf *P);
/* y = A x (matvecmul, sparse_matrices.f90:593-620) and x = A^T y (transvecmul, :622-656). */
int fair_mf_forward(fair_mf *P, const float *x, float *y);
int fair_mf_adjoint(fair_mf *P, const float *y, float *x);
int fair_mf_forward_dev(fair_mf *P, const float *d_x, float *d_y);
int fair_mf_adjoint_dev(fair_mf *P, const float *d_y, float *d_x);
/* Ray-pixel segments of positive length in one product (the operator's nonzeros). */
int fair_mf_segments(fair_mf *P, int64_t *count);
/* r = A x - b. */
int fair_mf_residual_dev(fair_mf *P, const float *d_x, const float *d_b, float *d_r);
/* normest, AIRMs.f90:1129-1173, on the operator (v0: n*n host start vector). */
int fair_mf_normest(fair_mf *P, const float *v0, double tol, int32_t maxit, double *sigma);
/* gradient_descent, AIRMs.f90:860-995, on the operator; d_x: x0 in, result out. */
int fair_mf_gd_dev(fair_mf *P, const float *d_b, int32_t maxiter, double rho, double tol, double sigma,
                   const fair_opts *opts, float *d_x, int32_t *iters);
/* Host-array form (x0 may be NULL: zero start). */
int fair_mf_gd(fair_mf *P, const float *b, int32_t maxiter, const float *x0, double rho, double tol,
               double sigma, const fair_opts *opts, float *xout, int32_t *iters);

/* ---- phantom gallery and white noise (per-pixel / per-sample work) ------------------------
 * phantomgallery.f90 builds its test images pixel by pixel over every shape; the shapes come
 * from short sequential loops over random_number draws, which the host caller runs as the
 * reference does and hands over here.  Images are column-major n x n (pixel (i, j) at i + j*n,
 * i the row), the layout of the reference's x(N,N); X (n) holds the coordinate of each column,
 * Y (n) that of each row, as meshgrid / intmeshgrid (phantomgallery.f90:17-82) lay them out.
 * All arrays are host arrays; each call returns after x is copied back. */
/* shepplogan (phantomgallery.f90:116-164): x = sum of A over the ellipses containing the pixel.
 * ell: ne x 7 (A, a^2, b^2, x0, y0, cos phi, sin phi). */
int fair_phantom_ellipses(int32_t n, const double *X, const double *Y, int32_t ne, const double *ell, double *x);
/* smooth (:166-199), threephases (:201-254), threephases_smooth (:256-328): x = sum over bumps of
 * w*exp(-(c*P(X-cx))/dx - (c*P(Y-cy))/dy), P(t) = t^2 (power 2) or |t|^3 (power 3).
 * bumps: nb x 6 (w, cx, cy, dx, dy, c). */
int fair_phantom_bumps(int32_t n, const double *X, const double *Y, int32_t nb, const double *bumps,
                       int32_t power, double *x);
/* grains (:529-597): x = 1-based index of the first seed point nearest in NORM2.  pts: np x 2. */
int fair_phantom_nearest(int32_t n, const double *X, const double *Y, int32_t np, const double *pts, double *x);
/* bubbles (:599-642): for each disk in order, pixels with (X-cx)^2+(Y-cy)^2 <= r2 take its value;
 * x in/out.  disks: nd x 4 (cx, cy, r2, value). */
int fair_phantom_last_disk(int32_t n, const double *X, const double *Y, int32_t nd, const double *disks,
                           double *x);
/* binary (:330-416): pixels with (X-cx)^2+(Y-cy)^2 <= r2 for some disk are set to value; x in/out.
 * disks: nd x 3 (cx, cy, r2). */
int fair_phantom_stamp_disks(int32_t n, const double *X, const double *Y, int32_t nd, const double *disks,
                             double value, double *x);
/* mst (:439-527, pow = 1): for each tree edge in order, pixels with
 * |c(X-m1) + s(Y-m2)| + |-s(X-m1) + c(Y-m2)| < r take its value; x in/out.
 * edges: ne x 6 (m1, m2, c, s, r, value). */
int fair_phantom_last_band(int32_t n, const double *X, const double *Y, int32_t ne, const double *edges,
                           double *x);
/* add_white_noise's per-sample Newton solve of Phi(e) = u (system_generation.f90:725-739), the
 * reference's step and stopping test; for |e| >= 1.5 the residual is formed with erfc (no
 * cancellation near 0 and 1).  Fails with "Newton method did not converge." as the reference
 * stops. */
int fair_normal_newton(int64_t m, const double *u, double *e);

#ifdef __cplusplus
}
#endif
#endif /* FAIR_HIP_H */
