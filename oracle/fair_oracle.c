/*
 * oracle/fair_oracle.c -- CPU restatement of the FAIR tools reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check and
 * bench.py's cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every function restates one reference procedure, in the reference's own operation
 * order, so that results are bit-identical to the reference compiled for x86-64 without
 * FMA (build with -ffp-contract=off, no -march).  Index conventions follow the reference:
 * IA holds m+1 zero-based row offsets, JA holds one-based column numbers.
 *
 * Pinning (what checks each function beyond retyping it):
 *   linspace, quicksort, fan/parallel ray tracers  -> golden vectors written by the compiled
 *       reference (oracle/_ref), analytic chord lengths, exact rational clipping at N<=8,
 *       axis-aligned unit rows, the odd-p empty half.
 *   matvec / transvec / transpose / row norms      -> golden vectors, scipy.sparse, dot-product
 *       identity, double transpose == row-sorted A.
 *   sirt / gd / normest / art steps / scramble    -> golden iterates from the compiled
 *       reference, dense numpy restatements at tiny N, Kaczmarz projection invariant,
 *       svd for sigma.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REF_PI (4.0 * atan(1.0))

/* Fortran MIN/MAX on reals as the reference compiler evaluates them (probed with flang):
 * min(a,b) = a<b ? a : b, max(a,b) = a>b ? a : b.  NaN in `a` therefore yields b. */
static inline double fmin_ref(double a, double b) { return a < b ? a : b; }
static inline double fmax_ref(double a, double b) { return a > b ? a : b; }

/* system_generation.f90:17-35 */
void oracle_linspace(double a, double b, int n, double *x) {
    double c = (b - a) / (double)(n - 1);
    double d = a - c;
    for (int i = 1; i <= n; i++) x[i - 1] = c * (double)i + d;
}

/* system_generation.f90:37-77 -- recursive Hoare partition with middle pivot.  a and ind are
 * addressed 1-based (callers pass base pointers minus one). */
static void quicksort_ref(double *a, int *ind, int f, int l) {
    double x = a[(f + l) / 2];
    int i = f, j = l;
    for (;;) {
        while (a[i] < x) i++;
        while (x < a[j]) j--;
        if (i >= j) break;
        double t = a[i]; a[i] = a[j]; a[j] = t;
        int m = ind[i]; ind[i] = ind[j]; ind[j] = m;
        i++;
        j--;
    }
    if (f < i - 1) quicksort_ref(a, ind, f, i - 1);
    if (j + 1 < l) quicksort_ref(a, ind, j + 1, l);
}

/* Per-ray scratch for the ray tracers. */
typedef struct {
    double *x, *y, *t, *xxy, *yxy, *tx_, *ty_;
    int *ind;
} ray_scratch;

static void scratch_alloc(ray_scratch *s, int n) {
    int nn = 2 * n + 2;
    s->x = malloc(sizeof(double) * (n + 1));
    s->y = malloc(sizeof(double) * (n + 1));
    s->t = malloc(sizeof(double) * (nn + 1));
    s->xxy = malloc(sizeof(double) * (nn + 1));
    s->yxy = malloc(sizeof(double) * (nn + 1));
    s->tx_ = malloc(sizeof(double) * (nn + 1));
    s->ty_ = malloc(sizeof(double) * (nn + 1));
    s->ind = malloc(sizeof(int) * (nn + 1));
}

static void scratch_free(ray_scratch *s) {
    free(s->x); free(s->y); free(s->t); free(s->xxy); free(s->yxy);
    free(s->tx_); free(s->ty_); free(s->ind);
}

/* One ray, common to fan and parallel beam: the body of system_generation.f90:162-203
 * (count) and :243-297 (fill).  (a,b) is the direction, (x0th,y0th) the start point.
 * excl is the parallel-beam boundary exclusion of :394/:466-467 (0 for fan beam).
 * Returns the entry count; writes entries when JA != NULL. */
static int trace_ray(int n, const double *xg, const double *yg, double a, double b, double x0th,
                     double y0th, int excl, ray_scratch *s, int32_t *JA, double *vA) {
    int nn = 2 * n + 2;
    double *t = s->t, *xxy = s->xxy, *yxy = s->yxy;
    int *ind = s->ind;
    /* tx = (x-x0th)/a ; yx = b*tx + y0th ; ty = (y-y0th)/b ; xy = a*ty + x0th */
    for (int k = 1; k <= n + 1; k++) {
        double tx = (xg[k - 1] - x0th) / a;
        double yx = b * tx + y0th;
        double ty = (yg[k - 1] - y0th) / b;
        double xy = a * ty + x0th;
        t[k] = tx;
        t[n + 1 + k] = ty;
        s->tx_[k] = xg[k - 1];      /* xxy = (/x, xy/) */
        s->tx_[n + 1 + k] = xy;
        s->ty_[k] = yx;             /* yxy = (/yx, y/) */
        s->ty_[n + 1 + k] = yg[k - 1];
    }
    for (int k = 1; k <= nn; k++) ind[k] = k;
    quicksort_ref(t, ind, 1, nn);
    for (int k = 1; k <= nn; k++) {
        xxy[k] = s->tx_[ind[k]];
        yxy[k] = s->ty_[ind[k]];
    }
    /* inclusive box filter, system_generation.f90:187-193 */
    double lo = -.5 * (double)n, hi = .5 * (double)n;
    int nint = 0;
    double *kx = s->tx_, *ky = s->ty_; /* reuse as temp_xxy / temp_yxy */
    for (int k = 1; k <= nn; k++) {
        if (xxy[k] >= lo && xxy[k] <= hi && yxy[k] >= lo && yxy[k] <= hi) {
            nint++;
            kx[nint] = xxy[k];
            ky[nint] = yxy[k];
        }
    }
    /* remove double points and emit, system_generation.f90:196-203 / :279-297 */
    int cnt = 0;
    for (int k = 1; k <= nint - 1; k++) {
        double dx = kx[k] - kx[k + 1];
        double dy = ky[k] - ky[k + 1];
        double tv = sqrt(dx * dx + dy * dy);
        if (tv > 1e-10 && !excl) {
            if (JA) {
                double xm = .5 * ((kx[k] + kx[k + 1]) + (double)n);
                double ym = .5 * ((ky[k] + ky[k + 1]) + (double)n);
                int fx = (int)floor(xm), fy = (int)floor(ym);
                JA[cnt] = fx * n + (n - fy);
                vA[cnt] = tv;
            }
            cnt++;
        }
    }
    return cnt;
}

/* Ray angles and detector offsets of the fan geometry, system_generation.f90:104-130
 * (including the odd-p left half `atan(sd/d) - .5`, :122). */
static void fan_omega(int p, double dew, double sd, double *omega) {
    const double pi = REF_PI;
    if (p % 2 == 1) {
        int h = (p - 1) / 2;
        for (int j = 1; j <= h; j++) {
            double dep = (double)(h + 1 - j) * dew;          /* depos(h:1:-1) */
            omega[j - 1] = atan(sd / dep) - .5;
        }
        omega[h] = 0.0;
        for (int k = 1; k <= h; k++) omega[h + k] = .5 * pi - atan(sd / ((double)k * dew));
    } else {
        int h = p / 2;
        for (int j = 1; j <= h; j++) {
            double dep = ((double)(h + 1 - j) - .5) * dew;
            omega[j - 1] = atan(sd / dep) - .5 * pi;
        }
        for (int k = 1; k <= h; k++) omega[h + k - 1] = .5 * pi - atan(sd / (((double)k - .5) * dew));
    }
}

/* fanlineartomo, system_generation.f90:79-307.  Fills IA (m+1) when non-NULL and JA/vA when
 * non-NULL (IA must then be given).  Returns nnz (int64: the reference's int32 is checked by
 * callers). */
int64_t oracle_fanlineartomo(int n, int p, int nth, double r_in, double dw_in, double sd_in,
                             const double *theta, int32_t *IA, int32_t *JA, double *vA) {
    double R = r_in * (double)n, dw = dw_in * (double)n, sd = sd_in * (double)n;
    double x0 = 0.0, y0 = R, dew = dw / (double)p;
    double *xg = malloc(sizeof(double) * (n + 1));
    for (int i = 1; i <= n + 1; i++) xg[i - 1] = (double)i - .5 * (double)n - 1.0;
    double *omega = malloc(sizeof(double) * p);
    fan_omega(p, dew, sd, omega);
    int64_t m = (int64_t)nth * p;
    int32_t *cnt = malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int fill = (JA != NULL);
    int64_t *off = NULL;
    if (fill) {
        off = malloc(sizeof(int64_t) * (m + 1));
        off[0] = 0;
        for (int64_t q = 0; q < m; q++) off[q + 1] = off[q] + (IA[q + 1] - IA[q]);
    }
#pragma omp parallel
    {
        ray_scratch s;
        scratch_alloc(&s, n);
#pragma omp for schedule(dynamic, 1)
        for (int i = 1; i <= nth; i++) {
            double st, ct;
            ct = cos(theta[i - 1]);
            st = sin(theta[i - 1]);
            double x0th = ct * x0 - st * y0;
            double y0th = st * x0 + ct * y0;
            double ath = -(x0th / R), bth = -(y0th / R);
            for (int j = 1; j <= p; j++) {
                double co = cos(omega[j - 1]), so = sin(omega[j - 1]);
                double a = co * ath - so * bth;
                double b = so * ath + co * bth;
                int64_t row = (int64_t)(i - 1) * p + (j - 1);
                if (fill)
                    trace_ray(n, xg, xg, a, b, x0th, y0th, 0, &s, JA + off[row], vA + off[row]);
                else
                    cnt[row] = trace_ray(n, xg, xg, a, b, x0th, y0th, 0, &s, NULL, NULL);
            }
        }
        scratch_free(&s);
    }
    int64_t nnz;
    if (!fill) {
        nnz = 0;
        if (IA) IA[0] = 0;
        for (int64_t q = 0; q < m; q++) {
            nnz += cnt[q];
            if (IA) IA[q + 1] = (int32_t)nnz;
        }
    } else {
        nnz = off[m];
        free(off);
    }
    free(cnt); free(omega); free(xg);
    return nnz;
}

/* paralleltomo, system_generation.f90:309-492 */
int64_t oracle_paralleltomo(int n, int p, int nth, const double *theta, int32_t *IA, int32_t *JA,
                            double *vA) {
    double d = (double)(p - 1);
    double *x0 = malloc(sizeof(double) * p), *y0 = malloc(sizeof(double) * p);
    oracle_linspace(-.5 * d, .5 * d, p, x0);
    for (int j = 0; j < p; j++) y0[j] = 0.0 * x0[j];
    double *xg = malloc(sizeof(double) * (n + 1));
    oracle_linspace(-.5 * (double)n, .5 * (double)n, n + 1, xg);
    int64_t m = (int64_t)nth * p;
    int32_t *cnt = malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int fill = (JA != NULL);
    int64_t *off = NULL;
    if (fill) {
        off = malloc(sizeof(int64_t) * (m + 1));
        off[0] = 0;
        for (int64_t q = 0; q < m; q++) off[q + 1] = off[q] + (IA[q + 1] - IA[q]);
    }
    double half = .5 * (double)n;
#pragma omp parallel
    {
        ray_scratch s;
        scratch_alloc(&s, n);
#pragma omp for schedule(dynamic, 1)
        for (int i = 1; i <= nth; i++) {
            double ct = cos(theta[i - 1]), st = sin(theta[i - 1]);
            double a = -st, b = ct;
            for (int j = 1; j <= p; j++) {
                double x0th = ct * x0[j - 1] - st * y0[j - 1];
                double y0th = st * x0[j - 1] + ct * y0[j - 1];
                int excl = (b == 0.0 && fabs(y0th - half) < 1e-15) ||
                           (a == 0.0 && fabs(x0th - half) < 1e-15);
                int64_t row = (int64_t)(i - 1) * p + (j - 1);
                if (fill)
                    trace_ray(n, xg, xg, a, b, x0th, y0th, excl, &s, JA + off[row], vA + off[row]);
                else
                    cnt[row] = trace_ray(n, xg, xg, a, b, x0th, y0th, excl, &s, NULL, NULL);
            }
        }
        scratch_free(&s);
    }
    int64_t nnz;
    if (!fill) {
        nnz = 0;
        if (IA) IA[0] = 0;
        for (int64_t q = 0; q < m; q++) {
            nnz += cnt[q];
            if (IA) IA[q + 1] = (int32_t)nnz;
        }
    } else {
        nnz = off[m];
        free(off);
    }
    free(cnt); free(x0); free(y0); free(xg);
    return nnz;
}

/* csr_matrowvec, sparse_matrices.f90:252-270: sequential dot in stored order. */
double oracle_rowvecmult(const int32_t *IA, const int32_t *JA, const double *vA, int k,
                         const double *v) {
    double f = 0.0;
    for (int32_t i = IA[k - 1]; i < IA[k]; i++) f = f + v[JA[i] - 1] * vA[i];
    return f;
}

/* MMRS, sparse_matrices.f90:349-368 */
void oracle_matvec(int m, const int32_t *IA, const int32_t *JA, const double *vA,
                   const double *x, double *y) {
    for (int k = 1; k <= m; k++) y[k - 1] = oracle_rowvecmult(IA, JA, vA, k, x);
}

/* MMCS as run single-threaded, sparse_matrices.f90:370-393 (through sparsetransmul :300-347):
 * scatter over rows in ascending order. */
void oracle_transvec(int m, int n, const int32_t *IA, const int32_t *JA, const double *vA,
                     const double *v, double *w) {
    for (int j = 0; j < n; j++) w[j] = 0.0;
    for (int k = 1; k <= m; k++)
        for (int32_t i = IA[k - 1]; i < IA[k]; i++) w[JA[i] - 1] = w[JA[i] - 1] + v[k - 1] * vA[i];
}

/* reorder, sparse_matrices.f90:492-570: counting-sort transpose.  IB has n+1 entries. */
void oracle_transpose(int m, int n, const int32_t *IA, const int32_t *JA, const double *vA,
                      int32_t *IB, int32_t *JB, double *vB) {
    int32_t nnz = IA[m];
    for (int j = 0; j <= n; j++) IB[j] = 0;
    for (int32_t i = 0; i < nnz; i++) IB[JA[i]] += 1;      /* B%IA(A%JA(i)) += 1 */
    int32_t cs = 0;
    for (int j = 0; j <= n; j++) { cs += IB[j]; IB[j] = cs; }
    for (int i = 1; i <= m; i++) {
        for (int32_t k = IA[i - 1]; k < IA[i]; k++) {
            int col = JA[k] - 1;
            int32_t dest = IB[col];                          /* one-based dest - 1 */
            JB[dest] = i;
            vB[dest] = vA[k];
            IB[col] += 1;
        }
    }
    int32_t last = 0;
    for (int c = 0; c < n; c++) { int32_t t = IB[c]; IB[c] = last; last = t; }
}

/* compute_row_norms, sparse_matrices.f90:740-780: (sum |a|^p)^(1/p) via pow. */
void oracle_row_norms(int m, const int32_t *IA, const double *vA, int p, double *nrA) {
    double e = 1.0 / (double)p;
    for (int k = 1; k <= m; k++) {
        double s = 0.0;
        for (int32_t i = IA[k - 1]; i < IA[k]; i++) {
            double v = fabs(vA[i]);
            double q = 1.0;
            for (int r = 0; r < p; r++) q = (r == 0) ? v : q * v;
            s = s + q;
        }
        nrA[k - 1] = pow(s, e);
    }
}

/* scalar_mult, sparse_matrices.f90:824-831 */
void oracle_scalar_mult(int64_t nnz, double c, double *vA) {
    for (int64_t i = 0; i < nnz; i++) vA[i] = c * vA[i];
}

/* add_mult_row, sparse_matrices.f90:795-822: x(JA) += c*vA then max(lb), min(ub) on the
 * touched entries. */
void oracle_add_mult_row(const int32_t *IA, const int32_t *JA, const double *vA, int k, double c,
                         double *w, int lwrb, double lb, int upb, double ub) {
    for (int32_t i = IA[k - 1]; i < IA[k]; i++) {
        int j = JA[i] - 1;
        w[j] = w[j] + c * vA[i];
        if (lwrb) w[j] = fmax_ref(w[j], lb);
        if (upb) w[j] = fmin_ref(w[j], ub);
    }
}

static void box_sirt(int n, double *xn, int upb, int lwrb, double ub, double lb) {
    /* AIRMs.f90:1093-1100 / :963-970: min(ub) then max(lb) on every entry */
    if (upb) for (int j = 0; j < n; j++) xn[j] = fmin_ref(xn[j], ub);
    if (lwrb) for (int j = 0; j < n; j++) xn[j] = fmax_ref(xn[j], lb);
}

/* sirt, AIRMs.f90:997-1127.  method 0 = cimmino, 1 = sart.  The matrix arrays are copied:
 * SART's double transpose re-sorts A in place in the reference (:1077-1079), which this
 * restatement applies to its private copy. */
void oracle_sirt(int method, int m, int n, const int32_t *IA_in, const int32_t *JA_in,
                 const double *vA_in, const double *b, int maxits, const double *x0, int upb,
                 int lwrb, double ub, double lb, double omg, double *xout) {
    int32_t nnz = IA_in[m];
    int32_t *IA = malloc(sizeof(int32_t) * (m + 1));
    int32_t *JA = malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
    double *vA = malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    memcpy(IA, IA_in, sizeof(int32_t) * (m + 1));
    memcpy(JA, JA_in, sizeof(int32_t) * nnz);
    memcpy(vA, vA_in, sizeof(double) * nnz);
    double *M = malloc(sizeof(double) * m), *D = malloc(sizeof(double) * n);
    double *x = malloc(sizeof(double) * n), *xn = malloc(sizeof(double) * n);
    double *r = malloc(sizeof(double) * m), *u = malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) x[j] = x0 ? x0[j] : 0.0;
    if (method == 0) {
        oracle_row_norms(m, IA, vA, 2, M);
        for (int i = 0; i < m; i++) M[i] = (double)m * (M[i] * M[i]);
        for (int j = 0; j < n; j++) D[j] = 1.0;
    } else {
        oracle_row_norms(m, IA, vA, 1, M);
        int32_t *IB = malloc(sizeof(int32_t) * (n + 1));
        int32_t *JB = malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
        double *vB = malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
        oracle_transpose(m, n, IA, JA, vA, IB, JB, vB);
        oracle_row_norms(n, IB, vB, 1, D);
        oracle_transpose(n, m, IB, JB, vB, IA, JA, vA);
        free(IB); free(JB); free(vB);
    }
    for (int it = 0; it < maxits; it++) {
        oracle_matvec(m, IA, JA, vA, x, r);
        for (int i = 0; i < m; i++) r[i] = (b[i] - r[i]) / M[i];
        oracle_transvec(m, n, IA, JA, vA, r, u);
        for (int j = 0; j < n; j++) xn[j] = x[j] + (omg * u[j]) / D[j];
        box_sirt(n, xn, upb, lwrb, ub, lb);
        memcpy(x, xn, sizeof(double) * n);
    }
    memcpy(xout, x, sizeof(double) * n);
    free(IA); free(JA); free(vA); free(M); free(D); free(x); free(xn); free(r); free(u);
}

static double norm2_plain(int n, const double *v) {
    double s = 0.0;
    for (int j = 0; j < n; j++) s += v[j] * v[j];
    return sqrt(s);
}

/* normest, AIRMs.f90:1129-1173, started from a supplied vector v0 (the reference draws it
 * with random_number). */
double oracle_normest(int m, int n, const int32_t *IA, const int32_t *JA, const double *vA,
                      const double *v0, double tol, int maxit, int *its) {
    double *v = malloc(sizeof(double) * n), *vn = malloc(sizeof(double) * n);
    double *w = malloc(sizeof(double) * m), *dv = malloc(sizeof(double) * n);
    memcpy(v, v0, sizeof(double) * n);
    int i;
    for (i = 1; i <= maxit; i++) {
        oracle_matvec(m, IA, JA, vA, v, w);
        oracle_transvec(m, n, IA, JA, vA, w, vn);
        double nv = norm2_plain(n, vn);
        for (int j = 0; j < n; j++) vn[j] = vn[j] / nv;
        for (int j = 0; j < n; j++) dv[j] = vn[j] - v[j];
        if (norm2_plain(n, dv) < tol) break;
        memcpy(v, vn, sizeof(double) * n);
    }
    if (its) *its = i;
    oracle_matvec(m, IA, JA, vA, v, w);
    double ww = 0.0, vv = 0.0;
    for (int k = 0; k < m; k++) ww += w[k] * w[k];
    for (int j = 0; j < n; j++) vv += v[j] * v[j];
    free(v); free(vn); free(w); free(dv);
    return sqrt(ww / vv);
}

/* gradient_descent, AIRMs.f90:860-995, with sigma (the normest result, before squaring)
 * supplied.  Returns the number of completed updates. */
int oracle_gd(int m, int n, const int32_t *IA, const int32_t *JA, const double *vA,
              const double *b, int maxiter, const double *x0, double rho, double tol,
              double sigma, int upb, int lwrb, double ub, double lb, double *xout) {
    double *r = malloc(sizeof(double) * m), *g = malloc(sizeof(double) * n);
    double *xn = malloc(sizeof(double) * n), *dx = malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) xout[j] = x0 ? x0[j] : 0.0;
    double s2 = sigma * sigma;
    for (int i = 0; i < m; i++) r[i] = -b[i];
    int done = 0;
    for (int it = 1; it <= maxiter; it++) {
        oracle_transvec(m, n, IA, JA, vA, r, g);
        double step = rho / s2;
        for (int j = 0; j < n; j++) xn[j] = xout[j] - step * g[j];
        box_sirt(n, xn, upb, lwrb, ub, lb);
        for (int j = 0; j < n; j++) dx[j] = xn[j] - xout[j];
        if (norm2_plain(n, dx) < tol) break;
        memcpy(xout, xn, sizeof(double) * n);
        done = it;
        oracle_matvec(m, IA, JA, vA, xout, r);
        for (int i = 0; i < m; i++) r[i] = r[i] - b[i];
    }
    free(r); free(g); free(xn); free(dx);
    return done;
}

/* basic_sweep's ART step, AIRMs.f90:840-856, applied to an explicit one-based row sequence.
 * nrA holds the damped squared row norms (AIRMs.f90:273-275). */
void oracle_art_steps(const int32_t *IA, const int32_t *JA, const double *vA, const double *b,
                      const double *nrA, const int32_t *rows, int64_t nsteps, double omg,
                      int upb, int lwrb, double ub, double lb, double *x) {
    for (int64_t s = 0; s < nsteps; s++) {
        int ri = rows[s];
        double inp = oracle_rowvecmult(IA, JA, vA, ri, x);
        double coeff = (omg * (b[ri - 1] - inp)) / nrA[ri - 1];
        oracle_add_mult_row(IA, JA, vA, ri, coeff, x, lwrb, lb, upb, ub);
    }
}

/* scramble, AIRMs.f90:1175-1212, consuming 2*nord supplied uniforms in draw order. */
void oracle_scramble(int nord, const double *u, int32_t *out) {
    for (int i = 0; i < nord; i++) out[i] = i + 1;
    int q = 0;
    for (int k = 0; k < 2; k++) {
        for (int i = 1; i <= nord; i++) {
            int j = 1 + (int)floor((double)nord * u[q++]);
            int32_t t = out[j - 1];
            out[j - 1] = out[i - 1];
            out[i - 1] = t;
        }
    }
}

/* Row-probability tables of art, AIRMs.f90:302-308: cumul(i) = cumsum(nrA(1:nord))/sum(nrA),
 * tail set to 1. */
void oracle_rowprob_cumul(int m, int nord, const double *nrA, double *cumul) {
    double cs = 0.0, tot = 0.0;
    for (int i = 0; i < nord; i++) { cs = cs + nrA[i]; cumul[i] = cs; }
    for (int i = 0; i < m; i++) tot = tot + nrA[i];
    for (int i = 0; i < nord; i++) cumul[i] = cumul[i] / tot;
    for (int i = nord; i < m; i++) cumul[i] = 1.0;
}

/* Linear search of basic_sweep, AIRMs.f90:830-835.  Returns a one-based row. */
int oracle_rowprob_pick(int m, const double *cumul, double u) {
    int ri = 1;
    while (ri <= m && cumul[ri - 1] < u) ri++;
    return ri;
}

/* shepplogan + meshgrid + vectorise, phantomgallery.f90:17-56, 116-164, 672-690.  x is N*N,
 * column-major (the vectorised image). */
void oracle_shepplogan(int N, double *x) {
    static const double e[10][6] = {
        {1.0, .6900, .9200, 0.0, 0.0, 0.0},       {-.8, .6624, .8740, 0.0, -.0184, 0.0},
        {-.2, .1100, .3100, .22, 0.0, -18.0},     {-.2, .1600, .4100, -.22, 0.0, 18.0},
        {.1, .2100, .2500, 0.0, .35, 0.0},        {.1, .0460, .0460, 0.0, .1, 0.0},
        {.1, .0460, .0460, 0.0, -.1, 0.0},        {.1, .0460, .0230, -.08, -.605, 0.0},
        {.1, .0230, .0230, 0.0, -.606, 0.0},      {.1, .0230, .0460, .06, -.605, 0.0}};
    const double pi = REF_PI;
    double *lx = malloc(sizeof(double) * N), *ly = malloc(sizeof(double) * N);
    oracle_linspace(-1.0, 1.0, N, lx);   /* xc(i,j) = lx(j) */
    oracle_linspace(1.0, -1.0, N, ly);   /* yc(i,j) = ly(i) */
    for (int q = 0; q < N * N; q++) x[q] = 0.0;
    for (int el = 0; el < 10; el++) {
        double A = e[el][0], a2 = e[el][1] * e[el][1], b2 = e[el][2] * e[el][2];
        double x0 = e[el][3], y0 = e[el][4];
        double phi = e[el][5] * pi / 180.0;
        double cp = cos(phi), sp = sin(phi);
        for (int j = 0; j < N; j++) {
            for (int i = 0; i < N; i++) {
                double xc = lx[j], yc = ly[i];
                double u = (xc - x0) * cp + (yc - y0) * sp;
                double v = (yc - y0) * cp - (xc - x0) * sp;
                if ((u * u) / a2 + (v * v) / b2 <= 1.0) x[i + j * N] = x[i + j * N] + A;
            }
        }
    }
    free(lx); free(ly);
}

/* art, AIRMs.f90:59-369 (without the xref early stop), drawing its uniforms from u[] in the
 * order the reference calls random_number: 2*nord for the set-up scramble (:297), then per
 * iteration 2*nord for scramble (:320) and, with rowprobability, one per row step (:830).
 * method 0 = kaczmarz, 1 = symkaczmarz, 2 = randkaczmarz.  Xout is nk x n (row k = the
 * iterate recorded at Kin[k]).  Returns the number of uniforms consumed, -1 for an odd
 * symkaczmarz K, -2 when nu uniforms are too few. */
int64_t oracle_art(int m, int n, const int32_t *IA, const int32_t *JA, const double *vA,
                   const double *b, int method, int nk, const int32_t *Kin, const double *x0,
                   const int32_t *order, int nord_in, int upb, int lwrb, double ub, double lb,
                   double omg, double dmp, int rowprob, const double *u, int64_t nu, double *Xout) {
    int isrand = (method == 2);
    int32_t *Kuse = malloc(sizeof(int32_t) * nk);
    for (int k = 0; k < nk; k++) Kuse[k] = Kin[k];
    if (method == 1) {
        for (int k = 0; k < nk; k++) if (Kin[k] % 2 == 1) { free(Kuse); return -1; }
        for (int k = 0; k < nk; k++) Kuse[k] = Kuse[k] / 2;
    }
    double *x = malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) x[j] = x0 ? x0[j] : 0.0;
    int32_t *ord = malloc(sizeof(int32_t) * (m + 1)), *tmp = malloc(sizeof(int32_t) * (m + 1));
    int nord;
    if (order) { nord = nord_in; for (int i = 0; i < nord; i++) ord[i] = order[i]; }
    else { nord = m; for (int i = 0; i < m; i++) ord[i] = i + 1; }
    double *nrA = malloc(sizeof(double) * m);
    oracle_row_norms(m, IA, vA, 2, nrA);
    double mx = -INFINITY;
    for (int i = 0; i < m; i++) { nrA[i] = nrA[i] * nrA[i]; }
    for (int i = 0; i < m; i++) if (nrA[i] > mx) mx = nrA[i];
    for (int i = 0; i < m; i++) nrA[i] = nrA[i] + dmp * mx;
    int l = nord; nord = 0;
    for (int i = 0; i < l; i++) if (nrA[ord[i] - 1] > 0) tmp[nord++] = ord[i];
    for (int i = 0; i < nord; i++) ord[i] = tmp[i];
    int64_t q = 0;
    double *cumul = malloc(sizeof(double) * m);
    int64_t need = isrand ? 2 * (int64_t)nord : 0;
    {
        int mk = 0;
        for (int k = 0; k < nk; k++) if (Kuse[k] > mk) mk = Kuse[k];
        if (isrand) need += (int64_t)mk * (2 * (int64_t)nord + (rowprob ? nord : 0));
    }
    if (need > nu) { free(Kuse); free(x); free(ord); free(tmp); free(nrA); return -2; }
    if (isrand) {
        oracle_scramble(nord, u + q, ord); q += 2 * (int64_t)nord;
        if (rowprob) oracle_rowprob_cumul(m, nord, nrA, cumul);
    }
    int maxiter = 0;
    for (int k = 0; k < nk; k++) if (Kuse[k] > maxiter) maxiter = Kuse[k];
    int32_t *rows = malloc(sizeof(int32_t) * (nord > 0 ? nord : 1));
    int rec = 0;
    for (int iter = 1; iter <= maxiter; iter++) {
        if (isrand) { oracle_scramble(nord, u + q, ord); q += 2 * (int64_t)nord; }
        for (int ii = 0; ii < nord; ii++) {
            if (isrand && rowprob) rows[ii] = oracle_rowprob_pick(m, cumul, u[q++]);
            else rows[ii] = ord[ii];
        }
        oracle_art_steps(IA, JA, vA, b, nrA, rows, nord, omg, upb, lwrb, ub, lb, x);
        if (method == 1) {
            for (int ii = 0; ii < nord; ii++) rows[ii] = ord[nord - 1 - ii];
            oracle_art_steps(IA, JA, vA, b, nrA, rows, nord, omg, upb, lwrb, ub, lb, x);
        }
        if (rec < nk && iter == Kuse[rec]) { memcpy(Xout + (int64_t)rec * n, x, sizeof(double) * n); rec++; }
    }
    free(Kuse); free(x); free(ord); free(tmp); free(nrA); free(cumul); free(rows);
    return q;
}

/* twin_art, AIRMs.f90:371-811: two Kaczmarz chains, x over the row order and y over the
 * reversed order (forward + reverse pairs for symkaczmarz), with the 'errorgauge' stop (stoprule
 * 0: slack iterations without a new minimum of ||x - y||, result (x + y)/2 at the minimum) or
 * the 'mutualstep' stop (stoprule 1: a 2x2 least-squares step along both chains' last steps,
 * exit when |alpha| + |beta| < 1e-5).  Uniform draws are consumed in the reference's order:
 * scramble (2*nord) at set-up and every iteration, one per row step with rowprobability.
 * Xstore / Ystore: n x nk (column k at Kuse(k)), EG: maxval(Kin) entries.  Returns the draws
 * consumed, -1 for odd Kin with symkaczmarz, -2 when u runs out, -3 for x0, y0 too close. */
static double nrm2_diff(const double *a, const double *b, int n) {
    double s = 0.0;
    for (int j = 0; j < n; j++) { double d = a[j] - b[j]; s += d * d; }
    return sqrt(s);
}

static double dot_n(const double *a, const double *b, int n) {
    double s = 0.0;
    for (int j = 0; j < n; j++) s += a[j] * b[j];
    return s;
}

typedef struct {
    const int32_t *IA, *JA;
    const double *vA, *b, *nrA, *cumul, *u;
    int m, nord, isrand, rowprob, sym, upb, lwrb;
    double omg, ub, lb;
    int64_t q, nu;
    int32_t *rows;
} TwinCtx;

/* basic_sweep (AIRMs.f90:813-858) over ord (reversed if rev); -2 when the draws run out */
static int twin_sweep(TwinCtx *c, const int32_t *ord, int rev, double *x) {
    for (int ii = 0; ii < c->nord; ii++) {
        if (c->isrand && c->rowprob) {
            if (c->q >= c->nu) return -2;
            c->rows[ii] = oracle_rowprob_pick(c->m, c->cumul, c->u[c->q++]);
        } else {
            c->rows[ii] = rev ? ord[c->nord - 1 - ii] : ord[ii];
        }
    }
    oracle_art_steps(c->IA, c->JA, c->vA, c->b, c->nrA, c->rows, c->nord, c->omg, c->upb, c->lwrb, c->ub, c->lb, x);
    return 0;
}

/* the x chain's sweeps of one iteration (x: ord then reversed), or the y chain's (y: reversed
 * then ord) */
static int twin_chain(TwinCtx *c, const int32_t *ord, int is_y, double *v) {
    if (twin_sweep(c, ord, is_y, v) < 0) return -2;
    if (c->sym && twin_sweep(c, ord, !is_y, v) < 0) return -2;
    return 0;
}

int64_t oracle_twin_art(int m, int n, const int32_t *IA, const int32_t *JA, const double *vA,
                        const double *b, int method, int stoprule, int nk, const int32_t *Kin,
                        const double *x0, const double *y0, const int32_t *order, int nord_in,
                        int slack, int upb, int lwrb, double ub, double lb, double omg, double dmp,
                        int rowprob, const double *u, int64_t nu, double *xout, int *final_it,
                        double *Xstore, double *Ystore, double *EG) {
    int isrand = (method == 2);
    int32_t *Kuse = malloc(sizeof(int32_t) * nk);
    int maxK = 0;
    for (int k = 0; k < nk; k++) { Kuse[k] = Kin[k]; if (Kin[k] > maxK) maxK = Kin[k]; }
    if (method == 1) {
        for (int k = 0; k < nk; k++) if (Kin[k] % 2 == 1) { free(Kuse); return -1; }
        for (int k = 0; k < nk; k++) Kuse[k] = Kuse[k] / 2;
    }
    for (int k = 0; k < maxK; k++) EG[k] = 0.0;
    int32_t *ord = malloc(sizeof(int32_t) * (m + 1)), *tmp = malloc(sizeof(int32_t) * (m + 1));
    int nord;
    if (order) { nord = nord_in; for (int i = 0; i < nord; i++) ord[i] = order[i]; }
    else { nord = m; for (int i = 0; i < m; i++) ord[i] = i + 1; }
    double *nrA = malloc(sizeof(double) * m), *cumul = malloc(sizeof(double) * m);
    oracle_row_norms(m, IA, vA, 2, nrA);                                  /* :604-606 */
    double mx = -INFINITY;
    for (int i = 0; i < m; i++) nrA[i] = nrA[i] * nrA[i];
    for (int i = 0; i < m; i++) if (nrA[i] > mx) mx = nrA[i];
    for (int i = 0; i < m; i++) nrA[i] = nrA[i] + dmp * mx;
    int l = nord; nord = 0;                                               /* :609-617 */
    for (int i = 0; i < l; i++) if (nrA[ord[i] - 1] > 0) tmp[nord++] = ord[i];
    for (int i = 0; i < nord; i++) ord[i] = tmp[i];
    TwinCtx c = {IA, JA, vA, b, nrA, cumul, u, m, nord, isrand, rowprob, method == 1, upb, lwrb, omg, ub, lb,
                 0, nu, malloc(sizeof(int32_t) * (nord > 0 ? nord : 1))};
    int64_t ret = 0;
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    double *v = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) v[j] = w[j] = 0.0;
    if (isrand) {                                                        /* :620-633 */
        if (c.q + 2 * (int64_t)nord > nu) { ret = -2; goto done; }
        oracle_scramble(nord, u + c.q, ord); c.q += 2 * (int64_t)nord;
        if (rowprob) oracle_rowprob_cumul(m, nord, nrA, cumul);
    }
    int startiter;
    if (x0 && y0) {                                                      /* :650-654 */
        for (int j = 0; j < n; j++) { x[j] = x0[j]; y[j] = y0[j]; }
        if (nrm2_diff(x0, y0, n) < 1e-10) { ret = -3; goto done; }
        startiter = 1;
    } else {                                                             /* :634-649, 655-666 */
        const double *s0 = x0 ? x0 : y0;
        for (int j = 0; j < n; j++) x[j] = y[j] = s0 ? s0[j] : 0.0;
        if (twin_chain(&c, ord, 0, x) < 0 || twin_chain(&c, ord, 1, y) < 0) { ret = -2; goto done; }
        EG[0] = nrm2_diff(x, y, n);
        startiter = 2;
    }
    int maxiter = 0;
    for (int k = 0; k < nk; k++) if (Kuse[k] > maxiter) maxiter = Kuse[k];
    int rec = 0, trck = 0, iter;
    double minEG = HUGE_VAL;
    for (iter = startiter; iter <= maxiter; iter++) {                    /* :672-770 */
        if (isrand) {
            if (c.q + 2 * (int64_t)nord > nu) { ret = -2; goto done; }
            oracle_scramble(nord, u + c.q, ord); c.q += 2 * (int64_t)nord;
        }
        if (stoprule == 1) memcpy(v, x, sizeof(double) * n);
        if (twin_chain(&c, ord, 0, x) < 0) { ret = -2; goto done; }
        if (stoprule == 1) for (int j = 0; j < n; j++) { v[j] = x[j] - v[j]; x[j] = x[j] - v[j]; }
        if (stoprule == 1) memcpy(w, y, sizeof(double) * n);
        if (twin_chain(&c, ord, 1, y) < 0) { ret = -2; goto done; }
        if (stoprule == 1) for (int j = 0; j < n; j++) { w[j] = y[j] - w[j]; y[j] = y[j] - w[j]; }
        if (stoprule == 0) {
            EG[iter - 1] = nrm2_diff(x, y, n);
            if (EG[iter - 1] < minEG) {
                minEG = EG[iter - 1];
                memcpy(v, x, sizeof(double) * n);
                memcpy(w, y, sizeof(double) * n);
                trck = 0;
                *final_it = iter;
            } else {
                trck = trck + 1;
                if (trck >= slack) break;
            }
        } else {
            double nw = sqrt(dot_n(w, w, n)), nv = sqrt(dot_n(v, v, n));
            for (int j = 0; j < n; j++) { w[j] = w[j] / nw; v[j] = v[j] / nv; }
            double m1 = dot_n(v, v, n), m2 = dot_n(w, w, n), m3 = -dot_n(v, w, n);
            double r1 = 0.0, r2 = 0.0;
            for (int j = 0; j < n; j++) { double d = x[j] - y[j]; r1 += v[j] * d; r2 += w[j] * d; }
            r1 = -r1;
            double det = m1 * m2 - m3 * m3;
            double alph = (r1 * m2 - r2 * m3) / det;
            double bet = (m1 * r2 - m3 * r1) / det;
            for (int j = 0; j < n; j++) { x[j] = x[j] + alph * v[j]; y[j] = y[j] + bet * w[j]; }
            EG[iter - 1] = nrm2_diff(x, y, n);
            if (fabs(alph) + fabs(bet) < 1e-5) break;
        }
        if (rec < nk && iter == Kuse[rec]) {
            memcpy(Xstore + (int64_t)rec * n, x, sizeof(double) * n);
            memcpy(Ystore + (int64_t)rec * n, y, sizeof(double) * n);
            rec++;
        }
    }
    if (stoprule == 0) { for (int j = 0; j < n; j++) xout[j] = .5 * (v[j] + w[j]); }
    else { for (int j = 0; j < n; j++) xout[j] = .5 * (x[j] + y[j]); *final_it = iter; }
    ret = c.q;
done:
    free(Kuse); free(ord); free(tmp); free(nrA); free(cumul); free(c.rows);
    free(x); free(y); free(v); free(w);
    return ret;
}

/* Number of non-empty rows (the `nord` of art for the natural order), used to size the
 * uniform stream. */
int oracle_count_nonempty(int m, const int32_t *IA) {
    int c = 0;
    for (int i = 0; i < m; i++) c += IA[i + 1] > IA[i];
    return c;
}
