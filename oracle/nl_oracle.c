/*
 * ORACLE -- test infrastructure, not product code.  See nl_oracle.h for the
 * rules that govern its use and for what each routine follows.
 *
 * Built with -ffp-contract=off: a*b+c is always two roundings here.
 *
 * Parity status of each routine (how it is pinned):
 *   or_spmv / or_spmv_omp   closed-form fixtures, dense brute force, scipy CSR
 *   or_d* (BLAS-1)          exact dyadic fixtures, math.fsum error bounds
 *   or_jacobi               closed form; the zero-diagonal fallback (1.0) is
 *                           this build's choice: parity unpinned upstream
 *   or_pcg                  linear_algebra.lua KAT ([[1,2],[2,3]] -> [-3,4]),
 *                           closed-form tridiagonal solves, finite
 *                           termination, b = 0 -> 0 iterations, scipy solve
 *   or_bicgstab             linear_algebra.lua KAT ([[1,2],[3,4]] ->
 *                           [-4,4.5]), scipy solve on nonsymmetric systems
 *   or_gmres                numpy solve on nonsymmetric systems, finite
 *                           termination with m = n, non-increasing true
 *                           residual over restarts, scipy's gmres solution
 *   or_ssor                 dense omega (2 - omega) (D + omega U)^-1 D
 *                           (D + omega L)^-1 x by scipy triangular solves,
 *                           omega = 1 symmetric Gauss-Seidel by hand on a
 *                           3x3 system, diagonal matrices (y = x / d);
 *                           OpenNL's own operation order is unpinned
 *                           (its source is absent)
 *   or_solve_ssor           the SSOR-preconditioned CG / BiCGSTAB / GMRES
 *                           against numpy solves
 *   or_dyn_*                duplicate accumulation, scipy COO->CSR
 *   or_ctx_*                dense A^T A from numpy, flat-grid LSCM closed form
 */
#include "nl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ SpMV */

static double row_dot(const uint32_t* colind, const double* val, uint32_t b,
                      uint32_t e, const double* x) {
    double s = 0.0;
    for (uint32_t jj = b; jj < e; ++jj) {
        double p = val[jj] * x[colind[jj]];
        s = s + p;
    }
    return s;
}

void or_spmv(uint32_t m, const uint32_t* rowptr, const uint32_t* colind,
             const double* val, const double* x, double* y) {
    for (uint32_t i = 0; i < m; ++i) {
        y[i] = row_dot(colind, val, rowptr[i], rowptr[i + 1], x);
    }
}

void or_spmv_omp(uint32_t m, const uint32_t* rowptr, const uint32_t* colind,
                 const double* val, const double* x, double* y) {
    long mm = (long)m;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < mm; ++i) {
        y[i] = row_dot(colind, val, rowptr[i], rowptr[i + 1], x);
    }
}

/* ---------------------------------------------------------------- BLAS-1 */

void or_dcopy(uint32_t n, const double* x, double* y) {
    if (n != 0) memmove(y, x, sizeof(double) * (size_t)n);
}

void or_dscal(uint32_t n, double a, double* x) {
    for (uint32_t i = 0; i < n; ++i) x[i] = a * x[i];
}

double or_ddot(uint32_t n, const double* x, const double* y) {
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double p = x[i] * y[i];
        s = s + p;
    }
    return s;
}

/* Documented as sqrt(dot(x,x)) at scene_graph/NL/blas.h:93-98. */
double or_dnrm2(uint32_t n, const double* x) { return sqrt(or_ddot(n, x, x)); }

void or_daxpy(uint32_t n, double a, const double* x, double* y) {
    for (uint32_t i = 0; i < n; ++i) {
        double p = a * x[i];
        y[i] = y[i] + p;
    }
}

/* Static slices, each summed left to right, slices combined in order. */
double or_ddot_omp(uint32_t n, const double* x, const double* y) {
#ifdef _OPENMP
    int nt = omp_get_max_threads();
    double part[512];
    if (nt > 512) nt = 512;
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        int T = omp_get_num_threads();
        uint64_t lo = (uint64_t)n * (uint64_t)t / (uint64_t)T;
        uint64_t hi = (uint64_t)n * (uint64_t)(t + 1) / (uint64_t)T;
        double s = 0.0;
        for (uint64_t i = lo; i < hi; ++i) {
            double p = x[i] * y[i];
            s = s + p;
        }
        part[t] = s;
        if (t == 0) {
            for (int k = T; k < nt; ++k) part[k] = 0.0;
        }
    }
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s = s + part[t];
    return s;
#else
    return or_ddot(n, x, y);
#endif
}

static void daxpy_omp(uint32_t n, double a, const double* x, double* y) {
    long nn = (long)n;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nn; ++i) {
        double p = a * x[i];
        y[i] = y[i] + p;
    }
}

static void dscal_omp(uint32_t n, double a, double* x) {
    long nn = (long)n;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nn; ++i) x[i] = a * x[i];
}

/* ---------------------------------------------------------------- Jacobi */

void or_jacobi(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
               const double* val, double* dinv) {
    for (uint32_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (uint32_t jj = rowptr[i]; jj < rowptr[i + 1]; ++jj) {
            if (colind[jj] == i) d = d + val[jj];
        }
        dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
    }
}

/* A preconditioner as the solvers below apply it: Jacobi (dinv), SSOR
 * (ssor != 0, the matrix and omega), or none (both unset). */
typedef struct {
    const double* dinv;
    int ssor;
    const uint32_t* rp;
    const uint32_t* ci;
    const double* a;
    double omega;
} prec_t;

static void precond(const prec_t* P, uint32_t n, const double* r, double* h,
                    int par) {
    if (P->ssor) {
        or_ssor(n, P->rp, P->ci, P->a, P->omega, r, h);
        return;
    }
    const double* dinv = P->dinv;
    if (dinv == NULL) {
        or_dcopy(n, r, h);
        return;
    }
    if (par) {
        long nn = (long)n;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nn; ++i) h[i] = dinv[i] * r[i];
    } else {
        for (uint32_t i = 0; i < n; ++i) h[i] = dinv[i] * r[i];
    }
}

/* SSOR (nlNewSSORPreconditioner(M, omega); omega = 1.5 at
 * scene_graph/NL/matrix.cpp:366-368).  OpenNL's source is not in the
 * reference tree; this is the published SSOR preconditioner written as
 * OpenNL applies it (a forward sweep, a diagonal scaling, a backward sweep
 * and a final scaling by 2 - omega):
 *   w_i = (x_i - sum_{j<i} a_ij w_j) * omega / d_i      i = 0 .. n-1
 *   w_i = w_i * (d_i / omega)
 *   y_i = (w_i - sum_{j>i} a_ij y_j) * omega / d_i      i = n-1 .. 0
 *   y_i = (2 - omega) * y_i
 * i.e. y = omega (2 - omega) (D + omega U)^-1 D (D + omega L)^-1 x with
 * A = L + D + U; each sum runs over row i's entries in CRS order, one
 * rounding per product and per addition; d_i is the row's diagonal entry
 * (the sum of its j == i entries). */
void or_ssor(uint32_t n, const uint32_t* rp, const uint32_t* ci,
             const double* a, double omega, const double* x, double* y) {
    double* d = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    double* w = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        double di = 0.0;
        for (uint32_t e = rp[i]; e < rp[i + 1]; ++e)
            if (ci[e] == i) di = di + a[e];
        d[i] = di;
    }
    for (uint32_t i = 0; i < n; ++i) {
        double S = 0.0;
        for (uint32_t e = rp[i]; e < rp[i + 1]; ++e)
            if (ci[e] < i) S = S + a[e] * w[ci[e]];
        w[i] = (x[i] - S) * omega / d[i];
    }
    for (uint32_t i = 0; i < n; ++i) w[i] = w[i] * (d[i] / omega);
    for (uint32_t k = n; k-- > 0;) {
        double S = 0.0;
        for (uint32_t e = rp[k]; e < rp[k + 1]; ++e)
            if (ci[e] > k) S = S + a[e] * y[ci[e]];
        y[k] = (w[k] - S) * omega / d[k];
    }
    for (uint32_t i = 0; i < n; ++i) y[i] = (2.0 - omega) * y[i];
    free(d);
    free(w);
}

/* ------------------------------------------------------------------- PCG */

typedef struct {
    int spmv_par, blas_par;
} variant_t;

static variant_t decode(int v) {
    variant_t r;
    r.spmv_par = (v >= 1);
    r.blas_par = (v >= 2);
    return r;
}

static void vspmv(variant_t v, uint32_t n, const uint32_t* rp,
                  const uint32_t* ci, const double* a, const double* x,
                  double* y) {
    if (v.spmv_par) or_spmv_omp(n, rp, ci, a, x, y);
    else or_spmv(n, rp, ci, a, x, y);
}

static double vdot(variant_t v, uint32_t n, const double* x, const double* y) {
    return v.blas_par ? or_ddot_omp(n, x, y) : or_ddot(n, x, y);
}

static void vaxpy(variant_t v, uint32_t n, double a, const double* x,
                  double* y) {
    if (v.blas_par) daxpy_omp(n, a, x, y);
    else or_daxpy(n, a, x, y);
}

static void vscal(variant_t v, uint32_t n, double a, double* x) {
    if (v.blas_par) dscal_omp(n, a, x);
    else or_dscal(n, a, x);
}

/*
 * Textbook PCG (Hestenes & Stiefel 1952; Saad, Iterative Methods for Sparse
 * Linear Systems, 2nd ed., Alg. 9.1), written with the residual convention
 * r = A x - b and the operation order assumed for OpenNL's CG:
 *   r = A x - b; d = P r; h = d; rh = r.h; err = r.r
 *   while err > eps^2 b.b and its < max_iter:
 *     Ad = A d; alpha = rh / d.Ad; x -= alpha d; r -= alpha Ad; h = P r
 *     rh' = r.h; beta = rh'/rh; rh = rh'; d = beta d; d += h; err = r.r
 * The stopping rule |Ax-b|/|b| < eps is scene_graph/NL/matrix.h:177-179 and
 * x0 comes from the caller (the wrapper zeroes it, matrix.cpp:387).
 */
static int pcg_core(uint32_t n, const uint32_t* rp, const uint32_t* ci,
                    const double* a, const prec_t* P, const double* b,
                    double* x, double eps, int max_iter, variant_t v,
                    int check, double* out_rel2) {
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    double* h = (double*)malloc(sizeof(double) * (size_t)n);
    double* Ad = (double*)malloc(sizeof(double) * (size_t)n);
    vspmv(v, n, rp, ci, a, x, r);
    vaxpy(v, n, -1.0, b, r);
    precond(P, n, r, d, v.blas_par);
    or_dcopy(n, d, h);
    double rh = vdot(v, n, r, h);
    double err = vdot(v, n, r, r);
    double bb = vdot(v, n, b, b);
    double thresh = (eps * eps) * bb;
    int its = 0;
    while ((!check || err > thresh) && its < max_iter) {
        vspmv(v, n, rp, ci, a, d, Ad);
        double dAd = vdot(v, n, d, Ad);
        double alpha = rh / dAd;
        vaxpy(v, n, -alpha, d, x);
        vaxpy(v, n, -alpha, Ad, r);
        precond(P, n, r, h, v.blas_par);
        double rh_new = vdot(v, n, r, h);
        double beta = rh_new / rh;
        rh = rh_new;
        vscal(v, n, beta, d);
        vaxpy(v, n, 1.0, h, d);
        ++its;
        err = vdot(v, n, r, r);
    }
    if (out_rel2) *out_rel2 = (bb > 0.0) ? err / bb : err;
    free(r);
    free(d);
    free(h);
    free(Ad);
    return its;
}

int or_pcg(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
           const double* val, const double* dinv, const double* b, double* x,
           double eps, int max_iter, int variant, double* out_rel2) {
    prec_t P = {dinv, 0, NULL, NULL, NULL, 0.0};
    return pcg_core(n, rowptr, colind, val, &P, b, x, eps, max_iter,
                    decode(variant), 1, out_rel2);
}

double or_pcg_timed_iters(uint32_t n, const uint32_t* rowptr,
                          const uint32_t* colind, const double* val,
                          const double* dinv, const double* b, double* x,
                          int iters, int variant) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    prec_t P = {dinv, 0, NULL, NULL, NULL, 0.0};
    pcg_core(n, rowptr, colind, val, &P, b, x, 0.0, iters, decode(variant),
             0, NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) +
           1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* -------------------------------------------------------------- BiCGSTAB */

static int bicgstab_core(uint32_t n, const uint32_t* rp, const uint32_t* ci,
                         const double* a, const prec_t* P, const double* b,
                         double* x, double eps, int max_iter, double* out_rel2) {
    size_t sz = sizeof(double) * (size_t)n;
    double *r = malloc(sz), *rt = malloc(sz), *p = malloc(sz), *v = malloc(sz);
    double *ph = malloc(sz), *s = malloc(sz), *sh = malloc(sz), *t = malloc(sz);
    /* r = b - A x */
    or_spmv(n, rp, ci, a, x, r);
    or_dscal(n, -1.0, r);
    or_daxpy(n, 1.0, b, r);
    or_dcopy(n, r, rt);
    memset(p, 0, sz);
    memset(v, 0, sz);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    double bb = or_ddot(n, b, b);
    double thresh = (eps * eps) * bb;
    double err = or_ddot(n, r, r);
    int its = 0;
    while (err > thresh && its < max_iter) {
        double rho_new = or_ddot(n, rt, r);
        if (rho_new == 0.0) break;
        double beta = (rho_new / rho) * (alpha / omega);
        /* p = r + beta (p - omega v) */
        or_daxpy(n, -omega, v, p);
        or_dscal(n, beta, p);
        or_daxpy(n, 1.0, r, p);
        precond(P, n, p, ph, 0);
        or_spmv(n, rp, ci, a, ph, v);
        double rtv = or_ddot(n, rt, v);
        if (rtv == 0.0) break;
        alpha = rho_new / rtv;
        /* s = r - alpha v */
        or_dcopy(n, r, s);
        or_daxpy(n, -alpha, v, s);
        precond(P, n, s, sh, 0);
        or_spmv(n, rp, ci, a, sh, t);
        double tt = or_ddot(n, t, t);
        or_daxpy(n, alpha, ph, x);
        ++its;
        if (tt == 0.0) {
            or_dcopy(n, s, r);
            err = or_ddot(n, r, r);
            break;
        }
        omega = or_ddot(n, t, s) / tt;
        or_daxpy(n, omega, sh, x);
        /* r = s - omega t */
        or_dcopy(n, s, r);
        or_daxpy(n, -omega, t, r);
        rho = rho_new;
        err = or_ddot(n, r, r);
        if (omega == 0.0) break;
    }
    if (out_rel2) *out_rel2 = (bb > 0.0) ? err / bb : err;
    free(r); free(rt); free(p); free(v); free(ph); free(s); free(sh); free(t);
    return its;
}

int or_bicgstab(uint32_t n, const uint32_t* rp, const uint32_t* ci,
                const double* a, const double* dinv, const double* b,
                double* x, double eps, int max_iter, double* out_rel2) {
    prec_t P = {dinv, 0, NULL, NULL, NULL, 0.0};
    return bicgstab_core(n, rp, ci, a, &P, b, x, eps, max_iter, out_rel2);
}

/* ---------------------------------------------------------------- GMRES */

/* Restarted GMRES(m) (Saad & Schultz 1986; Saad, "Iterative Methods for
 * Sparse Linear Systems", 2nd ed., Alg. 6.9 with Givens rotations as in
 * sec. 6.5.3), modified Gram-Schmidt Arnoldi, right preconditioning with
 * dinv (x = D^-1 u, A D^-1 u = b).  Stops when the Arnoldi residual
 * estimate |g_{j+1}|^2 <= eps^2 b.b, after max_iter Arnoldi steps in all, or
 * on a breakdown; a restart recomputes the true residual b - A x. */
static int gmres_core(uint32_t n, const uint32_t* rp, const uint32_t* ci,
                      const double* a, const prec_t* P, const double* b, double* x,
                      double eps, int max_iter, int m, double* out_rel2) {
    if (m < 1) m = 1;
    size_t sz = sizeof(double) * (size_t)n;
    double* V = malloc(sz * (size_t)(m + 1));
    double *w = malloc(sz), *z = malloc(sz), *u = malloc(sz);
    double* H = calloc((size_t)(m + 1) * (size_t)m, sizeof(double)); /* H[i*m + j] */
    double *cs = malloc(sizeof(double) * (size_t)m), *sn = malloc(sizeof(double) * (size_t)m);
    double *g = malloc(sizeof(double) * (size_t)(m + 1)), *y = malloc(sizeof(double) * (size_t)m);
    double bb = or_ddot(n, b, b);
    double thresh = (eps * eps) * bb;
    /* r = b - A x in V_0 */
    or_spmv(n, rp, ci, a, x, V);
    or_dscal(n, -1.0, V);
    or_daxpy(n, 1.0, b, V);
    double err = or_ddot(n, V, V);
    int its = 0;
    while (err > thresh && its < max_iter) {
        double beta = sqrt(err);
        or_dscal(n, 1.0 / beta, V);
        memset(g, 0, sizeof(double) * (size_t)(m + 1));
        g[0] = beta;
        int k = 0;
        int breakdown = 0;
        for (int j = 0; j < m && its < max_iter; ++j) {
            double* vj = V + (size_t)j * n;
            precond(P, n, vj, z, 0);
            or_spmv(n, rp, ci, a, z, w);
            for (int i = 0; i <= j; ++i) {
                double h = or_ddot(n, w, V + (size_t)i * n);
                H[i * m + j] = h;
                or_daxpy(n, -h, V + (size_t)i * n, w);
            }
            double hn = sqrt(or_ddot(n, w, w));
            H[(j + 1) * m + j] = hn;
            if (hn != 0.0) {
                or_dcopy(n, w, V + (size_t)(j + 1) * n);
                or_dscal(n, 1.0 / hn, V + (size_t)(j + 1) * n);
            } else {
                breakdown = 1;
            }
            for (int i = 0; i < j; ++i) {  /* earlier rotations on column j */
                double h0 = H[i * m + j], h1 = H[(i + 1) * m + j];
                H[i * m + j] = cs[i] * h0 + sn[i] * h1;
                H[(i + 1) * m + j] = -sn[i] * h0 + cs[i] * h1;
            }
            double h0 = H[j * m + j], h1 = H[(j + 1) * m + j];
            double den = hypot(h0, h1);
            cs[j] = den != 0.0 ? h0 / den : 1.0;
            sn[j] = den != 0.0 ? h1 / den : 0.0;
            H[j * m + j] = den;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++its;
            k = j + 1;
            if (g[j + 1] * g[j + 1] <= thresh || breakdown) break;
        }
        /* y = H[0:k,0:k]^-1 g[0:k]; x += D^-1 V y */
        for (int i = k - 1; i >= 0; --i) {
            double t = g[i];
            for (int l = i + 1; l < k; ++l) t -= H[i * m + l] * y[l];
            y[i] = H[i * m + i] != 0.0 ? t / H[i * m + i] : 0.0;
        }
        memset(u, 0, sz);
        for (int i = 0; i < k; ++i) or_daxpy(n, y[i], V + (size_t)i * n, u);
        precond(P, n, u, z, 0);
        or_daxpy(n, 1.0, z, x);
        or_spmv(n, rp, ci, a, x, V);
        or_dscal(n, -1.0, V);
        or_daxpy(n, 1.0, b, V);
        err = or_ddot(n, V, V);
        if (breakdown) break;
    }
    if (out_rel2) *out_rel2 = (bb > 0.0) ? err / bb : err;
    free(V); free(w); free(z); free(u); free(H); free(cs); free(sn); free(g); free(y);
    return its;
}

int or_gmres(uint32_t n, const uint32_t* rp, const uint32_t* ci,
             const double* a, const double* dinv, const double* b, double* x,
             double eps, int max_iter, int m, double* out_rel2) {
    prec_t P = {dinv, 0, NULL, NULL, NULL, 0.0};
    return gmres_core(n, rp, ci, a, &P, b, x, eps, max_iter, m, out_rel2);
}

int or_solve_ssor(int solver, uint32_t n, const uint32_t* rp, const uint32_t* ci,
                  const double* a, double omega, const double* b, double* x,
                  double eps, int max_iter, int m, double* out_rel2) {
    prec_t P = {NULL, 1, rp, ci, a, omega};
    if (solver == 0)
        return pcg_core(n, rp, ci, a, &P, b, x, eps, max_iter, decode(0), 1, out_rel2);
    if (solver == 1) return bicgstab_core(n, rp, ci, a, &P, b, x, eps, max_iter, out_rel2);
    return gmres_core(n, rp, ci, a, &P, b, x, eps, max_iter, m, out_rel2);
}

/* -------------------------------------------------------- dynamic matrix */

typedef struct {
    uint32_t size, cap;
    uint32_t* j;
    double* a;
} or_row;

struct or_dyn {
    uint32_t m, n;
    or_row* rows;
};

or_dyn* or_dyn_new(uint32_t m, uint32_t n) {
    or_dyn* M = (or_dyn*)calloc(1, sizeof(or_dyn));
    M->m = m;
    M->n = n;
    M->rows = (or_row*)calloc(m ? m : 1, sizeof(or_row));
    return M;
}

void or_dyn_free(or_dyn* M) {
    if (!M) return;
    for (uint32_t i = 0; i < M->m; ++i) {
        free(M->rows[i].j);
        free(M->rows[i].a);
    }
    free(M->rows);
    free(M);
}

void or_dyn_add(or_dyn* M, uint32_t i, uint32_t j, double a) {
    or_row* R = &M->rows[i];
    for (uint32_t k = 0; k < R->size; ++k) {
        if (R->j[k] == j) {
            R->a[k] = R->a[k] + a;
            return;
        }
    }
    if (R->size == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 4;
        R->j = (uint32_t*)realloc(R->j, sizeof(uint32_t) * R->cap);
        R->a = (double*)realloc(R->a, sizeof(double) * R->cap);
    }
    R->j[R->size] = j;
    R->a[R->size] = a;
    ++R->size;
}

uint64_t or_dyn_nnz(const or_dyn* M) {
    uint64_t s = 0;
    for (uint32_t i = 0; i < M->m; ++i) s += M->rows[i].size;
    return s;
}

void or_dyn_compress(const or_dyn* M, uint32_t* rowptr, uint32_t* colind,
                     double* val) {
    uint64_t pos = 0;
    rowptr[0] = 0;
    for (uint32_t i = 0; i < M->m; ++i) {
        const or_row* R = &M->rows[i];
        /* insertion sort by column: rows are short */
        uint64_t start = pos;
        for (uint32_t k = 0; k < R->size; ++k) {
            uint32_t j = R->j[k];
            double a = R->a[k];
            uint64_t q = pos;
            while (q > start && colind[q - 1] > j) {
                colind[q] = colind[q - 1];
                val[q] = val[q - 1];
                --q;
            }
            colind[q] = j;
            val[q] = a;
            ++pos;
        }
        rowptr[i + 1] = (uint32_t)pos;
    }
}

/* ------------------------------------------------------- context (rows) */

struct or_ctx {
    uint32_t nv;
    int least_squares;
    double* value;
    uint8_t* locked;
    uint32_t* free_index; /* variable -> free index, or UINT32_MAX */
    uint32_t nb_free;
    uint32_t cur_row;
    or_dyn* M;
    double* rhs;
    /* current row */
    uint32_t rsize, rcap;
    uint32_t* rvar;
    double* ra;
    double rrhs;
};

or_ctx* or_ctx_new(uint32_t nb_variables, int least_squares) {
    or_ctx* C = (or_ctx*)calloc(1, sizeof(or_ctx));
    C->nv = nb_variables;
    C->least_squares = least_squares;
    C->value = (double*)calloc(nb_variables ? nb_variables : 1, sizeof(double));
    C->locked = (uint8_t*)calloc(nb_variables ? nb_variables : 1, 1);
    C->free_index =
        (uint32_t*)calloc(nb_variables ? nb_variables : 1, sizeof(uint32_t));
    return C;
}

void or_ctx_free(or_ctx* C) {
    if (!C) return;
    free(C->value);
    free(C->locked);
    free(C->free_index);
    or_dyn_free(C->M);
    free(C->rhs);
    free(C->rvar);
    free(C->ra);
    free(C);
}

void or_ctx_set_variable(or_ctx* C, uint32_t i, double v) { C->value[i] = v; }
double or_ctx_get_variable(const or_ctx* C, uint32_t i) { return C->value[i]; }
void or_ctx_lock_variable(or_ctx* C, uint32_t i) { C->locked[i] = 1; }

void or_ctx_begin_matrix(or_ctx* C) {
    uint32_t f = 0;
    for (uint32_t i = 0; i < C->nv; ++i) {
        C->free_index[i] = C->locked[i] ? UINT32_MAX : f++;
    }
    C->nb_free = f;
    or_dyn_free(C->M);
    C->M = or_dyn_new(f, f);
    free(C->rhs);
    C->rhs = (double*)calloc(f ? f : 1, sizeof(double));
    C->cur_row = 0;
}

void or_ctx_begin_row(or_ctx* C) {
    C->rsize = 0;
    C->rrhs = 0.0;
}

void or_ctx_coefficient(or_ctx* C, uint32_t var, double a) {
    if (C->rsize == C->rcap) {
        C->rcap = C->rcap ? 2 * C->rcap : 8;
        C->rvar = (uint32_t*)realloc(C->rvar, sizeof(uint32_t) * C->rcap);
        C->ra = (double*)realloc(C->ra, sizeof(double) * C->rcap);
    }
    C->rvar[C->rsize] = var;
    C->ra[C->rsize] = a;
    ++C->rsize;
}

void or_ctx_right_hand_side(or_ctx* C, double v) { C->rrhs = v; }

/*
 * One equation row  sum_k a_k x_{v_k} = rhs.  Locked variables move to the
 * right-hand side in coefficient order: s = rhs - a_l x_l - ...  Then, for the
 * least-squares system, every ordered pair of free coefficients (f1, f2)
 * adds a_f1 a_f2 to M[f1][f2] and every free coefficient adds a_f s to
 * rhs[f]; for a square system the free coefficients form row cur_row.
 */
void or_ctx_end_row(or_ctx* C) {
    double s = C->rrhs;
    for (uint32_t k = 0; k < C->rsize; ++k) {
        uint32_t v = C->rvar[k];
        if (C->locked[v]) {
            double p = C->ra[k] * C->value[v];
            s = s - p;
        }
    }
    if (C->least_squares) {
        for (uint32_t k1 = 0; k1 < C->rsize; ++k1) {
            uint32_t f1 = C->free_index[C->rvar[k1]];
            if (f1 == UINT32_MAX) continue;
            for (uint32_t k2 = 0; k2 < C->rsize; ++k2) {
                uint32_t f2 = C->free_index[C->rvar[k2]];
                if (f2 == UINT32_MAX) continue;
                double p = C->ra[k1] * C->ra[k2];
                or_dyn_add(C->M, f1, f2, p);
            }
            double q = C->ra[k1] * s;
            C->rhs[f1] = C->rhs[f1] + q;
        }
    } else {
        uint32_t row = C->cur_row;
        if (row < C->nb_free) {
            for (uint32_t k = 0; k < C->rsize; ++k) {
                uint32_t f = C->free_index[C->rvar[k]];
                if (f == UINT32_MAX) continue;
                or_dyn_add(C->M, row, f, C->ra[k]);
            }
            C->rhs[row] = s;
        }
    }
    ++C->cur_row;
}

/* Rows r < nrows of arrays (entries [rp[r], rp[r+1]) of var/a, right-hand
 * side rhs[r], or 0 when rhs is NULL), each through begin_row /
 * coefficient / right_hand_side / end_row. */
void or_ctx_rows(or_ctx* C, uint32_t nrows, const uint32_t* rp, const uint32_t* var,
                 const double* a, const double* rhs) {
    for (uint32_t r = 0; r < nrows; ++r) {
        or_ctx_begin_row(C);
        for (uint32_t e = rp[r]; e < rp[r + 1]; ++e) or_ctx_coefficient(C, var[e], a[e]);
        if (rhs) or_ctx_right_hand_side(C, rhs[r]);
        or_ctx_end_row(C);
    }
}

uint32_t or_ctx_nb_free(const or_ctx* C) { return C->nb_free; }
uint64_t or_ctx_nnz(const or_ctx* C) { return or_dyn_nnz(C->M); }

void or_ctx_system(const or_ctx* C, uint32_t* rowptr, uint32_t* colind,
                   double* val, double* rhs) {
    or_dyn_compress(C->M, rowptr, colind, val);
    or_dcopy(C->nb_free, C->rhs, rhs);
}

int or_ctx_solve(or_ctx* C, double eps, int max_iter, double* out_rel2) {
    uint32_t n = C->nb_free;
    uint64_t nnz = or_dyn_nnz(C->M);
    uint32_t* rp = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint32_t* ci = (uint32_t*)malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
    double* a = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    double* b = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* x = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* dinv = (double*)malloc(sizeof(double) * (n ? n : 1));
    or_ctx_system(C, rp, ci, a, b);
    for (uint32_t v = 0; v < C->nv; ++v) {
        if (C->free_index[v] != UINT32_MAX) x[C->free_index[v]] = C->value[v];
    }
    or_jacobi(n, rp, ci, a, dinv);
    int its = or_pcg(n, rp, ci, a, dinv, b, x, eps, max_iter, 0, out_rel2);
    for (uint32_t v = 0; v < C->nv; ++v) {
        if (C->free_index[v] != UINT32_MAX) C->value[v] = x[C->free_index[v]];
    }
    free(rp); free(ci); free(a); free(b); free(x); free(dinv);
    return its;
}
