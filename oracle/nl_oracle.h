/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C, single-threaded (optionally OpenMP) restatement of the OpenNL
 * solve path that Graphite calls through its OGF::NL wrapper:
 * CRS storage and SpMV, the NLBlas level-1 routines, the Jacobi and SSOR
 * preconditioners, preconditioned CG, BiCGSTAB and GMRES(m), the dynamic
 * matrix, and the least-squares context API (normal equations with locked
 * variables moved to the RHS).
 *
 * Parity unpinned against OpenNL itself (its source is absent; each is
 * pinned by the mathematics instead, see nl_oracle.c): the Jacobi fallback
 * for a zero diagonal, GMRES's use of the preconditioner, and the summation
 * order inside OpenNL's SSOR sweeps.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline.  The product
 * (graphitethree_amd/lib/libnlhip.so) never links or calls it.
 *
 * The upstream OpenNL sources (geogram/src/lib/geogram/NL) are an empty
 * submodule in the reference snapshot (.gitmodules:1-4), so this file follows
 * the reference's call sites and textbook algorithms; every function names
 * what it follows.  Arithmetic is compiled with -ffp-contract=off, so every
 * element-wise update is a separately rounded multiply then add, and every
 * reduction is a left-to-right sum unless the name says _omp.
 */
#ifndef NL_ORACLE_H
#define NL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- CRS SpMV: y = A x, rows summed left to right.
 * Follows nlMultMatrixVector (scene_graph/NL/matrix.cpp:303-305). */
void or_spmv(uint32_t m, const uint32_t* rowptr, const uint32_t* colind,
             const double* val, const double* x, double* y);
/* Same per-row arithmetic, rows split into static OpenMP slices. */
void or_spmv_omp(uint32_t m, const uint32_t* rowptr, const uint32_t* colind,
                 const double* val, const double* x, double* y);

/* ---- NLBlas level 1, stride 1 (scene_graph/NL/blas.cpp:86,96,115,125,144) */
void   or_dcopy(uint32_t n, const double* x, double* y);
void   or_dscal(uint32_t n, double a, double* x);
double or_ddot(uint32_t n, const double* x, const double* y);
double or_dnrm2(uint32_t n, const double* x);
void   or_daxpy(uint32_t n, double a, const double* x, double* y);
double or_ddot_omp(uint32_t n, const double* x, const double* y);

/* ---- Jacobi preconditioner (scene_graph/NL/matrix.cpp:363-365):
 * dinv[i] = 1/a_ii, or 1 when the row has no (or a zero) diagonal entry. */
void or_jacobi(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
               const double* val, double* dinv);

/* ---- Preconditioned CG (nlSolveSystemIterative with NL_CG,
 * scene_graph/NL/matrix.cpp:390-393).  dinv == NULL means no preconditioner.
 * Stops when r.r <= eps^2 b.b (matrix.h:177-179) or after max_iter
 * iterations; x holds the initial guess on entry.  variant: 0 serial,
 * 1 OpenMP SpMV with serial BLAS-1, 2 OpenMP SpMV and BLAS-1.
 * Returns the iteration count; *out_err receives the final r.r / b.b. */
int or_pcg(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
           const double* val, const double* dinv, const double* b, double* x,
           double eps, int max_iter, int variant, double* out_rel2);

/* Runs exactly `iters` PCG iterations with no stopping test (CPU baseline
 * sampling); returns seconds spent in the loop. */
double or_pcg_timed_iters(uint32_t n, const uint32_t* rowptr,
                          const uint32_t* colind, const double* val,
                          const double* dinv, const double* b, double* x,
                          int iters, int variant);

/* ---- BiCGSTAB (NL_BICGSTAB, matrix.cpp:380-382), van der Vorst 1992,
 * right-preconditioned with dinv (NULL = none).  Same stopping rule. */
int or_bicgstab(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
                const double* val, const double* dinv, const double* b,
                double* x, double eps, int max_iter, double* out_rel2);

/* ---- GMRES(m) (NL_GMRES, matrix.cpp:377-379, with inner_iter = m = 5 at
 * matrix.cpp:385): restarted, modified Gram-Schmidt, right-preconditioned
 * with dinv (NULL = none); returns the number of Arnoldi steps.  Upstream's
 * own GMRES (geogram, absent here) is not available to compare with: this
 * is the textbook method. */
int or_gmres(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
             const double* val, const double* dinv, const double* b, double* x,
             double eps, int max_iter, int m, double* out_rel2);

/* ---- SSOR preconditioner (nlNewSSORPreconditioner, omega = 1.5 at
 * matrix.cpp:366-368): y = omega (2 - omega) (D + omega U)^-1 D
 * (D + omega L)^-1 x, as two triangular sweeps over the CRS rows (see
 * nl_oracle.c).  The diagonal must be nonzero. */
void or_ssor(uint32_t n, const uint32_t* rp, const uint32_t* ci,
             const double* a, double omega, const double* x, double* y);

/* CG (solver 0), BiCGSTAB (1) or GMRES(m) (2) as above, preconditioned by
 * SSOR(omega) instead of Jacobi. */
int or_solve_ssor(int solver, uint32_t n, const uint32_t* rp, const uint32_t* ci,
                  const double* a, double omega, const double* b, double* x,
                  double eps, int max_iter, int m, double* out_rel2);

/* ---- Dynamic matrix: duplicate (i,j) entries accumulate
 * (scene_graph/NL/matrix.h:111-112,122-123), then compression to CRS with
 * each row's columns in increasing order. */
typedef struct or_dyn or_dyn;
or_dyn*  or_dyn_new(uint32_t m, uint32_t n);
void     or_dyn_free(or_dyn* M);
void     or_dyn_add(or_dyn* M, uint32_t i, uint32_t j, double a);
uint64_t or_dyn_nnz(const or_dyn* M);
/* Writes rowptr[m+1], colind[nnz], val[nnz]. */
void     or_dyn_compress(const or_dyn* M, uint32_t* rowptr, uint32_t* colind,
                         double* val);

/* ---- Least-squares / row-system context (the nlBegin/nlEnd(NL_ROW),
 * nlCoefficient, nlSetVariable, nlLockVariable sequence of
 * plugins/OGF/WarpDrive/commands/voxel_grob_transport_commands.cpp:231-287).
 * least_squares != 0: each row sum_j a_j x_j = rhs contributes a^T a and
 * a^T (rhs - sum_locked a_l x_l) to the normal equations.  Otherwise each row
 * is one row of a square system over the free variables.  Free variables are
 * numbered in increasing variable order. */
typedef struct or_ctx or_ctx;
or_ctx*  or_ctx_new(uint32_t nb_variables, int least_squares);
void     or_ctx_free(or_ctx* C);
void     or_ctx_set_variable(or_ctx* C, uint32_t i, double v);
double   or_ctx_get_variable(const or_ctx* C, uint32_t i);
void     or_ctx_lock_variable(or_ctx* C, uint32_t i);
/* Call after all set/lock and before the first row. */
void     or_ctx_begin_matrix(or_ctx* C);
void     or_ctx_begin_row(or_ctx* C);
void     or_ctx_coefficient(or_ctx* C, uint32_t var, double a);
void     or_ctx_right_hand_side(or_ctx* C, double v);
void     or_ctx_end_row(or_ctx* C);
void     or_ctx_rows(or_ctx* C, uint32_t nrows, const uint32_t* rp, const uint32_t* var,
                     const double* a, const double* rhs);
uint32_t or_ctx_nb_free(const or_ctx* C);
uint64_t or_ctx_nnz(const or_ctx* C);
/* Compressed system over the free variables. */
void     or_ctx_system(const or_ctx* C, uint32_t* rowptr, uint32_t* colind,
                       double* val, double* rhs);
/* Solves with Jacobi-PCG from the current variable values; writes the
 * solution back into the free variables; returns iterations. */
int      or_ctx_solve(or_ctx* C, double eps, int max_iter, double* out_rel2);

#ifdef __cplusplus
}
#endif

#endif
