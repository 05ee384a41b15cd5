/*
 * OpenNL HIP extension for AMD Instinct MI355X (gfx950) -- libnlhip.so.
 *
 * Part 1 fills the role OpenNL's CUDA extension plays behind
 * nlInitExtension("CUDA") (state recorded as nl:CUDA in
 * lib/data/violin.obj:82): device NLBlas, device CRS matrices and a device
 * Jacobi preconditioner that nlSolveSystemIterative
 * (src/lib/OGF/scene_graph/NL/matrix.cpp:390-393) accepts in place of the
 * host ones.  Names follow the nlCUDA* pattern with HIP in place of CUDA;
 * upstream has no HIP extension, so these are this library's own exports.
 *
 * Part 2 is the re-entrant, handle-based API underneath: resident systems,
 * FP64 and mixed-precision solves, and their statistics.  Every function
 * returns NLHIP_OK (0) or a negative error code and never throws; the
 * message of the last error on the calling thread is nlhipLastError().
 */
#ifndef OPENNL_HIP_NL_HIP_H
#define OPENNL_HIP_NL_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "nl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Part 1: OpenNL extension entry points ----------------------------- */

/* nlInitExtension("HIP") also calls this. */
NLAPI NLboolean NLAPIENTRY nlInitExtension_HIP(void);
NLAPI NLboolean NLAPIENTRY nlExtensionIsInitialized_HIP(void);
/* Device-memory NLBlas: Malloc/Free/Memcpy on the GPU, BLAS-1 as kernels,
 * has_unique_memory = NL_FALSE. */
NLAPI NLBlas_t NLAPIENTRY nlHIPBlas(void);
/* Uploads a host CRS (or dynamic) matrix; the result is resident on the
 * device, its mult_func takes device pointers. */
NLAPI NLMatrix NLAPIENTRY nlHIPMatrixNewFromCRSMatrix(NLMatrix M);
/* Jacobi preconditioner of M, computed and kept on the device. */
NLAPI NLMatrix NLAPIENTRY nlHIPJacobiPreconditionerNewFromCRSMatrix(NLMatrix M);

/* ---- Part 2: handle-based API ------------------------------------------ */

#define NLHIP_OK               0
#define NLHIP_ERR_INVALID     -1   /* bad argument / malformed CRS          */
#define NLHIP_ERR_NO_DEVICE   -2   /* no usable gfx950 device               */
#define NLHIP_ERR_HIP         -3   /* HIP runtime error                     */
#define NLHIP_ERR_OOM         -4   /* device allocation failed              */
#define NLHIP_ERR_NONFINITE   -5   /* NaN / Inf in b or in the matrix       */
#define NLHIP_ERR_UNSUPPORTED -6   /* request outside what is implemented   */
#define NLHIP_ERR_BUDGET      -7   /* exceeds the configured memory budget  */
#define NLHIP_ERR_INJECTED    -8   /* failure forced by the test hook       */

#define NLHIP_PREC_FP64  0   /* FP64 Jacobi-PCG                              */
#define NLHIP_PREC_MIXED 1   /* FP32 inner PCG inside FP64 refinement; the
                                refinement keeps the iterate with the lowest
                                FP64 residual and ends back at it after 3
                                steps without a lower one, or when the
                                residual norm exceeds 100x the lowest       */
#define NLHIP_PREC_MIXED_FALLBACK 2 /* as MIXED, but a system whose FP32
                                       inner solve reaches inner_max_iter
                                       (> 0) without reaching inner_eps
                                       finishes with FP64 CG from its current
                                       iterate (LAPACK DSGESV's policy), as
                                       does one whose refinement ends as
                                       above (from its best iterate);
                                       single systems and batches; no
                                       counterpart in the reference */

#define NLHIP_SYSTEM_FP32 1u /* also keep an FP32 copy of the values         */
#define NLHIP_SYSTEM_NO_OFFSETS 2u /* read 32-bit column indices only (no
                                      16-bit segment offsets, no 2x2 blocks) */
#define NLHIP_SYSTEM_NO_BLOCKS 4u  /* no 2x2 block form (kept otherwise when
                                      rows 2I, 2I+1 share columns that come
                                      in pairs 2J, 2J+1, as in LSCM)       */

typedef struct nlhip_system_s* nlhip_system;

typedef struct {
    int32_t precision;        /* NLHIP_PREC_*                                */
    int32_t precond;          /* NL_PRECOND_JACOBI or NL_PRECOND_NONE        */
    double eps;               /* stop when |Ax-b| <= eps |b|                 */
    uint32_t max_iter;        /* iteration cap (FP64 CG / total inner)       */
    uint32_t check_every;     /* iterations per captured graph (0: auto)     */
    double inner_eps;         /* mixed: relative tolerance of each inner CG  */
    uint32_t max_outer;       /* mixed: refinement steps cap                 */
    uint32_t inner_max_iter;  /* mixed: cap per inner CG (0: max_iter)       */
    uint32_t confirm;         /* FP64 CG: when the recursive residual meets
                                 eps, recompute |b - A x| in FP64; while it
                                 is still above eps |b|, run another round:
                                 the same recurrence continued with the
                                 recursive threshold lowered by the measured
                                 gap (bitwise the iterates of a solve with a
                                 smaller threshold) when the true residual
                                 is within 2x of the recursive one, else a
                                 restart from the current iterate with the
                                 residual recomputed (residual replacement,
                                 for drift from rounding on long solves);
                                 stops when the true residual meets eps, at
                                 max_iter, after 8 rounds, or when a
                                 restart no longer lowers it (default 1;
                                 0: stop on the recursive residual alone, as
                                 OpenNL's CG does).  Batched and multi-RHS
                                 solves continue only (no restart).          */
} nlhip_solve_params;

typedef struct {
    uint32_t iterations;      /* CG iterations (mixed: sum of inner)         */
    uint32_t outer_iterations;/* mixed: refinement steps; FP64 CG: restarts
                                 of the confirmation                         */
    int32_t converged;        /* 1 when |Ax-b| <= eps |b| was reached        */
    int32_t status;           /* NLHIP_OK or error code                      */
    double rel_residual;      /* |b-Ax|/|b| recomputed in FP64 at the end    */
    double cg_rel_residual;   /* the solver's own recursive residual         */
    double device_seconds;    /* HIP-event time of the solve on the device   */
    double wall_seconds;      /* host time including staging                 */
    uint32_t extra_iterations;/* iterations run after the recursive residual
                                 first met eps (confirm); iterations minus
                                 this is where OpenNL's stopping rule stops  */
    uint32_t reserved;
} nlhip_solve_stats;

NLAPI void NLAPIENTRY nlhipSolveParamsDefault(nlhip_solve_params* p);

/* Copies a CRS matrix (rowptr[n+1], colind[nnz], val[nnz]; host or device
 * pointers) to the device.  Checks the structure (monotone rowptr, colind <
 * n) and that the values are finite. */
NLAPI int NLAPIENTRY nlhipSystemCreate(uint32_t n, const uint32_t* rowptr,
                                       const uint32_t* colind, const double* val,
                                       uint32_t flags, nlhip_system* out);
NLAPI int NLAPIENTRY nlhipSystemDestroy(nlhip_system s);
NLAPI int NLAPIENTRY nlhipSystemInfo(nlhip_system s, uint32_t* n, uint64_t* nnz,
                                     uint64_t* device_bytes);
/* Generalized symmetric eigenpairs A x = lambda M x: the nb_eigens
 * smallest, `shift` a lower bound of the spectrum (A - shift M positive
 * definite, as for a shift-invert solve).  Replaces the ARPACK shift-invert
 * solves behind spectral LSCM and manifold harmonics
 * (mesh_grob_surface_commands.cpp:761-763, mesh_grob_spectral_commands.cpp:
 * 74-116), by Chebyshev-filtered subspace iteration on the device (no
 * linear solves).  A symmetric; mass: the diagonal of M (n positive values,
 * host or device), NULL for the identity.
 * nb_eigens in [1, 24]; eps bounds |A x - lambda M x| / (max(|lambda|,
 * |shift|) |M x|) for each returned pair, or the residual is at the
 * rounding level of the operator (1000 u |M^-1 A| |M x|, for eigenvalues
 * near 0); max_iter outer iterations (0: 1000), each a degree-40 filter.  eigenvalues (nb_eigens, ascending) and eigenvectors (n x nb_eigens
 * column-major, M-orthonormal, host or device, may be NULL). */
typedef struct {
    uint32_t outer_iterations;
    uint32_t converged;        /* pairs below eps                       */
    uint64_t inner_iterations; /* filter degrees (block SpMMs), all steps */
    double max_rel_residual;
    double seconds;
} nlhip_eigen_stats;
NLAPI int NLAPIENTRY nlhipEigenSolve(nlhip_system a, const double* mass, uint32_t nb_eigens,
                                     double shift, double eps, uint32_t max_iter,
                                     double* eigenvalues, double* eigenvectors,
                                     nlhip_eigen_stats* stats);

/* Copies the resident CRS out (rowptr[n+1], colind[nnz], val[nnz]; host or
 * device pointers, any may be null). */
NLAPI int NLAPIENTRY nlhipSystemGetCRS(nlhip_system s, uint32_t* rowptr, uint32_t* colind,
                                       double* val);

/* Least-squares normal equations assembled on the device: the system a
 * least-squares OpenNL context builds from the same rows (replaces the
 * per-row accumulation of nlEnd(NL_ROW) under NL_LEAST_SQUARES and the
 * compression of nlEnd(NL_MATRIX); geogram's nl_context / NLRowColumn code,
 * driven at voxel_grob_transport_commands.cpp:236-279 and by the LSCM calls
 * at mesh_grob_surface_commands.cpp:724-779).  Row r is
 *   sum over e in [row_ptr[r], row_ptr[r+1]) of coef[e] x[var[e]] = row_rhs[s][r]
 * for each of nb_systems right-hand sides (row_rhs: nb_systems x nb_rows, or
 * NULL for zeros).  Locked variables (locked[v] != 0; NULL: none) keep
 * values[s][v] (nb_systems x nb_variables) and move to the right-hand side.
 * The system is over the free variables in index order; rhs (nb_systems x
 * nb_free, host or device, may be NULL) receives its right-hand sides, and
 * nb_free their number.  Every sum is formed in the order the context forms
 * it, so the CRS and rhs are bitwise those of the host assembly.  Inputs are
 * host or device arrays; flags as nlhipSystemCreate. */
NLAPI int NLAPIENTRY nlhipLeastSquaresCreate(uint32_t nb_variables, uint32_t nb_rows,
                                             const uint32_t* row_ptr, const uint32_t* var,
                                             const double* coef, uint32_t nb_systems,
                                             const double* row_rhs, const uint8_t* locked,
                                             const double* values, uint32_t flags,
                                             nlhip_system* out, double* rhs,
                                             uint32_t* nb_free);
/* The LSCM system of a triangle mesh built entirely on the device: per
 * non-degenerate triangle the two conformality rows of Levy et al. 2002
 * (what geogram's LSCM hands to a least-squares context for
 * mesh_make_atlas / parameterize_chart, mesh_grob_surface_commands.cpp:
 * 724-779), vertices pin0 and pin1 locked at (u0, v0) and (u1, v1), then
 * nlhipLeastSquaresCreate's assembly.  Variables are (u_v, v_v) = (2v, 2v+1);
 * the system is over the free ones in that order; rhs (nb_free) may be NULL.
 * points: nb_vertices x 3, triangles: nb_triangles x 3; host or device. */
NLAPI int NLAPIENTRY nlhipLSCMCreate(uint32_t nb_vertices, const double* points,
                                     uint32_t nb_triangles, const uint32_t* triangles,
                                     uint32_t pin0, double u0, double v0, uint32_t pin1,
                                     double u1, double v1, uint32_t flags, nlhip_system* out,
                                     double* rhs, uint32_t* nb_free);
/* Launch shape of the SpMV engine: rows per wavefront (2, 4, 8, 16: that many
 * rows share a 64-lane wavefront; 64: one row per lane, nonzeros streamed
 * through LDS, the default) and workgroups per launch.  65..79 select
 * variants of the streamed engine for launch-shape sweeps (chunk width,
 * load kind, column format, tile order, 2x2 blocks); they give the same sums
 * as 64. */
NLAPI int NLAPIENTRY nlhipSystemSetLaunch(nlhip_system s, int rows_per_wave,
                                          uint32_t workgroups);
NLAPI int NLAPIENTRY nlhipSystemGetLaunch(nlhip_system s, int* rows_per_wave,
                                          uint32_t* workgroups);
/* Autotuner: times `iters` CG iterations (0: 20) for each streamed SpMV
 * engine (2x2 blocks with 1, 2, 4, 8 blocks per lane per chunk; scalar with
 * 16/32-bit columns, contiguous or XCD-interleaved; same row sums) and 1, 2,
 * 4, 8 workgroups per CU, and keeps the fastest as the system's launch
 * shape (returned, with the seconds per iteration).  The workgroup count
 * fixes the order of the dot-product sums, so a tuned system's solves are
 * reproducible; they stay within the solver's tolerance of the untuned. */
NLAPI int NLAPIENTRY nlhipSystemAutotune(nlhip_system s, int fp32, uint32_t iters,
                                         int* rows_per_wave, uint32_t* workgroups,
                                         double* seconds_per_iteration);
/* Solves A x = b; x holds the initial guess.  b, x host or device. */
NLAPI int NLAPIENTRY nlhipSolve(nlhip_system s, const double* b, double* x,
                                const nlhip_solve_params* params,
                                nlhip_solve_stats* stats);
/* A solve advanced in slices (FP64 Jacobi-PCG, or plain CG with
 * NL_PRECOND_NONE): nlhipSolveBegin stages b and the initial guess x0 (host
 * or device) and forms the initial residual; each nlhipSolveAdvance runs at
 * most `iters` more iterations of the same recurrence and sets *finished
 * once the solve has stopped (stopping test, with params->confirm as for
 * nlhipSolve, or max_iter); nlhipSolveEnd copies the iterate into x (host or
 * device, may be NULL) and ends the solve.  The iterates, iteration count and
 * result are bitwise those of one nlhipSolve call with the same parameters,
 * however the iterations are sliced.  One sliced solve per system at a time;
 * nlhipSolve on the system in between ends it.  stats (may be NULL) describe
 * the solve so far.  No counterpart in the reference: a caller uses it to
 * bound the time of one call (time-sliced serving) or to read the iterate
 * of a long solve between slices. */
NLAPI int NLAPIENTRY nlhipSolveBegin(nlhip_system s, const double* b, const double* x0,
                                     const nlhip_solve_params* params);
NLAPI int NLAPIENTRY nlhipSolveAdvance(nlhip_system s, uint32_t iters, int* finished,
                                       nlhip_solve_stats* stats);
NLAPI int NLAPIENTRY nlhipSolveEnd(nlhip_system s, double* x, nlhip_solve_stats* stats);
/* nlhipSolve without any allocation, for callers that cannot allocate
 * device or pinned memory (serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  FP64 CG only (Jacobi or no
 * preconditioner, params->confirm as for nlhipSolve); b and x host or device.
 * Every work buffer lives in one fixed device array of the library, sized for
 * systems of at most NLHIP_NOALLOC_MAX_ROWS rows (larger ones are rejected
 * with NLHIP_ERR_INVALID) and held by each GPU the library runs on until the
 * process ends; the kernels are launched directly (no graph instantiation).
 * The result, iteration count and residuals are bitwise nlhipSolve's;
 * device_seconds is host wall time.  Calls that overlap on one GPU (any
 * systems) share the array and race: serialise them. */
#define NLHIP_NOALLOC_MAX_ROWS 2097152u
NLAPI int NLAPIENTRY nlhipSolveNoAlloc(nlhip_system s, const double* b, double* x,
                                       const nlhip_solve_params* params,
                                       nlhip_solve_stats* stats);
/* k right-hand sides at once (1 <= k <= 32), sharing every pass over the
 * matrix: B and X are column-major n x k (system j at B + j n, the layout of
 * OpenNL's NL_NB_SYSTEMS); X holds the initial guesses.  Each column has its
 * own scalars and stopping test.  per_rhs (k entries) and total may be null;
 * total.iterations is the largest count. */
NLAPI int NLAPIENTRY nlhipSolveMulti(nlhip_system s, uint32_t k, const double* B,
                                     double* X, const nlhip_solve_params* params,
                                     nlhip_solve_stats* per_rhs,
                                     nlhip_solve_stats* total);
/* y = A x (FP64), host or device pointers. */
NLAPI int NLAPIENTRY nlhipSpMV(nlhip_system s, const double* x, double* y);
/* Times `reps` back-to-back FP64 (fp32 = 0) or FP32 SpMVs on device-resident
 * vectors; returns the mean seconds per SpMV measured with HIP events. */
NLAPI int NLAPIENTRY nlhipBenchSpMV(nlhip_system s, int fp32, uint32_t reps,
                                    double* seconds_per_spmv);

/* Times `reps` plain device-to-device copies of `bytes` (a multiple of 16)
 * between two buffers it allocates, for a few grid sizes; returns the mean
 * seconds per copy of the fastest (HIP events).  The bandwidth reference the
 * solver kernels are compared with. */
NLAPI int NLAPIENTRY nlhipBenchCopy(uint64_t bytes, uint32_t reps, double* seconds_per_copy);
/* Times `reps` read-only passes over `bytes` (a multiple of 16) of a buffer it
 * allocates (16-byte loads: grid-stride, plain and non-temporal, or one
 * contiguous run per wavefront; a few grid sizes each); returns the mean
 * seconds per pass of the fastest.  The bandwidth reference for
 * read-dominated kernels such as the SpMV. */
NLAPI int NLAPIENTRY nlhipBenchRead(uint64_t bytes, uint32_t reps, double* seconds_per_pass);
/* Times `reps` passes of the 2x2-block SpMV's memory traffic without its
 * arithmetic: `blocks` 32-byte blocks (in the block engine's grouped
 * layout) and 2-byte column offsets streamed, a
 * 16-byte gather per block from an L2-resident vector, and per row pair
 * (every blocks_per_row blocks) a 16-byte read and, when writes != 0, a
 * 16-byte write; each wavefront streams one contiguous run; returns the
 * mean seconds per pass of the fastest of a few grid sizes.  The attainable
 * rate for the SpMV's access mix, which the read-only and copy references
 * bracket.  No counterpart in the reference. */
NLAPI int NLAPIENTRY nlhipBenchStreamMix(uint64_t blocks, uint32_t blocks_per_row, int writes,
                                         uint32_t reps, double* seconds_per_pass);

/* Runs `iters` PCG iterations (FP64, or FP32 when fp32 = 1, b = 1) launching
 * each kernel directly; seconds[0..2] receive the mean HIP-event duration of
 * the direction, SpMV and update launches. */
NLAPI int NLAPIENTRY nlhipProfileCG(nlhip_system s, int fp32, uint32_t iters,
                                    double* seconds);

/* ---- batched systems (texture-atlas charts) ------------------------------
 * Many independent systems solved together (mesh_make_atlas solves one LSCM
 * system per chart, src/lib/OGF/mesh/commands/mesh_grob_surface_commands.cpp:
 * 724-730).  The systems come packed as one block-diagonal CRS: system s owns
 * rows and columns [row_off[s], row_off[s+1]); rowptr and colind use global
 * indices.  Every entry must stay inside its own system's block.  Each
 * system's result is bitwise independent of the other systems in the batch.
 */
typedef struct nlhip_batch_s* nlhip_batch;

NLAPI int NLAPIENTRY nlhipBatchCreate(uint32_t nsys, const uint32_t* row_off,
                                      const uint32_t* rowptr, const uint32_t* colind,
                                      const double* val, uint32_t flags,
                                      nlhip_batch* out);
/* A texture atlas built on the device (mesh_make_atlas, mesh_grob_surface_
 * commands.cpp:724-730): the LSCM systems of nb_charts charts of one mesh,
 * as nlhipLSCMCreate builds each (rows of every non-degenerate triangle,
 * two pinned vertices per chart, device least-squares assembly), packed as
 * a batch.  Chart c owns vertices [vertex_off[c], vertex_off[c+1]) (host
 * array) and the triangles over them; points (3 per vertex) and triangles
 * (3 global vertex ids each) host or device; pins: 2 global vertex ids per
 * chart, pin_uv: (u0, v0, u1, v1) per chart (host).  System c is over the
 * free (u, v) of chart c in vertex order; rhs (n values, host or device,
 * may be NULL) receives the right-hand sides and n_out their number. */
NLAPI int NLAPIENTRY nlhipBatchLSCMCreate(uint32_t nb_charts, const uint32_t* vertex_off,
                                          const double* points, uint32_t nb_triangles,
                                          const uint32_t* triangles, const uint32_t* pins,
                                          const double* pin_uv, uint32_t flags,
                                          nlhip_batch* out, double* rhs, uint32_t* n_out);
NLAPI int NLAPIENTRY nlhipBatchDestroy(nlhip_batch b);
NLAPI int NLAPIENTRY nlhipBatchInfo(nlhip_batch b, uint32_t* nsys, uint32_t* n,
                                    uint64_t* nnz, uint32_t* tiles);
/* The packed CRS the batch holds, copied out (each pointer host or device,
 * may be NULL): row_off nsys + 1, rowptr n + 1, colind and val nnz values.
 * No counterpart in the reference (its matrices are host arrays); this is how
 * a caller inspects a batch the device assembled. */
NLAPI int NLAPIENTRY nlhipBatchGetCRS(nlhip_batch b, uint32_t* row_off, uint32_t* rowptr,
                                      uint32_t* colind, double* val);
/* Systems the last NLHIP_PREC_MIXED_FALLBACK solve finished in FP64. */
NLAPI int NLAPIENTRY nlhipBatchFallbacks(nlhip_batch b, uint32_t* n);
/* on = 0: solves enqueue every kernel launch directly instead of replaying
 * captured graphs (for tracing tools that cannot follow graph launches);
 * the same launches in the same order, so the same bits.  No counterpart in
 * the reference. */
NLAPI int NLAPIENTRY nlhipBatchSetGraphs(nlhip_batch b, int on);
/* on = 1: each run of iterations of the batched PCG (and of the FP32 inner
 * solves of its refinement) is one persistent cooperative launch, the
 * direction, SpMV and update phases separated by grid-wide barriers instead
 * of kernel boundaries, on as many workgroups as the device holds at once
 * (at most the batch's workgroup count).  Same kernels' arithmetic, same
 * bits as the three-launch path; on = 0 (default): captured graphs of three
 * launches per iteration. */
NLAPI int NLAPIENTRY nlhipBatchSetPersistent(nlhip_batch b, int on);
/* Workgroups of the grid that walks the active tiles. */
NLAPI int NLAPIENTRY nlhipBatchSetLaunch(nlhip_batch b, uint32_t workgroups);
/* Passes of a tile (64 rows or block rows each, 4 per tile) that one
 * wavefront of the SpMV streams: 1, 2 or 4, or 0 (the default) to choose
 * per launch from the number of active tiles.  Results do not depend on it. */
NLAPI int NLAPIENTRY nlhipBatchSetSpmvPasses(nlhip_batch b, uint32_t passes);
/* Autotuner: times `iters` (0: 50) FP64 iterations on rhs for workgroup
 * counts of 4, 16, 64 per CU and each SpMV run length, keeps the fastest
 * (returned, with seconds per iteration).  No result bit depends on it. */
NLAPI int NLAPIENTRY nlhipBatchAutotune(nlhip_batch b, const double* rhs, uint32_t iters,
                                        uint32_t* workgroups, uint32_t* passes,
                                        double* seconds);
/* Solves every system; b and x are the concatenated vectors (x holds the
 * initial guesses).  per_system (nsys entries) and total may be null; in
 * total, iterations is the largest count and outer_iterations the sum. */
NLAPI int NLAPIENTRY nlhipBatchSolve(nlhip_batch b, const double* rhs, double* x,
                                     const nlhip_solve_params* params,
                                     nlhip_solve_stats* per_system,
                                     nlhip_solve_stats* total);

/* ---- solve server ----------------------------------------------------------
 * Requests (one system, 1..32 right-hand sides) are copied at submission,
 * admitted first-come first-served when a slot is free and their device
 * footprint fits in the remaining budget, and run concurrently, each on its
 * own resident system and stream.  A request larger than the whole budget
 * finishes at once with NLHIP_ERR_BUDGET.  A failing request does not affect
 * the others.  Times are seconds since the server was created.
 */
typedef struct nlhip_server_s* nlhip_server;

typedef struct {
    uint64_t hbm_budget_bytes;  /* default 200 GiB (capped at the device's HBM) */
    uint32_t max_active;        /* concurrent requests (slots), default 8       */
} nlhip_server_config;

typedef struct {
    uint64_t id;
    int32_t status;
    uint32_t k;
    uint32_t n;
    uint64_t device_bytes;      /* footprint charged against the budget         */
    nlhip_solve_stats stats;    /* for k > 1: worst column, largest count       */
    double t_submit, t_admit, t_done;
} nlhip_request_result;

typedef struct {
    uint64_t budget_bytes, in_use_bytes, peak_in_use_bytes;
    uint32_t pending, running, peak_running;
    uint64_t completed, failed;
} nlhip_server_info;

NLAPI void NLAPIENTRY nlhipServerConfigDefault(nlhip_server_config* c);
NLAPI int NLAPIENTRY nlhipServerCreate(const nlhip_server_config* config,
                                       nlhip_server* out);
/* B (and X0, may be null) column-major n x k. */
NLAPI int NLAPIENTRY nlhipServerSubmit(nlhip_server s, uint32_t n, const uint32_t* rowptr,
                                       const uint32_t* colind, const double* val,
                                       uint32_t k, const double* B, const double* X0,
                                       const nlhip_solve_params* params, uint64_t* id);
/* A request against a system the caller keeps resident (nlhipSystemCreate,
 * nlhipLSCMCreate, nlhipLeastSquaresCreate, ...): only the k right-hand sides
 * B (column-major n x k; X0 may be NULL) travel, the matrix is not uploaded
 * again, and the request is charged only its solve workspace against the
 * budget.  Requests on one system run one at a time in arrival order; while
 * the system is busy later requests may be admitted first.  The result is
 * bitwise that of nlhipSolve (k = 1) or nlhipSolveMulti on the system.  The
 * system must stay alive, and no other call may use it concurrently, until
 * nlhipServerWait has returned for every request on it.  No counterpart in
 * the reference: its callers re-solve one matrix many times (the xyz
 * smoothing of mesh_grob_surface_commands.cpp:687-705, the Newton loop of
 * Transport_2d_04.lua:18-32). */
NLAPI int NLAPIENTRY nlhipServerSubmitSystem(nlhip_server s, nlhip_system system, uint32_t k,
                                             const double* B, const double* X0,
                                             const nlhip_solve_params* params, uint64_t* id);
/* Blocks until request `id` is finished (timeout_seconds <= 0: no limit),
 * copies its solution into X (n x k, may be null) and releases it.  Returns
 * the request's own status. */
NLAPI int NLAPIENTRY nlhipServerWait(nlhip_server s, uint64_t id, double* X,
                                     nlhip_request_result* result,
                                     double timeout_seconds);
NLAPI int NLAPIENTRY nlhipServerInfo(nlhip_server s, nlhip_server_info* info);
/* Finishes the admitted and waiting requests, then stops the workers. */
NLAPI int NLAPIENTRY nlhipServerDestroy(nlhip_server s);

NLAPI const char* NLAPIENTRY nlhipLastError(void);
/* Test hook: the next `count` solves fail with NLHIP_ERR_INJECTED before
 * touching the device (other requests are unaffected). */
NLAPI void NLAPIENTRY nlhipInjectFailures(int count);
/* Device name, CU count and HBM bytes of the device in use. */
NLAPI int NLAPIENTRY nlhipDeviceInfo(char* name, size_t name_len, int* cu_count,
                                     uint64_t* hbm_bytes);

#ifdef __cplusplus
}
#endif

#endif /* OPENNL_HIP_NL_HIP_H */
