/*
 * OpenNL C API subset, MI355X-native (libnlhip.so).
 *
 * Declares the part of OpenNL's C API that Graphite calls on its solve path,
 * with the names, argument meanings and ownership rules of those call sites,
 * so that they compile and link against this library unchanged:
 *
 *   src/lib/OGF/scene_graph/NL/matrix.cpp:121-397   matrix + solver API
 *   src/lib/OGF/scene_graph/NL/blas.cpp:66-146      NLBlas table
 *   plugins/OGF/WarpDrive/commands/voxel_grob_transport_commands.cpp:231-287
 *                                                   context (least squares) API
 *
 * The upstream headers (geogram/src/lib/geogram/NL/nl.h, nl_matrix.h,
 * nl_blas.h, nl_iterative_solvers.h, nl_preconditioners.h) are not in the
 * reference snapshot: enum VALUES here are this library's own, so the drop-in
 * is at source level (recompile against this header), not binary level.
 *
 * Every solve and every matrix-vector product runs on the GPU.  There is no
 * CPU solver behind this API: when no gfx950 device is usable, the calls that
 * compute report an error (NL_FALSE / 0 iterations / message on stderr).
 */
#ifndef OPENNL_HIP_NL_H
#define OPENNL_HIP_NL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLAPI
#define NLAPIENTRY

typedef unsigned int   NLenum;
typedef unsigned char  NLboolean;
typedef unsigned int   NLbitfield;
typedef void           NLvoid;
typedef signed char    NLbyte;
typedef short          NLshort;
typedef int            NLint;
typedef unsigned char  NLubyte;
typedef unsigned short NLushort;
typedef unsigned int   NLuint;
typedef long           NLlong;
typedef unsigned long  NLulong;
typedef int            NLsizei;
typedef float          NLfloat;
typedef double         NLdouble;
typedef void*          NLContext;

#define NL_FALSE 0x0
#define NL_TRUE  0x1

/* ---- nlSolverParameter / nlGet* names ---------------------------------- */
#define NL_SOLVER           0x100
#define NL_NB_VARIABLES     0x101
#define NL_LEAST_SQUARES    0x102
#define NL_MAX_ITERATIONS   0x103
#define NL_THRESHOLD        0x104
#define NL_OMEGA            0x105
#define NL_SYMMETRIC        0x106
#define NL_USED_ITERATIONS  0x107
#define NL_ERROR            0x108
#define NL_INNER_ITERATIONS 0x109
#define NL_ELAPSED_TIME     0x10a
#define NL_PRECONDITIONER   0x10b
#define NL_GFLOPS           0x10c
#define NL_NNZ              0x10d
#define NL_NB_SYSTEMS       0x10e

/* ---- solvers ----------------------------------------------------------- */
#define NL_SOLVER_DEFAULT   0x000
#define NL_CG               0x200
#define NL_BICGSTAB         0x201
#define NL_GMRES            0x202

/* ---- direct-solver extensions (factorisation kinds) -------------------- */
#define NL_SUPERLU_EXT           0x210
#define NL_PERM_SUPERLU_EXT      0x211
#define NL_SYMMETRIC_SUPERLU_EXT 0x212
#define NL_CHOLMOD_EXT           0x213

/* ---- preconditioners --------------------------------------------------- */
#define NL_PRECOND_NONE     0x000
#define NL_PRECOND_JACOBI   0x300
#define NL_PRECOND_SSOR     0x301
#define NL_PRECOND_USER     0x303

/* ---- nlEnable / nlDisable ---------------------------------------------- */
#define NL_NORMALIZE_ROWS   0x400
#define NL_VERBOSE          0x401
#define NL_NO_VARIABLES_INDIRECTION 0x402

/* ---- nlBegin / nlEnd --------------------------------------------------- */
#define NL_SYSTEM  0x0
#define NL_MATRIX  0x1
#define NL_ROW     0x2
#define NL_MATRIX_PATTERN 0x3

/* ---- matrix types and storage ------------------------------------------ */
#define NL_MATRIX_SPARSE_DYNAMIC 0x1001
#define NL_MATRIX_CRS            0x1002
#define NL_MATRIX_SUPERLU_EXT    0x1003
#define NL_MATRIX_CHOLDMOD_EXT   0x1004
#define NL_MATRIX_FUNCTION       0x1005
#define NL_MATRIX_OTHER          0x1006

#define NL_MATRIX_STORE_ROWS      1
#define NL_MATRIX_STORE_COLUMNS   2
#define NL_MATRIX_STORE_SYMMETRIC 4

/* ---- memory types used by NLBlas Malloc/Memcpy ------------------------- */
typedef enum { NL_HOST_MEMORY = 0, NL_DEVICE_MEMORY = 1 } NLmemoryType;

/* ========================================================================
 * Matrices  (scene_graph/NL/matrix.cpp:121-143,163,303-310)
 * ====================================================================== */
struct NLMatrixStruct;
typedef struct NLMatrixStruct* NLMatrix;
typedef void (*NLDestroyMatrixFunc)(NLMatrix M);
typedef void (*NLMultMatrixVectorFunc)(NLMatrix M, const double* x, double* y);

/* Common header of every matrix.  The wrapper reads m, n and type directly
 * (scene_graph/NL/matrix.cpp:130,134,138-140). */
struct NLMatrixStruct {
    NLuint m;
    NLuint n;
    NLenum type;
    NLDestroyMatrixFunc destroy_func;
    NLMultMatrixVectorFunc mult_func;
};

/* Coefficient of a dynamic row. */
typedef struct {
    NLuint index;
    NLdouble value;
} NLCoeff;

/* One row (or column) of a dynamic matrix: duplicates accumulate. */
typedef struct {
    NLuint size;
    NLuint capacity;
    NLCoeff* coeff;
} NLRowColumn;

/* Dynamic sparse matrix (NL_MATRIX_SPARSE_DYNAMIC): row lists plus the
 * diagonal, accumulate-on-add (scene_graph/NL/matrix.h:105-123). */
typedef struct {
    NLuint m;
    NLuint n;
    NLenum type;
    NLDestroyMatrixFunc destroy_func;
    NLMultMatrixVectorFunc mult_func;
    NLuint diag_size;
    NLenum storage;
    NLRowColumn* row;
    NLdouble* diag;
    void* hip_mirror;        /* device copy, rebuilt after any add */
    NLuint version;          /* incremented by every add */
} NLSparseMatrix;

/* Compressed Row Storage (NL_MATRIX_CRS). rowptr has m+1 entries. */
typedef struct {
    NLuint m;
    NLuint n;
    NLenum type;
    NLDestroyMatrixFunc destroy_func;
    NLMultMatrixVectorFunc mult_func;
    NLuint* rowptr;
    NLuint* colind;
    NLdouble* val;
    NLuint nslices;
    NLuint* sliceptr;
    NLboolean symmetric_storage;
    void* hip_mirror;        /* resident device copy (created on first use) */
} NLCRSMatrix;

NLAPI NLMatrix NLAPIENTRY nlSparseMatrixNew(NLuint m, NLuint n, NLenum storage);
NLAPI void NLAPIENTRY nlSparseMatrixAdd(NLSparseMatrix* M, NLuint i, NLuint j,
                                        NLdouble value);
NLAPI void NLAPIENTRY nlSparseMatrixZero(NLSparseMatrix* M);
NLAPI NLuint NLAPIENTRY nlSparseMatrixNNZ(NLSparseMatrix* M);
/* CRS matrix from caller arrays (copied). */
NLAPI NLMatrix NLAPIENTRY nlCRSMatrixNewFromArrays(NLuint m, NLuint n,
                                                   const NLuint* rowptr,
                                                   const NLuint* colind,
                                                   const NLdouble* val);
/* Replaces *M (dynamic) by its CRS compression; other types are untouched. */
NLAPI void NLAPIENTRY nlMatrixCompress(NLMatrix* M);
/* y = M x.  For device matrices (nlHIP*), x and y are device pointers; for
 * host matrices they may be host or device pointers. */
NLAPI void NLAPIENTRY nlMultMatrixVector(NLMatrix M, const double* x, double* y);
NLAPI void NLAPIENTRY nlDeleteMatrix(NLMatrix M);
/* Direct factorisations belong to the SUPERLU / CHOLMOD extensions, which
 * this library does not provide: returns NULL with an error message. */
NLAPI NLMatrix NLAPIENTRY nlMatrixFactorize(NLMatrix M, NLenum solver);

/* ========================================================================
 * Preconditioners  (scene_graph/NL/matrix.cpp:363-368)
 * ====================================================================== */
/* Jacobi: y = D^-1 x, with 1 where a row has no or a zero diagonal. */
NLAPI NLMatrix NLAPIENTRY nlNewJacobiPreconditioner(NLMatrix M);
/* SSOR(omega) preconditioner of a square matrix (matrix.cpp:366-368 passes
 * omega = 1.5): y = omega (2 - omega) (D + omega U)^-1 D (D + omega L)^-1 x,
 * two triangular sweeps run on the device level by level, bitwise the
 * sequential sweeps.  The matrix is copied (later changes to M do not
 * reach P); NULL with an error when a diagonal entry is zero or omega is 0.
 * nlSolveSystemIterative accepts it for NL_CG, NL_BICGSTAB and NL_GMRES. */
NLAPI NLMatrix NLAPIENTRY nlNewSSORPreconditioner(NLMatrix M, NLdouble omega);

/* ========================================================================
 * NLBlas  (scene_graph/NL/blas.cpp:66,86,96,115,125,144)
 * ====================================================================== */
struct NLBlas;
typedef struct NLBlas* NLBlas_t;

typedef void* (*FUNPTR_malloc)(NLBlas_t, NLmemoryType, size_t);
typedef void (*FUNPTR_free)(NLBlas_t, NLmemoryType, size_t, void*);
typedef void (*FUNPTR_memcpy)(NLBlas_t, void*, NLmemoryType, const void*,
                              NLmemoryType, size_t);
typedef void (*FUNPTR_dcopy)(NLBlas_t, int, const double*, int, double*, int);
typedef void (*FUNPTR_dscal)(NLBlas_t, int, double, double*, int);
typedef double (*FUNPTR_ddot)(NLBlas_t, int, const double*, int, const double*, int);
typedef double (*FUNPTR_dnrm2)(NLBlas_t, int, const double*, int);
typedef void (*FUNPTR_daxpy)(NLBlas_t, int, double, const double*, int, double*, int);
/* y = diag(a) x, element-wise (used for Jacobi). */
typedef void (*FUNPTR_dmul)(NLBlas_t, int, const double*, const double*, double*);
typedef void (*FUNPTR_dzero)(NLBlas_t, int, double*);

struct NLBlas {
    FUNPTR_malloc Malloc;
    FUNPTR_free Free;
    FUNPTR_memcpy Memcpy;
    FUNPTR_dcopy Dcopy;
    FUNPTR_dscal Dscal;
    FUNPTR_ddot Ddot;
    FUNPTR_dnrm2 Dnrm2;
    FUNPTR_daxpy Daxpy;
    FUNPTR_dmul Dmul;
    FUNPTR_dzero Dzero;
    NLboolean has_unique_memory;
    NLulong start_time;
    NLulong flops;
    NLulong used_ram[2];
    NLulong max_used_ram[2];
};

/* Host-memory BLAS.  Vectors live in host memory; every operation is staged
 * to the GPU and computed there (there is no CPU BLAS in this library). */
NLAPI NLBlas_t NLAPIENTRY nlHostBlas(void);
NLAPI void NLAPIENTRY nlBlasResetStats(NLBlas_t blas);
NLAPI double NLAPIENTRY nlBlasGFlops(NLBlas_t blas);

/* ========================================================================
 * Iterative solver  (scene_graph/NL/matrix.cpp:390-393)
 * ====================================================================== */
/* Solves M x = b.  P may be NULL (no preconditioner) or a Jacobi
 * preconditioner.  Stops when |Mx-b| <= eps |b| or after max_iter iterations;
 * x holds the initial guess on entry.  b and x may be host or device
 * pointers.  Returns the number of iterations (0 on error). */
NLAPI NLuint NLAPIENTRY nlSolveSystemIterative(NLBlas_t blas, NLMatrix M,
                                               NLMatrix P, NLdouble* b,
                                               NLdouble* x, NLenum solver,
                                               double eps, NLuint max_iter,
                                               NLuint inner_iter);

/* ========================================================================
 * Extensions  (scene_graph/NL/matrix.cpp:51-66,430-434)
 * ====================================================================== */
/* "HIP" initialises the gfx950 device (NL_TRUE if usable).  Any other name
 * (SUPERLU, CHOLMOD, ARPACK, MKL, CUDA) returns NL_FALSE. */
NLAPI NLboolean NLAPIENTRY nlInitExtension(const char* extension);
NLAPI NLboolean NLAPIENTRY nlExtensionIsInitialized(const char* extension);

/* ========================================================================
 * Context API  (voxel_grob_transport_commands.cpp:231-287)
 * ====================================================================== */
NLAPI NLContext NLAPIENTRY nlNewContext(void);
NLAPI void NLAPIENTRY nlDeleteContext(NLContext context);
NLAPI void NLAPIENTRY nlMakeCurrent(NLContext context);
NLAPI NLContext NLAPIENTRY nlGetCurrent(void);

NLAPI void NLAPIENTRY nlSolverParameterd(NLenum pname, NLdouble param);
NLAPI void NLAPIENTRY nlSolverParameteri(NLenum pname, NLint param);
NLAPI void NLAPIENTRY nlGetBooleanv(NLenum pname, NLboolean* params);
NLAPI void NLAPIENTRY nlGetDoublev(NLenum pname, NLdouble* params);
NLAPI void NLAPIENTRY nlGetIntegerv(NLenum pname, NLint* params);
NLAPI void NLAPIENTRY nlEnable(NLenum pname);
NLAPI void NLAPIENTRY nlDisable(NLenum pname);
NLAPI NLboolean NLAPIENTRY nlIsEnabled(NLenum pname);

NLAPI void NLAPIENTRY nlSetVariable(NLuint i, NLdouble value);
NLAPI void NLAPIENTRY nlMultiSetVariable(NLuint i, NLuint k, NLdouble value);
NLAPI NLdouble NLAPIENTRY nlGetVariable(NLuint i);
NLAPI NLdouble NLAPIENTRY nlMultiGetVariable(NLuint i, NLuint k);
NLAPI void NLAPIENTRY nlLockVariable(NLuint index);
NLAPI void NLAPIENTRY nlUnlockVariable(NLuint index);
NLAPI NLboolean NLAPIENTRY nlVariableIsLocked(NLuint index);

NLAPI void NLAPIENTRY nlBegin(NLenum primitive);
NLAPI void NLAPIENTRY nlEnd(NLenum primitive);
NLAPI void NLAPIENTRY nlCoefficient(NLuint i, NLdouble value);
NLAPI void NLAPIENTRY nlAddIJCoefficient(NLuint i, NLuint j, NLdouble value);
NLAPI void NLAPIENTRY nlRightHandSide(NLdouble value);
NLAPI void NLAPIENTRY nlMultiRightHandSide(NLuint k, NLdouble value);
NLAPI void NLAPIENTRY nlRowScaling(NLdouble value);

/* Solves the current context's system on the GPU.  NL_TRUE on success. */
NLAPI NLboolean NLAPIENTRY nlSolve(void);

#ifdef __cplusplus
}
#endif

#endif /* OPENNL_HIP_NL_H */
