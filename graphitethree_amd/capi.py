"""ctypes bindings of libnlhip.so (include/nl.h and include/nl_hip.h).

Signatures are declared here once; ``lib()`` loads the in-tree library and
raises when it is missing -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

from . import _native

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"

NL_FALSE, NL_TRUE = 0, 1
NL_SOLVER, NL_NB_VARIABLES, NL_LEAST_SQUARES = 0x100, 0x101, 0x102
NL_MAX_ITERATIONS, NL_THRESHOLD, NL_OMEGA, NL_SYMMETRIC = 0x103, 0x104, 0x105, 0x106
NL_USED_ITERATIONS, NL_ERROR, NL_INNER_ITERATIONS = 0x107, 0x108, 0x109
NL_ELAPSED_TIME, NL_PRECONDITIONER, NL_GFLOPS, NL_NNZ, NL_NB_SYSTEMS = (
    0x10a, 0x10b, 0x10c, 0x10d, 0x10e)
NL_SOLVER_DEFAULT, NL_CG, NL_BICGSTAB, NL_GMRES = 0x000, 0x200, 0x201, 0x202
NL_SUPERLU_EXT, NL_PERM_SUPERLU_EXT, NL_SYMMETRIC_SUPERLU_EXT, NL_CHOLMOD_EXT = (
    0x210, 0x211, 0x212, 0x213)
NL_PRECOND_NONE, NL_PRECOND_JACOBI, NL_PRECOND_SSOR = 0x000, 0x300, 0x301
NL_NORMALIZE_ROWS, NL_VERBOSE = 0x400, 0x401
NL_SYSTEM, NL_MATRIX, NL_ROW = 0x0, 0x1, 0x2
NL_MATRIX_SPARSE_DYNAMIC, NL_MATRIX_CRS, NL_MATRIX_OTHER = 0x1001, 0x1002, 0x1006
NL_MATRIX_STORE_ROWS = 1

NLHIP_OK = 0
NLHIP_ERR_INVALID, NLHIP_ERR_NO_DEVICE, NLHIP_ERR_HIP, NLHIP_ERR_OOM = -1, -2, -3, -4
NLHIP_ERR_NONFINITE, NLHIP_ERR_UNSUPPORTED, NLHIP_ERR_BUDGET, NLHIP_ERR_INJECTED = (
    -5, -6, -7, -8)
NLHIP_PREC_FP64, NLHIP_PREC_MIXED, NLHIP_PREC_MIXED_FALLBACK = 0, 1, 2
NLHIP_SYSTEM_FP32 = 1
NLHIP_SYSTEM_NO_OFFSETS = 2
NLHIP_SYSTEM_NO_BLOCKS = 4


class NLMatrixStruct(C.Structure):
    _fields_ = [("m", C.c_uint), ("n", C.c_uint), ("type", C.c_uint),
                ("destroy_func", C.c_void_p), ("mult_func", C.c_void_p)]


NLMatrix = C.POINTER(NLMatrixStruct)

_DCOPY = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int)
_DSCAL = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int)
_DDOT = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                    C.c_int)
_DNRM2 = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int)
_DAXPY = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int,
                     C.c_void_p, C.c_int)
_DMUL = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)
_DZERO = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_void_p)


class NLBlasStruct(C.Structure):
    _fields_ = [("Malloc", C.c_void_p), ("Free", C.c_void_p), ("Memcpy", C.c_void_p),
                ("Dcopy", _DCOPY), ("Dscal", _DSCAL), ("Ddot", _DDOT), ("Dnrm2", _DNRM2),
                ("Daxpy", _DAXPY), ("Dmul", _DMUL), ("Dzero", _DZERO),
                ("has_unique_memory", C.c_ubyte), ("start_time", C.c_ulong),
                ("flops", C.c_ulong), ("used_ram", C.c_ulong * 2),
                ("max_used_ram", C.c_ulong * 2)]


NLBlas_t = C.POINTER(NLBlasStruct)


class SolveParams(C.Structure):
    _fields_ = [("precision", C.c_int32), ("precond", C.c_int32), ("eps", C.c_double),
                ("max_iter", C.c_uint32), ("check_every", C.c_uint32),
                ("inner_eps", C.c_double), ("max_outer", C.c_uint32),
                ("inner_max_iter", C.c_uint32), ("confirm", C.c_uint32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("outer_iterations", C.c_uint32),
                ("converged", C.c_int32), ("status", C.c_int32),
                ("rel_residual", C.c_double), ("cg_rel_residual", C.c_double),
                ("device_seconds", C.c_double), ("wall_seconds", C.c_double),
                ("extra_iterations", C.c_uint32), ("reserved", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class EigenStats(C.Structure):
    _fields_ = [("outer_iterations", C.c_uint32), ("converged", C.c_uint32),
                ("inner_iterations", C.c_uint64), ("max_rel_residual", C.c_double),
                ("seconds", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("k", C.c_uint32),
                ("n", C.c_uint32), ("device_bytes", C.c_uint64), ("stats", SolveStats),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class ServerInfo(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("in_use_bytes", C.c_uint64),
                ("peak_in_use_bytes", C.c_uint64), ("pending", C.c_uint32),
                ("running", C.c_uint32), ("peak_running", C.c_uint32),
                ("completed", C.c_uint64), ("failed", C.c_uint64)]


_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p

SIGNATURES = {
    # nl.h
    "nlSparseMatrixNew": (NLMatrix, [C.c_uint, C.c_uint, C.c_uint]),
    "nlSparseMatrixAdd": (None, [NLMatrix, C.c_uint, C.c_uint, C.c_double]),
    "nlSparseMatrixZero": (None, [NLMatrix]),
    "nlSparseMatrixNNZ": (C.c_uint, [NLMatrix]),
    "nlCRSMatrixNewFromArrays": (NLMatrix, [C.c_uint, C.c_uint, _vp, _vp, _vp]),
    "nlMatrixCompress": (None, [C.POINTER(NLMatrix)]),
    "nlMultMatrixVector": (None, [NLMatrix, _vp, _vp]),
    "nlDeleteMatrix": (None, [NLMatrix]),
    "nlMatrixFactorize": (NLMatrix, [NLMatrix, C.c_uint]),
    "nlNewJacobiPreconditioner": (NLMatrix, [NLMatrix]),
    "nlNewSSORPreconditioner": (NLMatrix, [NLMatrix, C.c_double]),
    "nlHostBlas": (NLBlas_t, []),
    "nlBlasResetStats": (None, [NLBlas_t]),
    "nlBlasGFlops": (C.c_double, [NLBlas_t]),
    "nlSolveSystemIterative": (C.c_uint, [NLBlas_t, NLMatrix, NLMatrix, _vp, _vp, C.c_uint,
                                          C.c_double, C.c_uint, C.c_uint]),
    "nlInitExtension": (C.c_ubyte, [C.c_char_p]),
    "nlExtensionIsInitialized": (C.c_ubyte, [C.c_char_p]),
    "nlNewContext": (_vp, []),
    "nlDeleteContext": (None, [_vp]),
    "nlMakeCurrent": (None, [_vp]),
    "nlGetCurrent": (_vp, []),
    "nlSolverParameterd": (None, [C.c_uint, C.c_double]),
    "nlSolverParameteri": (None, [C.c_uint, C.c_int]),
    "nlGetBooleanv": (None, [C.c_uint, C.POINTER(C.c_ubyte)]),
    "nlGetDoublev": (None, [C.c_uint, _f64p]),
    "nlGetIntegerv": (None, [C.c_uint, C.POINTER(C.c_int)]),
    "nlEnable": (None, [C.c_uint]),
    "nlDisable": (None, [C.c_uint]),
    "nlIsEnabled": (C.c_ubyte, [C.c_uint]),
    "nlSetVariable": (None, [C.c_uint, C.c_double]),
    "nlMultiSetVariable": (None, [C.c_uint, C.c_uint, C.c_double]),
    "nlGetVariable": (C.c_double, [C.c_uint]),
    "nlMultiGetVariable": (C.c_double, [C.c_uint, C.c_uint]),
    "nlLockVariable": (None, [C.c_uint]),
    "nlUnlockVariable": (None, [C.c_uint]),
    "nlVariableIsLocked": (C.c_ubyte, [C.c_uint]),
    "nlBegin": (None, [C.c_uint]),
    "nlEnd": (None, [C.c_uint]),
    "nlCoefficient": (None, [C.c_uint, C.c_double]),
    "nlAddIJCoefficient": (None, [C.c_uint, C.c_uint, C.c_double]),
    "nlRightHandSide": (None, [C.c_double]),
    "nlMultiRightHandSide": (None, [C.c_uint, C.c_double]),
    "nlRowScaling": (None, [C.c_double]),
    "nlSolve": (C.c_ubyte, []),
    # nl_hip.h
    "nlInitExtension_HIP": (C.c_ubyte, []),
    "nlExtensionIsInitialized_HIP": (C.c_ubyte, []),
    "nlHIPBlas": (NLBlas_t, []),
    "nlHIPMatrixNewFromCRSMatrix": (NLMatrix, [NLMatrix]),
    "nlHIPJacobiPreconditionerNewFromCRSMatrix": (NLMatrix, [NLMatrix]),
    "nlhipSolveParamsDefault": (None, [C.POINTER(SolveParams)]),
    "nlhipSystemCreate": (C.c_int, [C.c_uint32, _vp, _vp, _vp, C.c_uint32,
                                    C.POINTER(_vp)]),
    "nlhipSystemDestroy": (C.c_int, [_vp]),
    "nlhipSystemGetCRS": (C.c_int, [_vp, _vp, _vp, _vp]),
    "nlhipEigenSolve": (C.c_int, [_vp, _vp, C.c_uint32, C.c_double, C.c_double, C.c_uint32,
                                  _vp, _vp, C.POINTER(EigenStats)]),
    "nlhipLSCMCreate": (C.c_int, [C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, C.c_double,
                                  C.c_double, C.c_uint32, C.c_double, C.c_double, C.c_uint32,
                                  C.POINTER(_vp), _vp, _u32p]),
    "nlhipLeastSquaresCreate": (C.c_int, [C.c_uint32, C.c_uint32, _vp, _vp, _vp, C.c_uint32,
                                          _vp, _vp, _vp, C.c_uint32, C.POINTER(_vp), _vp,
                                          _u32p]),
    "nlhipSystemInfo": (C.c_int, [_vp, _u32p, C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]),
    "nlhipSystemSetLaunch": (C.c_int, [_vp, C.c_int, C.c_uint32]),
    "nlhipSystemGetLaunch": (C.c_int, [_vp, C.POINTER(C.c_int), _u32p]),
    "nlhipSystemAutotune": (C.c_int, [_vp, C.c_int, C.c_uint32, C.POINTER(C.c_int), _u32p,
                                      _f64p]),
    "nlhipBatchAutotune": (C.c_int, [_vp, _vp, C.c_uint32, _u32p, _u32p, _f64p]),
    "nlhipSolve": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveParams),
                             C.POINTER(SolveStats)]),
    "nlhipServerSubmitSystem": (C.c_int, [_vp, _vp, C.c_uint32, _vp, _vp,
                                          C.POINTER(SolveParams), C.POINTER(C.c_uint64)]),
    "nlhipSolveNoAlloc": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveParams),
                                    C.POINTER(SolveStats)]),
    "nlhipSolveBegin": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveParams)]),
    "nlhipSolveAdvance": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_int),
                                    C.POINTER(SolveStats)]),
    "nlhipSolveEnd": (C.c_int, [_vp, _vp, C.POINTER(SolveStats)]),
    "nlhipSolveMulti": (C.c_int, [_vp, C.c_uint32, _vp, _vp, C.POINTER(SolveParams),
                                  C.POINTER(SolveStats), C.POINTER(SolveStats)]),
    "nlhipSpMV": (C.c_int, [_vp, _vp, _vp]),
    "nlhipBenchSpMV": (C.c_int, [_vp, C.c_int, C.c_uint32, _f64p]),
    "nlhipBenchCopy": (C.c_int, [C.c_uint64, C.c_uint32, _f64p]),
    "nlhipBenchRead": (C.c_int, [C.c_uint64, C.c_uint32, _f64p]),
    "nlhipBenchStreamMix": (C.c_int, [C.c_uint64, C.c_uint32, C.c_int, C.c_uint32, _f64p]),
    "nlhipProfileCG": (C.c_int, [_vp, C.c_int, C.c_uint32, _f64p]),
    "nlhipBatchCreate": (C.c_int, [C.c_uint32, _vp, _vp, _vp, _vp, C.c_uint32,
                                   C.POINTER(_vp)]),
    "nlhipBatchDestroy": (C.c_int, [_vp]),
    "nlhipBatchGetCRS": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "nlhipBatchLSCMCreate": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint32, _vp, _vp, _vp,
                                       C.c_uint32, C.POINTER(_vp), _vp, _u32p]),
    "nlhipBatchInfo": (C.c_int, [_vp, _u32p, _u32p, C.POINTER(C.c_uint64), _u32p]),
    "nlhipBatchSetLaunch": (C.c_int, [_vp, C.c_uint32]),
    "nlhipBatchSetGraphs": (C.c_int, [_vp, C.c_int]),
    "nlhipBatchSetPersistent": (C.c_int, [_vp, C.c_int]),
    "nlhipBatchFallbacks": (C.c_int, [_vp, _u32p]),
    "nlhipBatchSetSpmvPasses": (C.c_int, [_vp, C.c_uint32]),
    "nlhipBatchSolve": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveParams),
                                  C.POINTER(SolveStats), C.POINTER(SolveStats)]),
    "nlhipServerConfigDefault": (None, [C.POINTER(ServerConfig)]),
    "nlhipServerCreate": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "nlhipServerSubmit": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32, _vp, _vp,
                                    C.POINTER(SolveParams), C.POINTER(C.c_uint64)]),
    "nlhipServerWait": (C.c_int, [_vp, C.c_uint64, _vp, C.POINTER(RequestResult),
                                  C.c_double]),
    "nlhipServerInfo": (C.c_int, [_vp, C.POINTER(ServerInfo)]),
    "nlhipServerDestroy": (C.c_int, [_vp]),
    "nlhipLastError": (C.c_char_p, []),
    "nlhipInjectFailures": (None, [C.c_int]),
    "nlhipDeviceInfo": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                  C.POINTER(C.c_uint64)]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = _native.load("libnlhip.so")
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Every function declared in include/*.h."""
    names = []
    for h in sorted(INCLUDE.glob("*.h")):
        names += re.findall(r"NLAPIENTRY\s+([A-Za-z_0-9]+)\s*\(", h.read_text())
    return names


def last_error() -> str:
    return (lib().nlhipLastError() or b"").decode()


class NLHIPError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        super().__init__(f"{where} failed ({code}): {last_error()}")


def check(rc: int, where: str) -> None:
    if rc != NLHIP_OK:
        raise NLHIPError(rc, where)


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def default_params(**kw) -> SolveParams:
    p = SolveParams()
    lib().nlhipSolveParamsDefault(C.byref(p))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = {"fp64": NLHIP_PREC_FP64, "mixed": NLHIP_PREC_MIXED,
                 "mixed-fallback": NLHIP_PREC_MIXED_FALLBACK}[v]
        if k == "precond" and isinstance(v, str):
            v = {"jacobi": NL_PRECOND_JACOBI, "none": NL_PRECOND_NONE}[v.lower()]
        setattr(p, k, v)
    return p


class System:
    """A CRS system resident on the GPU (nlhipSystemCreate)."""

    def __init__(self, rowptr, colind, val, fp32: bool = False, offsets: bool = True,
                 blocks: bool = True):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.uint32)
        self.colind = np.ascontiguousarray(colind, dtype=np.uint32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        flags = ((NLHIP_SYSTEM_FP32 if fp32 else 0) | (0 if offsets else NLHIP_SYSTEM_NO_OFFSETS)
                 | (0 if blocks else NLHIP_SYSTEM_NO_BLOCKS))
        h = C.c_void_p()
        check(lib().nlhipSystemCreate(self.n, ptr(self.rowptr), ptr(self.colind),
                                      ptr(self.val), flags, C.byref(h)), "nlhipSystemCreate")
        self._h = h

    @classmethod
    def from_crs(cls, S, fp32: bool = False, offsets: bool = True,
                 blocks: bool = True) -> "System":
        return cls(S.rowptr, S.colind, S.val, fp32=fp32, offsets=offsets, blocks=blocks)

    @classmethod
    def least_squares(cls, nb_variables: int, row_ptr, var, coef, row_rhs=None, locked=None,
                      values=None, fp32: bool = False):
        """Normal equations of the rows assembled on the device
        (nlhipLeastSquaresCreate).  row_rhs: (k, rows) or (rows,); values:
        (k, nb_variables) or (nb_variables,).  Returns (system, rhs (k, nfree))."""
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        var = np.ascontiguousarray(var, dtype=np.uint32)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        nrows = len(row_ptr) - 1
        k = 1
        if row_rhs is not None:
            row_rhs = np.ascontiguousarray(np.atleast_2d(row_rhs), dtype=np.float64)
            k = row_rhs.shape[0]
        if values is not None:
            values = np.ascontiguousarray(np.atleast_2d(values), dtype=np.float64)
            k = max(k, values.shape[0])
        if locked is not None:
            locked = np.ascontiguousarray(locked, dtype=np.uint8)
        nfree_locked = nb_variables - (int(np.count_nonzero(locked)) if locked is not None else 0)
        rhs = np.zeros((k, max(nfree_locked, 1)))
        h = C.c_void_p()
        nfree = C.c_uint32()
        flags = NLHIP_SYSTEM_FP32 if fp32 else 0
        check(lib().nlhipLeastSquaresCreate(
            nb_variables, nrows, ptr(row_ptr), ptr(var), ptr(coef), k,
            ptr(row_rhs) if row_rhs is not None else None,
            ptr(locked) if locked is not None else None,
            ptr(values) if values is not None else None, flags, C.byref(h), ptr(rhs),
            C.byref(nfree)), "nlhipLeastSquaresCreate")
        S = cls.__new__(cls)
        S._h = h
        S.n = int(nfree.value)
        return S, rhs[:, :S.n]

    @classmethod
    def lscm(cls, points, triangles, pins, pin_uv, fp32: bool = False):
        """LSCM system of a mesh built on the device (nlhipLSCMCreate).
        Returns (system, rhs)."""
        P = np.ascontiguousarray(points, dtype=np.float64)
        T = np.ascontiguousarray(triangles, dtype=np.uint32)
        nv, nt = len(P), len(T)
        rhs = np.zeros(max(2 * nv - 4, 1))
        h = C.c_void_p()
        nfree = C.c_uint32()
        check(lib().nlhipLSCMCreate(nv, ptr(P), nt, ptr(T), int(pins[0]), float(pin_uv[0]),
                                    float(pin_uv[1]), int(pins[1]), float(pin_uv[2]),
                                    float(pin_uv[3]), NLHIP_SYSTEM_FP32 if fp32 else 0,
                                    C.byref(h), ptr(rhs), C.byref(nfree)), "nlhipLSCMCreate")
        S = cls.__new__(cls)
        S._h = h
        S.n = int(nfree.value)
        return S, rhs[:S.n]

    def eigen(self, k: int, mass=None, shift: float = 0.0, eps: float = 1e-8,
              max_iter: int = 1000):
        """k eigenpairs of A x = lambda M x nearest above `shift`
        (nlhipEigenSolve).  Returns (values (k,), vectors (k, n), stats)."""
        vals = np.zeros(k)
        vecs = np.zeros((k, self.n))
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        st = EigenStats()
        check(lib().nlhipEigenSolve(self._h, ptr(m) if m is not None else None, k, float(shift),
                                    float(eps), int(max_iter), ptr(vals), ptr(vecs),
                                    C.byref(st)), "nlhipEigenSolve")
        return vals, vecs, st

    def crs(self):
        """The resident CRS (rowptr, colind, val) copied back."""
        n, nnz, _ = self.info()
        rp = np.empty(n + 1, dtype=np.uint32)
        ci = np.empty(max(nnz, 1), dtype=np.uint32)
        va = np.empty(max(nnz, 1), dtype=np.float64)
        check(lib().nlhipSystemGetCRS(self._h, ptr(rp), ptr(ci), ptr(va)), "nlhipSystemGetCRS")
        return rp, ci[:nnz], va[:nnz]

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().nlhipSystemDestroy(self._h)
            self._h = None

    __del__ = close

    def launch(self):
        """(rows per wavefront, workgroups) of the SpMV engine."""
        rpw, wg = C.c_int(), C.c_uint32()
        check(lib().nlhipSystemGetLaunch(self._h, C.byref(rpw), C.byref(wg)), "get launch")
        return int(rpw.value), int(wg.value)

    def set_launch(self, rows_per_wave: int, workgroups: int):
        check(lib().nlhipSystemSetLaunch(self._h, rows_per_wave, workgroups), "set launch")

    def autotune(self, fp32: bool = False, iters: int = 20):
        """Times the launch shapes (nlhipSystemAutotune); keeps and returns the
        fastest (rows per wavefront, workgroups, seconds per CG iteration)."""
        rpw, wg, t = C.c_int(), C.c_uint32(), C.c_double()
        check(lib().nlhipSystemAutotune(self._h, 1 if fp32 else 0, iters, C.byref(rpw),
                                        C.byref(wg), C.byref(t)), "nlhipSystemAutotune")
        return int(rpw.value), int(wg.value), float(t.value)

    def info(self):
        n, nnz, b = C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(lib().nlhipSystemInfo(self._h, C.byref(n), C.byref(nnz), C.byref(b)), "info")
        return int(n.value), int(nnz.value), int(b.value)

    def solve(self, b, x0=None, **params):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        p = default_params(**params)
        st = SolveStats()
        check(lib().nlhipSolve(self._h, ptr(b), ptr(x), C.byref(p), C.byref(st)), "nlhipSolve")
        return x, st

    def solve_begin(self, b, x0=None, **params):
        """Opens a sliced FP64 solve (nlhipSolveBegin); advance it with
        solve_advance and close it with solve_end."""
        self._sliced_b = np.ascontiguousarray(b, dtype=np.float64)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        p = default_params(**params)
        check(lib().nlhipSolveBegin(self._h, ptr(self._sliced_b),
                                    None if x0 is None else ptr(x0), C.byref(p)),
              "nlhipSolveBegin")

    def solve_advance(self, iters: int):
        """At most `iters` more iterations; returns (finished, stats so far)."""
        fin = C.c_int()
        st = SolveStats()
        check(lib().nlhipSolveAdvance(self._h, int(iters), C.byref(fin), C.byref(st)),
              "nlhipSolveAdvance")
        return bool(fin.value), st

    def solve_end(self):
        """Closes the sliced solve; returns (x, stats)."""
        x = np.empty(self.n)
        st = SolveStats()
        check(lib().nlhipSolveEnd(self._h, ptr(x), C.byref(st)), "nlhipSolveEnd")
        return x, st

    def solve_multi(self, B, X0=None, **params):
        """B: (k, n) right-hand sides (row j = system j).  Returns (X, per, total)."""
        B = np.ascontiguousarray(B, dtype=np.float64)
        k = B.shape[0]
        X = np.zeros_like(B) if X0 is None else np.array(X0, dtype=np.float64, copy=True)
        p = default_params(**params)
        per = (SolveStats * k)()
        tot = SolveStats()
        check(lib().nlhipSolveMulti(self._h, k, ptr(B), ptr(X), C.byref(p), per, C.byref(tot)),
              "nlhipSolveMulti")
        return X, list(per), tot

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        check(lib().nlhipSpMV(self._h, ptr(x), ptr(y)), "nlhipSpMV")
        return y

    def bench_spmv(self, fp32: bool = False, reps: int = 20) -> float:
        s = C.c_double()
        check(lib().nlhipBenchSpMV(self._h, 1 if fp32 else 0, reps, C.byref(s)),
              "nlhipBenchSpMV")
        return float(s.value)


class Batch:
    """Many independent systems packed block-diagonally (nlhipBatchCreate)."""

    def __init__(self, systems=None, fp32: bool = False, packed=None):
        """`systems`: sequence of objects with rowptr, colind, val (local CRS);
        or `packed` = (row_off, rowptr, colind, val) already block-diagonal."""
        if packed is not None:
            self._create(*packed, fp32=fp32)
            return
        row_off = [0]
        rps, cis, vas = [], [], []
        nnz_off = 0
        for S in systems:
            n = len(S.rowptr) - 1
            rps.append(np.asarray(S.rowptr[:-1], dtype=np.int64) + nnz_off)
            cis.append(np.asarray(S.colind, dtype=np.int64) + row_off[-1])
            vas.append(np.asarray(S.val, dtype=np.float64))
            nnz_off += int(S.rowptr[-1])
            row_off.append(row_off[-1] + n)
        self._create(np.array(row_off, dtype=np.uint32),
                     np.concatenate(rps + [np.array([nnz_off])]).astype(np.uint32),
                     np.concatenate(cis).astype(np.uint32), np.concatenate(vas), fp32=fp32)

    def _create(self, row_off, rowptr, colind, val, fp32: bool):
        self.row_off = np.ascontiguousarray(row_off, dtype=np.uint32)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.uint32)
        self.colind = np.ascontiguousarray(colind, dtype=np.uint32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.nsys = len(self.row_off) - 1
        self.n = int(self.row_off[-1])
        h = C.c_void_p()
        check(lib().nlhipBatchCreate(self.nsys, ptr(self.row_off), ptr(self.rowptr),
                                     ptr(self.colind), ptr(self.val),
                                     NLHIP_SYSTEM_FP32 if fp32 else 0, C.byref(h)),
              "nlhipBatchCreate")
        self._h = h

    @classmethod
    def lscm_atlas(cls, vertex_off, points, triangles, pins, pin_uv, fp32: bool = False):
        """An atlas's LSCM systems built on the device from its chart meshes
        (nlhipBatchLSCMCreate).  Returns (batch, rhs)."""
        vo = np.ascontiguousarray(vertex_off, dtype=np.uint32)
        P = np.ascontiguousarray(points, dtype=np.float64)
        T = np.ascontiguousarray(triangles, dtype=np.uint32)
        pn = np.ascontiguousarray(pins, dtype=np.uint32)
        uv = np.ascontiguousarray(pin_uv, dtype=np.float64)
        nc = len(vo) - 1
        rhs = np.zeros(max(2 * len(P), 1))
        h = C.c_void_p()
        nfree = C.c_uint32()
        check(lib().nlhipBatchLSCMCreate(nc, ptr(vo), ptr(P), len(T), ptr(T), ptr(pn), ptr(uv),
                                         NLHIP_SYSTEM_FP32 if fp32 else 0, C.byref(h), ptr(rhs),
                                         C.byref(nfree)), "nlhipBatchLSCMCreate")
        B = cls.__new__(cls)
        B._h = h
        B.nsys, B.n = nc, int(nfree.value)
        B.row_off = np.empty(nc + 1, dtype=np.uint32)
        check(lib().nlhipBatchGetCRS(h, ptr(B.row_off), None, None, None), "nlhipBatchGetCRS")
        return B, rhs[:B.n]

    def crs(self):
        """(row_off, rowptr, colind, val) the batch holds (nlhipBatchGetCRS)."""
        nsys, n, nnz, _ = self.info()
        ro = np.empty(nsys + 1, dtype=np.uint32)
        rp = np.empty(n + 1, dtype=np.uint32)
        ci = np.empty(max(nnz, 1), dtype=np.uint32)
        va = np.empty(max(nnz, 1), dtype=np.float64)
        check(lib().nlhipBatchGetCRS(self._h, ptr(ro), ptr(rp), ptr(ci), ptr(va)),
              "nlhipBatchGetCRS")
        return ro, rp, ci[:nnz], va[:nnz]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().nlhipBatchDestroy(self._h)
            self._h = None

    __del__ = close

    def info(self):
        ns, n, nnz, t = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_uint32()
        check(lib().nlhipBatchInfo(self._h, C.byref(ns), C.byref(n), C.byref(nnz),
                                   C.byref(t)), "nlhipBatchInfo")
        return int(ns.value), int(n.value), int(nnz.value), int(t.value)

    def set_launch(self, workgroups: int):
        check(lib().nlhipBatchSetLaunch(self._h, workgroups), "nlhipBatchSetLaunch")

    def fallbacks(self) -> int:
        """Systems the last "mixed-fallback" solve finished in FP64."""
        n = C.c_uint32()
        check(lib().nlhipBatchFallbacks(self._h, C.byref(n)), "nlhipBatchFallbacks")
        return int(n.value)

    def set_persistent(self, on: bool):
        check(lib().nlhipBatchSetPersistent(self._h, 1 if on else 0), "nlhipBatchSetPersistent")

    def set_graphs(self, on: bool):
        """False: enqueue every launch directly (nlhipBatchSetGraphs)."""
        check(lib().nlhipBatchSetGraphs(self._h, 1 if on else 0), "nlhipBatchSetGraphs")

    def autotune(self, b, iters: int = 50):
        """Times grid and SpMV run length (nlhipBatchAutotune); keeps and returns
        the fastest (workgroups, passes, seconds per iteration)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        wg, ps, t = C.c_uint32(), C.c_uint32(), C.c_double()
        check(lib().nlhipBatchAutotune(self._h, ptr(b), iters, C.byref(wg), C.byref(ps),
                                       C.byref(t)), "nlhipBatchAutotune")
        return int(wg.value), int(ps.value), float(t.value)

    def set_spmv_passes(self, passes: int):
        """SpMV passes per wavefront: 1, 2, 4, or 0 (by active tiles)."""
        check(lib().nlhipBatchSetSpmvPasses(self._h, passes), "nlhipBatchSetSpmvPasses")

    def split(self, v):
        return [v[self.row_off[s]:self.row_off[s + 1]] for s in range(self.nsys)]

    def solve(self, b, x0=None, **params):
        """b: concatenated right-hand sides.  Returns (x, per-system stats, total)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        p = default_params(**params)
        per = (SolveStats * self.nsys)()
        tot = SolveStats()
        check(lib().nlhipBatchSolve(self._h, ptr(b), ptr(x), C.byref(p), per, C.byref(tot)),
              "nlhipBatchSolve")
        return x, list(per), tot


class Server:
    """Solve server (nlhipServerCreate): budgeted admission, concurrent runs."""

    def __init__(self, hbm_budget_bytes: int | None = None, max_active: int = 8):
        c = ServerConfig()
        lib().nlhipServerConfigDefault(C.byref(c))
        if hbm_budget_bytes is not None:
            c.hbm_budget_bytes = hbm_budget_bytes
        c.max_active = max_active
        h = C.c_void_p()
        check(lib().nlhipServerCreate(C.byref(c), C.byref(h)), "nlhipServerCreate")
        self._h = h
        self._n = {}

    def submit(self, S, B, X0=None, **params) -> int:
        """S: CRS system (rowptr, colind, val); B: (k, n) or (n,)."""
        rp = np.ascontiguousarray(S.rowptr, dtype=np.uint32)
        ci = np.ascontiguousarray(S.colind, dtype=np.uint32)
        va = np.ascontiguousarray(S.val, dtype=np.float64)
        B = np.ascontiguousarray(np.atleast_2d(B), dtype=np.float64)
        k, n = B.shape
        x0 = None if X0 is None else np.ascontiguousarray(np.atleast_2d(X0), dtype=np.float64)
        p = default_params(**params)
        rid = C.c_uint64()
        check(lib().nlhipServerSubmit(self._h, n, ptr(rp), ptr(ci), ptr(va), k, ptr(B),
                                      None if x0 is None else ptr(x0), C.byref(p),
                                      C.byref(rid)), "nlhipServerSubmit")
        self._n[int(rid.value)] = (k, n, None)
        return int(rid.value)

    def submit_system(self, system: "System", B, X0=None, **params) -> int:
        """A request on a resident System (nlhipServerSubmitSystem): only the
        right-hand sides B ((k, n) or (n,)) travel.  The System is kept alive
        until the request has been waited for."""
        B = np.ascontiguousarray(np.atleast_2d(B), dtype=np.float64)
        k, n = B.shape
        if n != system.n:
            raise ValueError("right-hand sides do not match the system")
        x0 = None if X0 is None else np.ascontiguousarray(np.atleast_2d(X0), dtype=np.float64)
        p = default_params(**params)
        rid = C.c_uint64()
        check(lib().nlhipServerSubmitSystem(self._h, system.handle, k, ptr(B),
                                            None if x0 is None else ptr(x0), C.byref(p),
                                            C.byref(rid)), "nlhipServerSubmitSystem")
        self._n[int(rid.value)] = (k, n, system)
        return int(rid.value)

    def wait(self, rid: int, timeout: float = 0.0):
        """Returns (X (k, n), RequestResult); raises NLHIPError on failure."""
        k, n = self._n.pop(rid)[:2]
        X = np.empty((k, n))
        res = RequestResult()
        rc = lib().nlhipServerWait(self._h, rid, ptr(X), C.byref(res), timeout)
        if rc != NLHIP_OK:
            err = NLHIPError(rc, "nlhipServerWait")
            err.result = res
            raise err
        return X, res

    def info(self) -> ServerInfo:
        i = ServerInfo()
        check(lib().nlhipServerInfo(self._h, C.byref(i)), "nlhipServerInfo")
        return i

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().nlhipServerDestroy(self._h)
            self._h = None

    __del__ = close


def profile_cg(system: "System", iters: int = 50, fp32: bool = False):
    """Mean (direction, spmv, update) kernel seconds over `iters` iterations."""
    t = (C.c_double * 3)()
    check(lib().nlhipProfileCG(system.handle, 1 if fp32 else 0, iters, t), "nlhipProfileCG")
    return float(t[0]), float(t[1]), float(t[2])


def bench_copy(nbytes: int, reps: int = 20) -> float:
    """Mean seconds of a plain device copy of `nbytes` (HIP events)."""
    t = C.c_double()
    check(lib().nlhipBenchCopy(nbytes, reps, C.byref(t)), "nlhipBenchCopy")
    return float(t.value)


def bench_read(nbytes: int, reps: int = 20) -> float:
    """Mean seconds of a read-only pass over `nbytes` of device memory."""
    t = C.c_double()
    check(lib().nlhipBenchRead(nbytes, reps, C.byref(t)), "nlhipBenchRead")
    return float(t.value)


def bench_stream_mix(blocks: int, blocks_per_row: int, writes: bool = True,
                     reps: int = 20) -> float:
    """Mean seconds of one pass of the block SpMV's traffic without its
    arithmetic (nlhipBenchStreamMix)."""
    t = C.c_double()
    check(lib().nlhipBenchStreamMix(blocks, blocks_per_row, 1 if writes else 0, reps,
                                    C.byref(t)), "nlhipBenchStreamMix")
    return float(t.value)


def device_info():
    name = C.create_string_buffer(256)
    cu, hbm = C.c_int(), C.c_uint64()
    check(lib().nlhipDeviceInfo(name, 256, C.byref(cu), C.byref(hbm)), "nlhipDeviceInfo")
    return name.value.decode(), int(cu.value), int(hbm.value)
