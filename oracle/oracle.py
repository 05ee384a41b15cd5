"""ORACLE -- test infrastructure, not product code.

ctypes bindings for ``liboracle.so`` (``nl_oracle.c``), the CPU restatement of
the OpenNL solve path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only to check
or time against; the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NL_ORACLE_LIB points at another build of nl_oracle.c (the sanitizer build
# of oracle/Makefile's `sanitized` target).
_LIB_PATH = os.environ.get("NL_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-B", "-C", _HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built (run oracle/Makefile)")
        L = C.CDLL(_LIB_PATH)
        sig = {
            "or_spmv": (None, [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p]),
            "or_spmv_omp": (None, [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p]),
            "or_dcopy": (None, [C.c_uint32, _f64p, _f64p]),
            "or_dscal": (None, [C.c_uint32, C.c_double, _f64p]),
            "or_ddot": (C.c_double, [C.c_uint32, _f64p, _f64p]),
            "or_ddot_omp": (C.c_double, [C.c_uint32, _f64p, _f64p]),
            "or_dnrm2": (C.c_double, [C.c_uint32, _f64p]),
            "or_daxpy": (None, [C.c_uint32, C.c_double, _f64p, _f64p]),
            "or_jacobi": (None, [C.c_uint32, _u32p, _u32p, _f64p, _f64p]),
            "or_pcg": (C.c_int, [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p,
                                 _f64p, C.c_double, C.c_int, C.c_int, _f64p]),
            "or_pcg_timed_iters": (C.c_double, [C.c_uint32, _u32p, _u32p, _f64p,
                                                _f64p, _f64p, _f64p, C.c_int,
                                                C.c_int]),
            "or_bicgstab": (C.c_int, [C.c_uint32, _u32p, _u32p, _f64p, _f64p,
                                      _f64p, _f64p, C.c_double, C.c_int, _f64p]),
            "or_gmres": (C.c_int, [C.c_uint32, _u32p, _u32p, _f64p, _f64p,
                                   _f64p, _f64p, C.c_double, C.c_int, C.c_int, _f64p]),
            "or_ssor": (None, [C.c_uint32, _u32p, _u32p, _f64p, C.c_double, _f64p, _f64p]),
            "or_solve_ssor": (C.c_int, [C.c_int, C.c_uint32, _u32p, _u32p, _f64p,
                                        C.c_double, _f64p, _f64p, C.c_double, C.c_int,
                                        C.c_int, _f64p]),
            "or_dyn_new": (C.c_void_p, [C.c_uint32, C.c_uint32]),
            "or_dyn_free": (None, [C.c_void_p]),
            "or_dyn_add": (None, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double]),
            "or_dyn_nnz": (C.c_uint64, [C.c_void_p]),
            "or_dyn_compress": (None, [C.c_void_p, _u32p, _u32p, _f64p]),
            "or_ctx_new": (C.c_void_p, [C.c_uint32, C.c_int]),
            "or_ctx_free": (None, [C.c_void_p]),
            "or_ctx_set_variable": (None, [C.c_void_p, C.c_uint32, C.c_double]),
            "or_ctx_get_variable": (C.c_double, [C.c_void_p, C.c_uint32]),
            "or_ctx_lock_variable": (None, [C.c_void_p, C.c_uint32]),
            "or_ctx_begin_matrix": (None, [C.c_void_p]),
            "or_ctx_begin_row": (None, [C.c_void_p]),
            "or_ctx_coefficient": (None, [C.c_void_p, C.c_uint32, C.c_double]),
            "or_ctx_right_hand_side": (None, [C.c_void_p, C.c_double]),
            "or_ctx_end_row": (None, [C.c_void_p]),
            "or_ctx_rows": (None, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
            "or_ctx_nb_free": (C.c_uint32, [C.c_void_p]),
            "or_ctx_nnz": (C.c_uint64, [C.c_void_p]),
            "or_ctx_system": (None, [C.c_void_p, _u32p, _u32p, _f64p, _f64p]),
            "or_ctx_solve": (C.c_int, [C.c_void_p, C.c_double, C.c_int, _f64p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class CRS:
    """A CRS matrix as three numpy arrays (rowptr u32, colind u32, val f64)."""

    def __init__(self, rowptr, colind, val, ncols=None):
        self.rowptr = _u32(rowptr)
        self.colind = _u32(colind)
        self.val = _f64(val)
        self.m = len(self.rowptr) - 1
        self.n = self.m if ncols is None else int(ncols)

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def args(self):
        return (self.m, _p(self.rowptr, C.c_uint32), _p(self.colind, C.c_uint32),
                _p(self.val, C.c_double))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_array((self.val, self.colind.astype(np.int64),
                             self.rowptr.astype(np.int64)), shape=(self.m, self.n))


def spmv(A: CRS, x, parallel: bool = False) -> np.ndarray:
    x = _f64(x)
    y = np.empty(A.m, dtype=np.float64)
    f = lib().or_spmv_omp if parallel else lib().or_spmv
    f(*A.args(), _p(x, C.c_double), _p(y, C.c_double))
    return y


def ddot(x, y, parallel: bool = False) -> float:
    x, y = _f64(x), _f64(y)
    f = lib().or_ddot_omp if parallel else lib().or_ddot
    return float(f(len(x), _p(x, C.c_double), _p(y, C.c_double)))


def dnrm2(x) -> float:
    x = _f64(x)
    return float(lib().or_dnrm2(len(x), _p(x, C.c_double)))


def daxpy(a: float, x, y: np.ndarray) -> None:
    x = _f64(x)
    assert y.dtype == np.float64 and y.flags.c_contiguous
    lib().or_daxpy(len(x), float(a), _p(x, C.c_double), _p(y, C.c_double))


def dscal(a: float, x: np.ndarray) -> None:
    assert x.dtype == np.float64 and x.flags.c_contiguous
    lib().or_dscal(len(x), float(a), _p(x, C.c_double))


def dcopy(x, y: np.ndarray) -> None:
    x = _f64(x)
    lib().or_dcopy(len(x), _p(x, C.c_double), _p(y, C.c_double))


def jacobi(A: CRS) -> np.ndarray:
    d = np.empty(A.m, dtype=np.float64)
    lib().or_jacobi(*A.args(), _p(d, C.c_double))
    return d


def pcg(A: CRS, b, x0=None, eps=1e-10, max_iter=5000, precond="jacobi",
        variant=0):
    """Jacobi-PCG (or plain CG with precond=None).  Returns (x, iters, rel_res)."""
    b = _f64(b)
    x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
    dinv = jacobi(A) if precond == "jacobi" else None
    rel2 = C.c_double(0.0)
    it = lib().or_pcg(*A.args(), None if dinv is None else _p(dinv, C.c_double),
                      _p(b, C.c_double), _p(x, C.c_double), float(eps),
                      int(max_iter), int(variant), C.byref(rel2))
    return x, int(it), float(np.sqrt(rel2.value))


def pcg_timed_iters(A: CRS, b, iters: int, variant: int = 1) -> float:
    """Seconds for exactly `iters` Jacobi-PCG iterations (no stopping test)."""
    b = _f64(b)
    x = np.zeros(A.m)
    dinv = jacobi(A)
    return float(lib().or_pcg_timed_iters(*A.args(), _p(dinv, C.c_double),
                                          _p(b, C.c_double), _p(x, C.c_double),
                                          int(iters), int(variant)))


def bicgstab(A: CRS, b, x0=None, eps=1e-10, max_iter=5000, precond=None):
    b = _f64(b)
    x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
    dinv = jacobi(A) if precond == "jacobi" else None
    rel2 = C.c_double(0.0)
    it = lib().or_bicgstab(*A.args(), None if dinv is None else _p(dinv, C.c_double),
                           _p(b, C.c_double), _p(x, C.c_double), float(eps),
                           int(max_iter), C.byref(rel2))
    return x, int(it), float(np.sqrt(rel2.value))


def gmres(A: CRS, b, x0=None, eps=1e-10, max_iter=5000, m=5, precond=None):
    """Restarted GMRES(m).  Returns (x, Arnoldi steps, relative residual)."""
    b = _f64(b)
    x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
    dinv = jacobi(A) if precond == "jacobi" else None
    rel2 = C.c_double(0.0)
    it = lib().or_gmres(*A.args(), None if dinv is None else _p(dinv, C.c_double),
                        _p(b, C.c_double), _p(x, C.c_double), float(eps),
                        int(max_iter), int(m), C.byref(rel2))
    return x, int(it), float(np.sqrt(rel2.value))


def ssor(A: CRS, x, omega: float = 1.5) -> np.ndarray:
    """y = SSOR(omega)^-1 x: forward sweep, diagonal scaling, backward sweep,
    times (2 - omega)."""
    x = _f64(x)
    y = np.zeros(A.m)
    lib().or_ssor(*A.args(), float(omega), _p(x, C.c_double), _p(y, C.c_double))
    return y


SOLVERS = {"cg": 0, "bicgstab": 1, "gmres": 2}


def solve_ssor(A: CRS, b, solver: str = "cg", x0=None, eps=1e-10, max_iter=5000, m=5,
               omega: float = 1.5):
    """CG / BiCGSTAB / GMRES(m) preconditioned by SSOR(omega).
    Returns (x, iterations, relative residual)."""
    b = _f64(b)
    x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
    rel2 = C.c_double(0.0)
    it = lib().or_solve_ssor(SOLVERS[solver], *A.args(), float(omega), _p(b, C.c_double),
                             _p(x, C.c_double), float(eps), int(max_iter), int(m),
                             C.byref(rel2))
    return x, int(it), float(np.sqrt(rel2.value))


class Dyn:
    """Dynamic (row-accumulating) sparse matrix."""

    def __init__(self, m: int, n: int):
        self.m, self.n = m, n
        self._h = lib().or_dyn_new(m, n)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_dyn_free(self._h)
            self._h = None

    def add(self, i: int, j: int, a: float) -> None:
        lib().or_dyn_add(self._h, i, j, float(a))

    def compress(self) -> CRS:
        nnz = int(lib().or_dyn_nnz(self._h))
        rp = np.empty(self.m + 1, dtype=np.uint32)
        ci = np.empty(max(nnz, 1), dtype=np.uint32)
        va = np.empty(max(nnz, 1), dtype=np.float64)
        lib().or_dyn_compress(self._h, _p(rp, C.c_uint32), _p(ci, C.c_uint32),
                              _p(va, C.c_double))
        return CRS(rp, ci[:nnz], va[:nnz], self.n)


class Context:
    """Row-system / least-squares context (nlBegin/nlEnd(NL_ROW) semantics)."""

    def __init__(self, nb_variables: int, least_squares: bool = True):
        self.nv = nb_variables
        self._h = lib().or_ctx_new(nb_variables, 1 if least_squares else 0)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_ctx_free(self._h)
            self._h = None

    def set_variable(self, i, v):
        lib().or_ctx_set_variable(self._h, i, float(v))

    def get_variable(self, i) -> float:
        return float(lib().or_ctx_get_variable(self._h, i))

    def lock_variable(self, i):
        lib().or_ctx_lock_variable(self._h, i)

    def begin_matrix(self):
        lib().or_ctx_begin_matrix(self._h)

    def row(self, coeffs, rhs: float = 0.0):
        L = lib()
        L.or_ctx_begin_row(self._h)
        for v, a in coeffs:
            L.or_ctx_coefficient(self._h, int(v), float(a))
        if rhs != 0.0:
            L.or_ctx_right_hand_side(self._h, float(rhs))
        L.or_ctx_end_row(self._h)

    def rows(self, rp, var, coef, rhs=None):
        """Rows from arrays (as many row() calls, looped in C)."""
        rp = _u32(rp)
        var = _u32(var)
        coef = _f64(coef)
        rhs = None if rhs is None else _f64(rhs)
        lib().or_ctx_rows(self._h, len(rp) - 1, rp.ctypes.data, var.ctypes.data,
                          coef.ctypes.data, None if rhs is None else rhs.ctypes.data)

    def system(self):
        n = int(lib().or_ctx_nb_free(self._h))
        nnz = int(lib().or_ctx_nnz(self._h))
        rp = np.empty(n + 1, dtype=np.uint32)
        ci = np.empty(max(nnz, 1), dtype=np.uint32)
        va = np.empty(max(nnz, 1), dtype=np.float64)
        b = np.empty(max(n, 1), dtype=np.float64)
        lib().or_ctx_system(self._h, _p(rp, C.c_uint32), _p(ci, C.c_uint32),
                            _p(va, C.c_double), _p(b, C.c_double))
        return CRS(rp, ci[:nnz], va[:nnz], n), b[:n]

    def solve(self, eps=1e-10, max_iter=5000):
        rel2 = C.c_double(0.0)
        it = lib().or_ctx_solve(self._h, float(eps), int(max_iter), C.byref(rel2))
        return int(it), float(np.sqrt(rel2.value))
