"""Python mirror of Graphite's OGF::NL scripting classes over libnlhip.so.

Same classes, methods, argument meanings, defaults and error behaviour as
the reference wrapper:

* ``Library``  src/lib/OGF/scene_graph/NL/library.h:54-94, library.cpp:44-57
* ``Vector``   vector.h:53-208 (owned storage, or a zero-copy view of a numpy
  buffer as plugins/OGF/gompy/interpreter/nl_vector_interop.cpp:100-176 does)
* ``Matrix``   matrix.h:51-234, matrix.cpp:120-440
* ``Blas``     blas.h:55-112, blas.cpp:65-148

Errors are reported like ``Logger::err("NL")`` followed by a plain return:
a message on the ``NL`` logger, nothing raised, no side effect.  One quirk is
not reproduced: ``Blas::dot`` and ``Blas::axpy`` check the type of ``x``
twice and never ``y`` (blas.cpp:105,134); here ``y`` is checked.

Every product and solve runs on the GPU through the OpenNL C API.
"""
from __future__ import annotations

import ctypes as C
import logging

import numpy as np

from . import capi

log = logging.getLogger("NL")


def _err(msg: str) -> None:
    log.error(msg)


# ------------------------------------------------------------------ Vector

class Vector:
    """A contiguous vector of ``size`` items of ``dimension`` elements."""

    def __init__(self, size: int = 0, dimension: int = 1, element_type=np.float64,
                 data: np.ndarray | None = None, read_only: bool = False):
        if data is not None:
            if not data.flags.c_contiguous:
                raise ValueError("Vector views need a C-contiguous array")
            self._a = data.reshape(-1)
            self._owns = False
            self._dim = dimension
            self._read_only = read_only
        else:
            self._a = np.zeros(size * dimension, dtype=element_type)
            self._owns = True
            self._dim = dimension
            self._read_only = False

    @classmethod
    def view(cls, array: np.ndarray, dimension: int = 1, read_only: bool = False) -> "Vector":
        """Zero-copy view of a numpy buffer (gompy's python_to_nl_vector)."""
        return cls(data=array, dimension=dimension, read_only=read_only)

    # properties (vector.h:135-174)
    @property
    def nb_elements(self) -> int:
        return self._a.size

    @property
    def dimension(self) -> int:
        return self._dim

    @property
    def size(self) -> int:
        return self._a.size // max(self._dim, 1)

    @property
    def element_type(self):
        return self._a.dtype

    @property
    def read_only(self) -> bool:
        return self._read_only

    @property
    def data(self) -> np.ndarray:
        return self._a

    def __len__(self):
        return self.nb_elements

    def __getitem__(self, i):
        return self._a[i]

    def __setitem__(self, i, v):
        if self._read_only:
            _err("NL::Vector is read-only")
            return
        self._a[i] = v

    def resize(self, new_size: int, new_dim: int = 1, element_type=None) -> None:
        """vector.cpp:123-164: previous values are not kept; views cannot resize."""
        if (new_size == self.size and new_dim == self._dim and
                (element_type is None or np.dtype(element_type) == self._a.dtype)):
            return
        if not self._owns or self._read_only:
            _err("Cannot resize, Vector does not own memory")
            return
        dt = np.float64 if element_type is None else element_type
        self._a = np.zeros(new_size * new_dim, dtype=dt)
        self._dim = new_dim

    def _is_double(self) -> bool:
        return self._a.dtype == np.float64

    def _ptr(self):
        return C.c_void_p(self._a.ctypes.data)


# ------------------------------------------------------------------ Blas

class Blas:
    """BLAS level 1 on Vectors (nlHostBlas(): staged to the GPU)."""

    def __init__(self):
        self._impl = capi.lib().nlHostBlas()

    def _t(self):
        return self._impl.contents

    @staticmethod
    def _check_type(v: Vector, fn: str, arg: str) -> bool:
        if not v._is_double():
            _err(f"{fn} arg {arg} is a vector of {v.element_type} (expected double)")
            return False
        return True

    def copy(self, x: Vector, y: Vector) -> None:
        if not (self._check_type(x, "Blas::copy", "x") and self._check_type(y, "Blas::copy", "y")):
            return
        if x.nb_elements != y.nb_elements:
            _err("Blas::copy Vectors x and y have different sizes")
            return
        self._t().Dcopy(self._impl, x.nb_elements, x._ptr(), 1, y._ptr(), 1)

    def scal(self, a: float, x: Vector) -> None:
        if not self._check_type(x, "Blas::scal", "x"):
            return
        self._t().Dscal(self._impl, x.nb_elements, float(a), x._ptr(), 1)

    def dot(self, x: Vector, y: Vector) -> float:
        if not (self._check_type(x, "Blas::dot", "x") and self._check_type(y, "Blas::dot", "y")):
            return 0.0
        if x.nb_elements != y.nb_elements:
            _err("Blas::dot Vectors x and y have different sizes")
            return 0.0
        return float(self._t().Ddot(self._impl, x.nb_elements, x._ptr(), 1, y._ptr(), 1))

    def nrm2(self, x: Vector) -> float:
        if not self._check_type(x, "Blas::nrm2", "x"):
            return 0.0
        return float(self._t().Dnrm2(self._impl, x.nb_elements, x._ptr(), 1))

    def axpy(self, a: float, x: Vector, y: Vector) -> None:
        if not (self._check_type(x, "Blas::axpy", "x") and self._check_type(y, "Blas::axpy", "y")):
            return
        if x.nb_elements != y.nb_elements:
            _err("Blas::axpy Vectors x and y have different sizes")
            return
        self._t().Daxpy(self._impl, x.nb_elements, float(a), x._ptr(), 1, y._ptr(), 1)


# ------------------------------------------------------------------ Matrix

class Matrix:
    """Sparse matrix: Dynamic (add_coefficient) then CRS after compress()."""

    Dynamic, CRS, Factorized = "Dynamic", "CRS", "Factorized"
    SuperLU, SuperLU_perm, SuperLU_sym, Cholmod = (
        "SuperLU", "SuperLU_perm", "SuperLU_sym", "Cholmod")
    CG, GMRES, BiCGSTAB = "CG", "GMRES", "BiCGSTAB"
    PRECOND_NONE, Jacobi, SSOR = "None", "Jacobi", "SSOR"

    def __init__(self, m: int, n: int):
        L = capi.lib()
        self._impl = L.nlSparseMatrixNew(m, n, capi.NL_MATRIX_STORE_ROWS)
        if not self._impl:
            raise RuntimeError(capi.last_error())

    def __del__(self):
        if getattr(self, "_impl", None):
            capi.lib().nlDeleteMatrix(self._impl)
            self._impl = None

    @property
    def m(self) -> int:
        return int(self._impl.contents.m)

    @property
    def n(self) -> int:
        return int(self._impl.contents.n)

    @property
    def format(self) -> str:
        t = self._impl.contents.type
        if t == capi.NL_MATRIX_SPARSE_DYNAMIC:
            return Matrix.Dynamic
        if t == capi.NL_MATRIX_CRS:
            return Matrix.CRS
        return Matrix.Factorized

    def _name(self) -> str:
        return f"Matrix({self.m},{self.n})"

    def add_coefficient(self, i: int, j: int, a: float) -> None:
        """matrix.cpp:146-164."""
        if self.format != Matrix.Dynamic:
            _err("Matrix::add_coefficient() called on wrong matrix type")
            return
        if i < 0 or j < 0 or i >= self.m or j >= self.n:
            _err(f"{self._name()}::add_coefficient({i},{j}) index out of bound")
            return
        capi.lib().nlSparseMatrixAdd(self._impl, i, j, float(a))

    def add_coefficients(self, I: Vector, J: Vector, A: Vector, ignore_OOB: bool = False) -> None:
        """matrix.cpp:167-261: integer I, J; double A; duplicates accumulate."""
        ints = (np.uint32, np.int32)
        if I.element_type not in ints:
            _err(f"{self._name()}::add_coefficients() I is not an integer vector")
            return
        if J.element_type not in ints:
            _err(f"{self._name()}::add_coefficients() J is not an integer vector")
            return
        if not A._is_double():
            _err(f"{self._name()}::add_coefficients() A is not a vector of doubles")
            return
        if I.dimension != 1 or J.dimension != 1 or A.dimension != 1:
            _err(f"{self._name()}::add_coefficients() wrong dimension")
            return
        if I.nb_elements != J.nb_elements or J.nb_elements != A.nb_elements:
            _err(f"{self._name()}::add_coefficients() vector size mismatch")
            return
        if self.format != Matrix.Dynamic:
            _err("Matrix::add_coefficients() called on wrong matrix type")
            return
        ii = I.data.astype(np.int64)
        jj = J.data.astype(np.int64)
        aa = A.data
        ok = (ii >= 0) & (jj >= 0) & (ii < self.m) & (jj < self.n)
        if not ignore_OOB and not ok.all():
            _err(f"{self._name()}::add_coefficients() coefficient larger than matrix size")
            return
        add = capi.lib().nlSparseMatrixAdd
        for k in np.nonzero(ok)[0]:
            add(self._impl, int(ii[k]), int(jj[k]), float(aa[k]))

    def add_coefficients_to_diagonal(self, A: Vector) -> None:
        """matrix.cpp:263-294."""
        if not A._is_double():
            _err(f"{self._name()}::add_coefficients() A is not a vector of doubles")
            return
        if A.dimension != 1:
            _err(f"{self._name()}::add_coefficients() wrong dimension")
            return
        if A.nb_elements > self.m or A.nb_elements > self.n:
            _err(f"{self._name()}::add_coefficients() wrong size")
            return
        if self.format != Matrix.Dynamic:
            _err("Matrix::add_coefficients_to_diagonal() called on wrong matrix type")
            return
        add = capi.lib().nlSparseMatrixAdd
        for k in range(A.nb_elements):
            add(self._impl, k, k, float(A.data[k]))

    def _check_vector(self, v: Vector, check_m: bool, check_n: bool, fn: str, arg: str) -> bool:
        """matrix.cpp:74-113."""
        full = f"OGF::NL::{self._name()}::{fn} {arg}({v.nb_elements})"
        if not v._is_double():
            _err(f"{full} is not a vector of doubles")
            return False
        if check_m and v.nb_elements != self.m:
            _err(f"{full} vector size does not match number of rows")
            return False
        if check_n and v.nb_elements != self.n:
            _err(f"{full} vector size does not match number of columns")
            return False
        return True

    def mult(self, x: Vector, y: Vector) -> None:
        """matrix.cpp:296-306: y = M x (on the GPU)."""
        if not self._check_vector(x, False, True, "mult", "x"):
            return
        if not self._check_vector(y, True, False, "mult", "x"):
            return
        capi.lib().nlMultMatrixVector(self._impl, x._ptr(), y._ptr())

    def compress(self) -> None:
        """matrix.cpp:308-310."""
        # nlMatrixCompress replaces the handle in place (dynamic -> CRS)
        capi.lib().nlMatrixCompress(C.byref(self._impl))

    def factorize(self, factorization: str = "SuperLU_perm"):
        """matrix.cpp:312-337: direct solvers are extensions this library lacks."""
        _err("Could not factorize matrix")
        return None

    def solve_iterative(self, b: Vector, x: Vector, solver: str = "BiCGSTAB",
                        precond: str = "None", max_iter: int = 1000, eps: float = 1e-3) -> None:
        """matrix.cpp:340-398: x is zeroed, then nlSolveSystemIterative."""
        if self.m != self.n:
            _err("Matrix::solve_iterative() matrix is not square")
            return
        if not self._check_vector(b, True, True, "solve_iterative", "b"):
            return
        if not self._check_vector(x, True, True, "solve_iterative", "x"):
            return
        L = capi.lib()
        P = None
        if precond == Matrix.Jacobi:
            P = L.nlNewJacobiPreconditioner(self._impl)
        elif precond == Matrix.SSOR:
            P = L.nlNewSSORPreconditioner(self._impl, 1.5)
            if not P:
                return
        nl_solver = {Matrix.CG: capi.NL_CG, Matrix.GMRES: capi.NL_GMRES,
                     Matrix.BiCGSTAB: capi.NL_BICGSTAB}.get(solver, capi.NL_BICGSTAB)
        inner_iter = 5
        x.data[:] = 0.0
        L.nlSolveSystemIterative(L.nlHostBlas(), self._impl, P if P else None, b._ptr(),
                                 x._ptr(), nl_solver, float(eps), int(max_iter), inner_iter)
        if P:
            L.nlDeleteMatrix(P)

    def solve_direct(self, b: Vector, x: Vector, factorization: str = "SuperLU_perm") -> None:
        """matrix.cpp:401-424."""
        if self.m != self.n:
            _err("Matrix::solve_direct() matrix is not square")
            return
        if not self._check_vector(b, True, True, "solve_iterative", "b"):
            return
        if not self._check_vector(x, True, True, "solve_iterative", "x"):
            return
        F = self.factorize(factorization)
        if F is None:
            return

    def solve_symmetric(self, b: Vector, x: Vector, direct_solver: bool = False) -> None:
        """matrix.cpp:426-440: CG + Jacobi, max_iter = 5 n, eps = 1e-3."""
        L = capi.lib()
        if direct_solver:
            if L.nlInitExtension(b"CHOLMOD"):
                self.solve_direct(b, x, Matrix.Cholmod)
                return
            if L.nlInitExtension(b"SUPERLU"):
                self.solve_direct(b, x, Matrix.SuperLU_sym)
                return
        self.solve_iterative(b, x, Matrix.CG, Matrix.Jacobi, self.n * 5, 1e-3)


# ------------------------------------------------------------------ Library

class Library:
    """The ``NL`` object scripts use: vectors, matrices and the BLAS."""

    def __init__(self):
        self._blas = Blas()

    def create_vector(self, dim: int = 0) -> Vector:
        return Vector() if dim == 0 else Vector(dim)

    def create_matrix(self, m: int, n: int) -> Matrix:
        return Matrix(m, n)

    @property
    def blas(self) -> Blas:
        return self._blas
