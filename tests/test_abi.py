"""The C-ABI library: it loads, exports every function include/*.h declares,
keeps the host-side OpenNL semantics (dynamic matrices, compression, the
OGF::NL wrapper's error behaviour) and refuses to compute without a GPU.
No kernel is launched here."""
import ctypes as C
import logging
import subprocess

import numpy as np
import pytest

from graphitethree_amd import _native, capi, nl
from tests.conftest import gpu_present


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    declared = capi.declared_symbols()
    assert len(declared) >= 60
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(capi.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", _native.path("libnlhip.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert set(declared) <= exported


def test_workloads_library_loads():
    from graphitethree_amd import workloads as W
    assert W.lib().wl_sys_n


@pytest.mark.skipif(gpu_present(), reason="checks the no-GPU behaviour")
def test_no_device_fails_loudly():
    L = capi.lib()
    assert L.nlInitExtension(b"HIP") == 0
    assert "device" in capi.last_error()
    with pytest.raises(capi.NLHIPError) as e:
        capi.System([0, 1], [0], [2.0])
    assert e.value.code == capi.NLHIP_ERR_NO_DEVICE
    M = L.nlSparseMatrixNew(2, 2, capi.NL_MATRIX_STORE_ROWS)
    L.nlSparseMatrixAdd(M, 0, 0, 2.0)
    L.nlSparseMatrixAdd(M, 1, 1, 2.0)
    b = np.array([1.0, 1.0]); x = np.zeros(2)
    its = L.nlSolveSystemIterative(L.nlHostBlas(), M, None, capi.ptr(b), capi.ptr(x),
                                   capi.NL_CG, 1e-10, 100, 5)
    assert its == 0 and not x.any()
    L.nlDeleteMatrix(M)


def test_other_extensions_are_absent():
    L = capi.lib()
    for name in (b"SUPERLU", b"CHOLMOD", b"ARPACK", b"CUDA", b"MKL"):
        assert L.nlInitExtension(name) == 0


def _crs_arrays(M):
    class CRS(C.Structure):
        _fields_ = [("m", C.c_uint), ("n", C.c_uint), ("type", C.c_uint),
                    ("destroy_func", C.c_void_p), ("mult_func", C.c_void_p),
                    ("rowptr", C.POINTER(C.c_uint)), ("colind", C.POINTER(C.c_uint)),
                    ("val", C.POINTER(C.c_double))]
    c = C.cast(M, C.POINTER(CRS)).contents
    nnz = c.rowptr[c.m]
    return (np.array(c.rowptr[:c.m + 1]), np.array(c.colind[:nnz]), np.array(c.val[:nnz]))


def test_dynamic_matrix_compress_matches_oracle():
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    I = rng.integers(0, 20, 300); J = rng.integers(0, 20, 300); V = rng.standard_normal(300)
    L = capi.lib()
    M = L.nlSparseMatrixNew(20, 20, capi.NL_MATRIX_STORE_ROWS)
    D = O.Dyn(20, 20)
    for i, j, v in zip(I, J, V):
        L.nlSparseMatrixAdd(M, int(i), int(j), float(v))
        D.add(int(i), int(j), float(v))
    assert M.contents.type == capi.NL_MATRIX_SPARSE_DYNAMIC
    ref = D.compress()
    assert L.nlSparseMatrixNNZ(M) == ref.nnz
    L.nlMatrixCompress(C.byref(M))
    assert M.contents.type == capi.NL_MATRIX_CRS
    rp, ci, va = _crs_arrays(M)
    assert np.array_equal(rp, ref.rowptr) and np.array_equal(ci, ref.colind)
    assert np.array_equal(va, ref.val)
    L.nlDeleteMatrix(M)


def test_wrapper_error_behaviour(caplog):
    NL = nl.Library()
    M = NL.create_matrix(2, 3)
    assert M.format == nl.Matrix.Dynamic and (M.m, M.n) == (2, 3)
    with caplog.at_level(logging.ERROR, logger="NL"):
        M.add_coefficient(2, 0, 1.0)
        assert "index out of bound" in caplog.text
        caplog.clear()
        xi = nl.Vector(3, element_type=np.int32)
        y = nl.Vector(2)
        M.mult(xi, y)
        assert "not a vector of doubles" in caplog.text
        caplog.clear()
        M.solve_iterative(nl.Vector(2), nl.Vector(2))
        assert "not square" in caplog.text
        caplog.clear()
        v = nl.Vector.view(np.zeros(4))
        v.resize(8)
        assert "does not own memory" in caplog.text and v.nb_elements == 4
        caplog.clear()
        assert NL.blas.dot(nl.Vector(3), nl.Vector(4)) == 0.0
        assert "different sizes" in caplog.text
        caplog.clear()
        assert M.factorize("SuperLU") is None
        assert "Could not factorize" in caplog.text
    M.compress()
    assert M.format == nl.Matrix.CRS
    with caplog.at_level(logging.ERROR, logger="NL"):
        M.add_coefficient(0, 0, 1.0)
        assert "wrong matrix type" in caplog.text


def test_vector_view_is_zero_copy():
    a = np.arange(6, dtype=np.float64)
    v = nl.Vector.view(a, dimension=2)
    assert v.size == 3 and v.dimension == 2 and v.nb_elements == 6
    v[1] = 42.0
    assert a[1] == 42.0


def test_context_state_errors_without_gpu(capfd):
    L = capi.lib()
    ctx = L.nlNewContext()
    L.nlSolverParameteri(capi.NL_NB_VARIABLES, 3)
    L.nlSolverParameteri(capi.NL_LEAST_SQUARES, capi.NL_TRUE)
    L.nlBegin(capi.NL_SYSTEM)
    L.nlBegin(capi.NL_ROW)  # outside NL_MATRIX
    assert "outside NL_MATRIX" in capi.last_error()
    L.nlLockVariable(0)
    L.nlSetVariable(0, 1.0)
    L.nlBegin(capi.NL_MATRIX)
    L.nlBegin(capi.NL_ROW)
    L.nlCoefficient(0, 1.0)
    L.nlCoefficient(1, -1.0)
    L.nlEnd(capi.NL_ROW)
    L.nlCoefficient(7, 1.0)  # outside a row
    assert "outside NL_ROW" in capi.last_error()
    L.nlEnd(capi.NL_MATRIX)
    L.nlEnd(capi.NL_SYSTEM)
    v = C.c_int()
    # least-squares rows are assembled on the device: no device, an error
    L.nlGetIntegerv(capi.NL_NNZ, C.byref(v))
    assert "no HIP device" in capi.last_error()
    assert L.nlGetVariable(0) == 1.0 and L.nlVariableIsLocked(0)
    L.nlDeleteContext(ctx)
    assert not L.nlGetCurrent()


def test_context_rows_then_ij_coefficients_accumulate_on_host():
    """An nlAddIJCoefficient after least-squares rows sends the rows through
    the host accumulation first (OpenNL's order of additions)."""
    L = capi.lib()
    ctx = L.nlNewContext()
    L.nlSolverParameteri(capi.NL_NB_VARIABLES, 3)
    L.nlSolverParameteri(capi.NL_LEAST_SQUARES, capi.NL_TRUE)
    L.nlBegin(capi.NL_SYSTEM)
    L.nlLockVariable(0)
    L.nlSetVariable(0, 1.0)
    L.nlBegin(capi.NL_MATRIX)
    L.nlBegin(capi.NL_ROW)
    L.nlCoefficient(0, 1.0)
    L.nlCoefficient(1, -1.0)
    L.nlEnd(capi.NL_ROW)
    L.nlAddIJCoefficient(1, 1, 0.0)
    L.nlEnd(capi.NL_MATRIX)
    L.nlEnd(capi.NL_SYSTEM)
    v = C.c_int()
    L.nlGetIntegerv(capi.NL_NNZ, C.byref(v))
    assert v.value == 1   # one free-free entry (1,1)
    L.nlDeleteContext(ctx)


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N > 1 exits with a message before touching a
    device (this build runs on one GPU)."""
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)
