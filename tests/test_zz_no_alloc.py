"""nlhipSolveNoAlloc (the solve whose work buffers live in the library's
fixed device array) against nlhipSolve on the same systems: bitwise the same
iterate, iteration counts and residuals, on two calls in a row (the second on
a smaller system, so stale contents of the array from the first call would
show), with host and with device vectors, and the size limit rejected with
an error."""
import ctypes as C

import numpy as np
import pytest

from graphitethree_amd import capi
from graphitethree_amd import workloads as W

pytestmark = pytest.mark.gpu


def no_alloc(L, S, b, x0=None, **params):
    x = np.zeros(S.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    p = capi.default_params(**params)
    st = capi.SolveStats()
    capi.check(L.nlhipSolveNoAlloc(S.handle, capi.ptr(b), capi.ptr(x), C.byref(p), C.byref(st)),
               "nlhipSolveNoAlloc")
    return x, st


def same(a, sa, b, sb):
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert sa.iterations == sb.iterations and sa.extra_iterations == sb.extra_iterations
    assert sa.outer_iterations == sb.outer_iterations and sa.converged == sb.converged
    assert sa.rel_residual == sb.rel_residual and sa.cg_rel_residual == sb.cg_rel_residual


def test_no_alloc_matches_nlhip_solve(hip):
    L = hip
    big = W.lscm_grid(160, seed=3)     # 51,198 rows
    small = W.lscm_grid(40, seed=4)    # 3,198 rows
    for Sw in (big, small):
        S = capi.System.from_crs(Sw)
        xr, sr = S.solve(Sw.b, eps=1e-10, max_iter=100000)
        xa, sa = no_alloc(L, S, Sw.b, eps=1e-10, max_iter=100000)
        print(f"n={Sw.n}: {sa.iterations} it, rel res {sa.rel_residual:.3e}")
        assert sa.converged == 1 and sa.rel_residual <= 1e-10
        same(xa, sa, xr, sr)
        # warm start and no preconditioner
        xr2, sr2 = S.solve(Sw.b, x0=xr * 0.5, precond="none", eps=1e-8, max_iter=100000)
        xa2, sa2 = no_alloc(L, S, Sw.b, x0=xr * 0.5, precond="none", eps=1e-8, max_iter=100000)
        same(xa2, sa2, xr2, sr2)
        S.close()


def test_no_alloc_device_vectors(hip):
    L = hip
    Sw = W.lscm_grid(50, seed=5)
    S = capi.System.from_crs(Sw)
    xr, sr = S.solve(Sw.b)
    blas = L.nlHIPBlas()
    t = blas.contents
    Malloc = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_int, C.c_size_t)(t.Malloc)
    Memcpy = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                         C.c_size_t)(t.Memcpy)
    Free = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_size_t, C.c_void_p)(t.Free)
    nb = 8 * Sw.n
    db, dx = Malloc(blas, 1, nb), Malloc(blas, 1, nb)
    zero = np.zeros(Sw.n)
    Memcpy(blas, db, 1, Sw.b.ctypes.data, 0, nb)
    Memcpy(blas, dx, 1, zero.ctypes.data, 0, nb)
    p = capi.default_params()
    st = capi.SolveStats()
    capi.check(L.nlhipSolveNoAlloc(S.handle, C.c_void_p(db), C.c_void_p(dx), C.byref(p),
                                   C.byref(st)), "nlhipSolveNoAlloc")
    x = np.empty(Sw.n)
    Memcpy(blas, x.ctypes.data, 0, dx, 1, nb)
    Free(blas, 1, nb, db)
    Free(blas, 1, nb, dx)
    same(x, st, xr, sr)
    S.close()


def test_no_alloc_rejects_what_it_cannot_run(hip):
    L = hip
    Sw = W.lscm_grid(20, seed=1)
    S = capi.System.from_crs(Sw, fp32=True)
    with pytest.raises(capi.NLHIPError) as e:
        no_alloc(L, S, Sw.b, precision="mixed")
    assert e.value.code == capi.NLHIP_ERR_UNSUPPORTED
    S.close()
    n = 2_097_153                      # one row past the fixed array
    rp = np.arange(n + 1, dtype=np.uint32)
    S = capi.System(rp, np.arange(n, dtype=np.uint32), np.full(n, 2.0))
    with pytest.raises(capi.NLHIPError) as e:
        no_alloc(L, S, np.ones(n))
    assert e.value.code == capi.NLHIP_ERR_INVALID
    assert "2097152" in capi.last_error()
    S.close()
    # the array still works after the rejections
    Sw = W.lscm_grid(30, seed=2)
    S = capi.System.from_crs(Sw)
    xr, sr = S.solve(Sw.b)
    xa, sa = no_alloc(L, S, Sw.b)
    same(xa, sa, xr, sr)
    S.close()
