"""GPU Jacobi-PCG parity with the oracle's restatement of OpenNL's CG.

Bar: the same stopping rule (|Ax-b| <= eps |b| or max_iter), the same
iteration count (within a few percent: reductions are summed in another order),
and a relative difference |x_gpu - x_cpu| / |x_cpu| <= 1e-8 at eps = 1e-10.
The forward error against a scipy direct solve is checked where one is
affordable."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from graphitethree_amd import capi, nl
from graphitethree_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
REL_DIFF = 1e-8


def same_iters(a, b) -> bool:
    """Stopping iterations agree up to reduction-order rounding.  On these
    ill-conditioned LSCM systems the residual approaches 1e-10 slowly, so a
    different summation order of the dot products can move the crossing by a
    few percent; the solution bar (REL_DIFF) is the real check."""
    a, b = int(a), int(b)
    return abs(a - b) <= max(2, 0.05 * max(a, b))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_lua_kat_solve_symmetric(hip):
    k = json.load(open(os.path.join(GOLD, "lua_kats.json")))["symmetric"]
    NL = nl.Library()
    B = NL.create_vector(2)
    B[0], B[1] = k["B"]
    M = NL.create_matrix(2, 2)
    for i in range(2):
        for j in range(2):
            M.add_coefficient(i, j, k["M"][i][j])
    X = nl.Vector(2)
    M.solve_symmetric(B, X)
    np.testing.assert_allclose(X.data, k["X"], rtol=0, atol=1e-12)
    Bcheck = NL.create_vector(2)
    M.mult(X, Bcheck)
    np.testing.assert_allclose(Bcheck.data, k["B"], rtol=0, atol=1e-12)


def test_lua_kat_bicgstab(hip):
    k = json.load(open(os.path.join(GOLD, "lua_kats.json")))["general"]
    NL = nl.Library()
    B = NL.create_vector(2)
    B[0], B[1] = k["B"]
    M = NL.create_matrix(2, 2)
    for i in range(2):
        for j in range(2):
            M.add_coefficient(i, j, k["M"][i][j])
    X = nl.Vector(2)
    M.solve_iterative(B, X, solver="BiCGSTAB", precond="None", max_iter=1000, eps=0.001)
    np.testing.assert_allclose(X.data, k["X"], atol=1e-3 * np.linalg.norm(k["X"]))
    ref, _, _ = O.bicgstab(O.CRS([0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 3.0, 4.0]), k["B"],
                           eps=1e-3, max_iter=1000)
    assert rel(X.data, ref) < 1e-12


@pytest.mark.parametrize("k", [12, 40, 80])
def test_pcg_parity_lscm(hip, k):
    Sw = W.lscm_grid(k, seed=1)
    A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
    x_cpu, it_cpu, rel_cpu = O.pcg(A, Sw.b, eps=1e-10, max_iter=5000)
    S = capi.System.from_crs(Sw)
    x_gpu, st = S.solve(Sw.b, eps=1e-10, max_iter=5000)
    print(f"k={k} n={Sw.n} it cpu/gpu {it_cpu}/{st.iterations} (+{st.extra_iterations} "
          f"confirming) rel_res {st.rel_residual:.3e} rel_diff {rel(x_gpu, x_cpu):.3e}")
    # iterations at OpenNL's (recursive) stopping rule, before the confirmation
    assert same_iters(st.iterations - st.extra_iterations, it_cpu)
    assert st.converged == 1 and st.rel_residual <= 1e-10
    r = Sw.b - O.spmv(A, x_gpu)
    assert np.linalg.norm(r) / np.linalg.norm(Sw.b) <= 1e-10
    assert rel(x_gpu, x_cpu) <= REL_DIFF
    if k == 12:
        z = np.load(os.path.join(GOLD, "lscm_k12.npz"))
        assert np.array_equal(z["val"], Sw.val)
        assert rel(x_gpu, z["x"]) < 1e-8


def test_pcg_deterministic_and_warm_start(hip):
    Sw = W.lscm_grid(30, seed=9)
    S = capi.System.from_crs(Sw)
    x1, s1 = S.solve(Sw.b)
    x2, s2 = S.solve(Sw.b)
    assert np.array_equal(x1, x2) and s1.iterations == s2.iterations
    x3, s3 = S.solve(Sw.b, x0=x1)
    assert s3.iterations == 0 and np.array_equal(x3, x1)


def test_pcg_launch_shapes_agree(hip):
    Sw = W.lscm_grid(36, seed=4)
    S = capi.System.from_crs(Sw)
    base, st0 = S.solve(Sw.b)
    for rpw, wg in [(2, 7), (8, 64), (16, 1000), (64, 3)]:
        S.set_launch(rpw, wg)
        x, st = S.solve(Sw.b)
        assert same_iters(st.iterations, st0.iterations)
        assert rel(x, base) <= REL_DIFF


def test_pcg_edge_cases(hip):
    Sw = W.lscm_grid(10, seed=1)
    S = capi.System.from_crs(Sw)
    x, st = S.solve(np.zeros(Sw.n))
    assert st.iterations == 0 and not x.any() and st.converged == 1
    x, st = S.solve(Sw.b, max_iter=1)
    assert st.iterations == 1 and st.converged == 0
    b = Sw.b.copy()
    b[3] = np.nan
    with pytest.raises(capi.NLHIPError) as e:
        S.solve(b)
    assert e.value.code == capi.NLHIP_ERR_NONFINITE
    with pytest.raises(capi.NLHIPError):
        capi.System([0, 2, 1], [0, 1], [1.0, 1.0])        # rowptr not monotone
    with pytest.raises(capi.NLHIPError):
        capi.System([0, 1, 2], [0, 5], [1.0, 1.0])        # colind out of range
    with pytest.raises(capi.NLHIPError) as e:
        capi.System([0, 1, 2], [0, 1], [1.0, np.inf])
    assert e.value.code == capi.NLHIP_ERR_NONFINITE
    x, st = capi.System([0, 1], [0], [4.0]).solve([2.0])   # n = 1
    assert x[0] == 0.5


def test_injected_failure_is_reported(hip):
    Sw = W.lscm_grid(8, seed=1)
    S = capi.System.from_crs(Sw)
    hip.nlhipInjectFailures(1)
    with pytest.raises(capi.NLHIPError) as e:
        S.solve(Sw.b)
    assert e.value.code == capi.NLHIP_ERR_INJECTED
    x, st = S.solve(Sw.b)
    assert st.converged == 1


def test_mixed_precision_refinement(hip):
    """FP32 inner PCG inside FP64 refinement: the true residual (recomputed
    by the oracle's SpMV) meets 1e-10 and the solution matches the oracle's
    FP64 PCG within 1e-8, like the FP64 GPU solve."""
    Sw = W.lscm_grid(30, seed=2)
    S = capi.System.from_crs(Sw, fp32=True)
    A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
    x_cpu, _, _ = O.pcg(A, Sw.b, eps=1e-10, max_iter=20000)
    xm, sm = S.solve(Sw.b, precision="mixed", eps=1e-10, max_iter=200000, inner_eps=1e-4)
    print(f"mixed it {sm.iterations} outer {sm.outer_iterations} "
          f"rel_res {sm.rel_residual:.2e} diff vs oracle {rel(xm, x_cpu):.2e}")
    assert sm.converged == 1 and sm.rel_residual <= 1e-10
    assert sm.outer_iterations >= 2
    r = O.spmv(A, xm) - Sw.b
    assert np.linalg.norm(r) / np.linalg.norm(Sw.b) <= 1e-10
    assert rel(xm, x_cpu) <= REL_DIFF


@pytest.mark.parametrize("k", [80, 150])
def test_mixed_precision_parity_larger(hip, k):
    """The same on LSCM grids of 12,796 and 44,996 unknowns (FP32 needs
    3-6x the FP64 iterations there)."""
    Sw = W.lscm_grid(k, seed=2)
    S = capi.System.from_crs(Sw, fp32=True)
    A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
    x_cpu, _, _ = O.pcg(A, Sw.b, eps=1e-10, max_iter=200000, variant=1)
    for prec, kw in (("mixed", {}), ("mixed-fallback", {"inner_max_iter": 20})):
        x, st = S.solve(Sw.b, precision=prec, eps=1e-10, max_iter=2000000, inner_eps=1e-4, **kw)
        tr = float(np.linalg.norm(O.spmv(A, x, parallel=True) - Sw.b) / np.linalg.norm(Sw.b))
        print(f"k={k} {prec}: {st.iterations} it, outer {st.outer_iterations}, true residual "
              f"{tr:.2e}, diff vs oracle {rel(x, x_cpu):.2e}")
        assert st.converged == 1 and tr <= 1e-10
        assert rel(x, x_cpu) <= REL_DIFF
    S.close()


def test_mixed_refinement_never_returns_a_worse_iterate(hip):
    """A refinement step whose FP32 correction does not lower the FP64
    residual is undone and ends the refinement: started from a converged FP64
    solution and asked for an unreachable 1e-18, the mixed solve stops
    without converging and its true residual is no larger than the one it
    started from.  With the FP64 fallback the solve continues in FP64 from
    there and stays at the 1e-10 bar."""
    Sw = W.lscm_grid(60, seed=3)
    S = capi.System.from_crs(Sw, fp32=True)
    x64, s64 = S.solve(Sw.b, eps=1e-10, max_iter=50000)
    assert s64.converged == 1
    xm, sm = S.solve(Sw.b, x0=x64, precision="mixed", eps=1e-18, max_iter=20000,
                     inner_eps=1e-4)
    assert sm.converged == 0 and np.all(np.isfinite(xm))
    assert sm.rel_residual <= s64.rel_residual, (sm.rel_residual, s64.rel_residual)
    xf, sf = S.solve(Sw.b, x0=x64, precision="mixed-fallback", eps=1e-18, max_iter=20000,
                     inner_eps=1e-4, inner_max_iter=50)
    assert sf.converged == 0 and np.all(np.isfinite(xf)) and sf.rel_residual <= 1e-10
    S.close()


def test_mixed_with_fp64_fallback(hip):
    """NLHIP_PREC_MIXED_FALLBACK on one system: an FP32 inner solve that uses
    its whole cap hands over to FP64 CG from the current iterate; a cap no
    inner solve reaches gives the plain mixed solve, bit for bit."""
    Sw = W.lscm_grid(30, seed=2)
    S = capi.System.from_crs(Sw, fp32=True)
    A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
    x_cpu, _, _ = O.pcg(A, Sw.b, eps=1e-10, max_iter=20000)
    xf, sf = S.solve(Sw.b, precision="mixed-fallback", eps=1e-10, max_iter=200000,
                     inner_eps=1e-4, inner_max_iter=20)
    assert sf.converged == 1 and sf.rel_residual <= 1e-10
    assert sf.outer_iterations == 1           # the first inner solve hit its cap
    r = O.spmv(A, xf) - Sw.b
    assert np.linalg.norm(r) / np.linalg.norm(Sw.b) <= 1e-10
    assert rel(xf, x_cpu) <= REL_DIFF
    xm, _ = S.solve(Sw.b, precision="mixed", eps=1e-10, max_iter=200000, inner_eps=1e-4,
                    inner_max_iter=100000)
    xn, sn = S.solve(Sw.b, precision="mixed-fallback", eps=1e-10, max_iter=200000,
                     inner_eps=1e-4, inner_max_iter=100000)
    assert np.array_equal(xm.view(np.uint64), xn.view(np.uint64))
    with pytest.raises(capi.NLHIPError):
        S.solve(Sw.b, precision="mixed-fallback", inner_max_iter=0)
    S.close()


def test_fp64_confirmation_restarts_on_drift(hip):
    """At eps = 1e-13 the recursive residual of the 80 x 80 LSCM system drifts
    below the true one; the confirmation restarts from the iterate (residual
    replacement) until the true residual, recomputed by the oracle's SpMV,
    meets eps.  With confirm = 0 the solve stops where OpenNL's rule stops."""
    Sw = W.lscm_grid(80, seed=2)
    S = capi.System.from_crs(Sw)
    A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
    x, st = S.solve(Sw.b, eps=1e-13, max_iter=100000)
    tr = float(np.linalg.norm(O.spmv(A, x) - Sw.b) / np.linalg.norm(Sw.b))
    x0, st0 = S.solve(Sw.b, eps=1e-13, max_iter=100000, confirm=0)
    print(f"confirm: {st.iterations} it ({st.outer_iterations} restarts, +{st.extra_iterations}),"
          f" true {tr:.3e}; OpenNL rule: {st0.iterations} it, true {st0.rel_residual:.3e}")
    assert st.converged == 1 and tr <= 1e-13 and st.rel_residual <= 1e-13
    assert st0.extra_iterations == 0 and st0.outer_iterations == 0
    assert st.iterations - st.extra_iterations == st0.iterations
    S.close()


def test_solve_system_iterative_host_api(hip):
    L = hip
    Sw = W.lscm_grid(25, seed=6)
    M = L.nlCRSMatrixNewFromArrays(Sw.n, Sw.n, capi.ptr(Sw.rowptr), capi.ptr(Sw.colind),
                                   capi.ptr(Sw.val))
    P = L.nlNewJacobiPreconditioner(M)
    b = Sw.b.copy()
    x = np.zeros(Sw.n)
    its = L.nlSolveSystemIterative(L.nlHostBlas(), M, P, capi.ptr(b), capi.ptr(x),
                                   capi.NL_CG, 1e-10, 5000, 5)
    xr, itr, _ = O.pcg(O.CRS(Sw.rowptr, Sw.colind, Sw.val), Sw.b, eps=1e-10, max_iter=5000)
    assert same_iters(its, itr) and rel(x, xr) <= REL_DIFF
    # no preconditioner
    x0 = np.zeros(Sw.n)
    its0 = L.nlSolveSystemIterative(L.nlHostBlas(), M, None, capi.ptr(b), capi.ptr(x0),
                                    capi.NL_CG, 1e-10, 20000, 5)
    xr0, itr0, _ = O.pcg(O.CRS(Sw.rowptr, Sw.colind, Sw.val), Sw.b, eps=1e-10,
                         max_iter=20000, precond=None)
    assert same_iters(its0, itr0) and rel(x0, xr0) <= REL_DIFF
    L.nlDeleteMatrix(P)
    L.nlDeleteMatrix(M)


def test_gmres_host_api_and_wrapper(hip):
    """nlSolveSystemIterative(NL_GMRES, inner_iter = m) against the oracle's
    GMRES(m), with and without Jacobi, on a nonsymmetric system; and the
    wrapper's solve_iterative(..., "GMRES") on the general lua KAT."""
    L = hip
    import scipy.sparse as sp
    n = 400
    rng = np.random.default_rng(17)
    A = sp.random(n, n, density=6.0 / n, random_state=17, format="csr")
    A = (A + sp.diags(4.0 + rng.random(n))).tocsr()
    A.sort_indices()
    rp, ci, va = A.indptr.astype(np.uint32), A.indices.astype(np.uint32), A.data
    b = rng.standard_normal(n)
    M = L.nlCRSMatrixNewFromArrays(n, n, capi.ptr(rp), capi.ptr(ci), capi.ptr(va))
    P = L.nlNewJacobiPreconditioner(M)
    for prec, pname in ((None, None), (P, "jacobi")):
        x = np.zeros(n)
        its = L.nlSolveSystemIterative(L.nlHostBlas(), M, prec, capi.ptr(b), capi.ptr(x),
                                       capi.NL_GMRES, 1e-11, 5000, 5)
        xr, itr, relr = O.gmres(O.CRS(rp, ci, va), b, eps=1e-11, max_iter=5000, m=5,
                                precond=pname)
        assert its > 0 and same_iters(its, itr), (its, itr)
        assert np.linalg.norm(A @ x - b) <= 1.01e-11 * np.linalg.norm(b)
        assert rel(x, xr) <= REL_DIFF
    L.nlDeleteMatrix(P)
    L.nlDeleteMatrix(M)
    k = json.load(open(os.path.join(GOLD, "lua_kats.json")))["general"]
    NL = nl.Library()
    B = NL.create_vector(2)
    B[0], B[1] = k["B"]
    Mw = NL.create_matrix(2, 2)
    for i in range(2):
        for j in range(2):
            Mw.add_coefficient(i, j, k["M"][i][j])
    X = nl.Vector(2)
    Mw.solve_iterative(B, X, solver="GMRES", precond="None", max_iter=1000, eps=1e-12)
    np.testing.assert_allclose(X.data, k["X"], atol=1e-9)


def test_hip_extension_objects(hip):
    L = hip
    Sw = W.lscm_grid(22, seed=3)
    M = L.nlCRSMatrixNewFromArrays(Sw.n, Sw.n, capi.ptr(Sw.rowptr), capi.ptr(Sw.colind),
                                   capi.ptr(Sw.val))
    H = L.nlHIPMatrixNewFromCRSMatrix(M)
    P = L.nlHIPJacobiPreconditionerNewFromCRSMatrix(M)
    assert H.contents.type == capi.NL_MATRIX_OTHER
    b = Sw.b.copy()
    x = np.zeros(Sw.n)
    its = L.nlSolveSystemIterative(L.nlHIPBlas(), H, P, capi.ptr(b), capi.ptr(x),
                                   capi.NL_CG, 1e-10, 5000, 5)
    xr, itr, _ = O.pcg(O.CRS(Sw.rowptr, Sw.colind, Sw.val), Sw.b, eps=1e-10, max_iter=5000)
    assert same_iters(its, itr) and rel(x, xr) <= REL_DIFF
    for m in (P, H, M):
        L.nlDeleteMatrix(m)


def _context_lscm(L, mesh, pins, uv, eps, max_iter):
    coeffs, ok = mesh.lscm_rows()
    L.nlNewContext()
    L.nlSolverParameteri(capi.NL_NB_VARIABLES, 2 * mesh.nv)
    L.nlSolverParameteri(capi.NL_LEAST_SQUARES, capi.NL_TRUE)
    L.nlSolverParameteri(capi.NL_SOLVER, capi.NL_CG)
    L.nlSolverParameteri(capi.NL_PRECONDITIONER, capi.NL_PRECOND_JACOBI)
    L.nlSolverParameterd(capi.NL_THRESHOLD, eps)
    L.nlSolverParameteri(capi.NL_MAX_ITERATIONS, max_iter)
    L.nlBegin(capi.NL_SYSTEM)
    for p, (u, v) in zip(pins, uv):
        L.nlSetVariable(2 * p, u)
        L.nlSetVariable(2 * p + 1, v)
        L.nlLockVariable(2 * p)
        L.nlLockVariable(2 * p + 1)
    L.nlBegin(capi.NL_MATRIX)
    for t in range(mesh.nt):
        if not ok[t]:
            continue
        tri = mesh.triangles[t]
        var = [2 * int(tri[e // 2]) + (e & 1) for e in range(6)]
        for row in (coeffs[t, :6], coeffs[t, 6:]):
            L.nlBegin(capi.NL_ROW)
            for v_, a in zip(var, row):
                L.nlCoefficient(v_, float(a))
            L.nlEnd(capi.NL_ROW)
    L.nlEnd(capi.NL_MATRIX)
    L.nlEnd(capi.NL_SYSTEM)
    assert L.nlSolve() == capi.NL_TRUE
    out = np.array([L.nlGetVariable(i) for i in range(2 * mesh.nv)])
    used = C.c_int()
    L.nlGetIntegerv(capi.NL_USED_ITERATIONS, C.byref(used))
    L.nlDeleteContext(L.nlGetCurrent())
    return out, used.value


def test_context_api_lscm_chart(hip):
    """The mesh_compute_LSCM call sequence through nlBegin/nlEnd(NL_ROW)."""
    mesh = W.chart_mesh(400, seed=11)
    pins = mesh.extreme_pins()
    uv = [(mesh.points[p, 0], mesh.points[p, 1]) for p in pins]
    gpu, used = _context_lscm(hip, mesh, pins, uv, 1e-10, 5000)
    ctx = O.Context(2 * mesh.nv, least_squares=True)
    coeffs, ok = mesh.lscm_rows()
    for p, (u, v) in zip(pins, uv):
        ctx.set_variable(2 * p, u); ctx.set_variable(2 * p + 1, v)
        ctx.lock_variable(2 * p); ctx.lock_variable(2 * p + 1)
    ctx.begin_matrix()
    for t in range(mesh.nt):
        tri = mesh.triangles[t]
        var = [2 * int(tri[e // 2]) + (e & 1) for e in range(6)]
        ctx.row(list(zip(var, coeffs[t, :6])))
        ctx.row(list(zip(var, coeffs[t, 6:])))
    it, _ = ctx.solve(eps=1e-10, max_iter=5000)
    cpu = np.array([ctx.get_variable(i) for i in range(2 * mesh.nv)])
    assert same_iters(used, it)
    assert rel(gpu, cpu) <= REL_DIFF
    assert gpu[2 * pins[0]] == uv[0][0]


def test_context_api_periodic_voxel_least_squares(hip):
    """WarpDrive's voxel interpolation (voxel_grob_transport_commands.cpp:
    231-287): periodic 7-point Laplacian rows, data voxels locked."""
    L = hip
    nu = nv = nw = 6
    N = nu * nv * nw
    rng = np.random.default_rng(3)
    data = rng.random(N)
    locked = rng.random(N) < 0.3
    idx = lambda u, v, w: (w * nv + v) * nu + u
    L.nlNewContext()
    L.nlSolverParameteri(capi.NL_LEAST_SQUARES, capi.NL_TRUE)
    L.nlSolverParameteri(capi.NL_NB_VARIABLES, N)
    L.nlSolverParameterd(capi.NL_THRESHOLD, 1e-12)
    L.nlSolverParameteri(capi.NL_MAX_ITERATIONS, 10000)
    L.nlBegin(capi.NL_SYSTEM)
    ctx = O.Context(N, least_squares=True)
    for i in range(N):
        L.nlSetVariable(i, data[i]); ctx.set_variable(i, data[i])
        if locked[i]:
            L.nlLockVariable(i); ctx.lock_variable(i)
    L.nlBegin(capi.NL_MATRIX)
    ctx.begin_matrix()
    for w in range(nw):
        for v in range(nv):
            for u in range(nu):
                row = [(idx(u, v, w), -6.0), (idx((u - 1) % nu, v, w), 1.0),
                       (idx((u + 1) % nu, v, w), 1.0), (idx(u, (v - 1) % nv, w), 1.0),
                       (idx(u, (v + 1) % nv, w), 1.0), (idx(u, v, (w - 1) % nw), 1.0),
                       (idx(u, v, (w + 1) % nw), 1.0)]
                L.nlBegin(capi.NL_ROW)
                for j, a in row:
                    L.nlCoefficient(j, a)
                L.nlEnd(capi.NL_ROW)
                ctx.row(row)
    L.nlEnd(capi.NL_MATRIX)
    L.nlEnd(capi.NL_SYSTEM)
    assert L.nlSolve() == capi.NL_TRUE
    gpu = np.array([L.nlGetVariable(i) for i in range(N)])
    L.nlDeleteContext(L.nlGetCurrent())
    # the oracle context starts from the same set values (warm start)
    ctx.solve(eps=1e-12, max_iter=10000)
    cpu = np.array([ctx.get_variable(i) for i in range(N)])
    assert rel(gpu, cpu) <= REL_DIFF
    assert np.array_equal(gpu[locked], data[locked])


def test_autotuned_system_matches_oracle(hip):
    """The autotuner picks a streamed engine and a workgroup count by timing;
    the tuned solve stays within the tolerance of the oracle and gives the
    same bits on every solve."""
    Sw = W.lscm_grid(60, seed=7)
    S = capi.System.from_crs(Sw)
    rpw, wg, t = S.autotune(iters=10)
    assert rpw in (68, 69, 70, 71, 73, 74, 75, 76, 77, 78, 79) and wg >= 1 and t > 0
    assert S.launch() == (rpw, wg)
    x1, st = S.solve(Sw.b, eps=1e-10, max_iter=20000)
    x2, _ = S.solve(Sw.b, eps=1e-10, max_iter=20000)
    assert np.array_equal(x1, x2)
    xr, itr, _ = O.pcg(O.CRS(Sw.rowptr, Sw.colind, Sw.val), Sw.b, eps=1e-10, max_iter=20000)
    assert same_iters(st.iterations - st.extra_iterations, itr) and rel(x1, xr) <= REL_DIFF
    S.close()


def test_concurrent_solves_from_two_threads(hip):
    """Two threads solving on two systems at once (their graph captures take
    the process-wide capture lock) and two threads multiplying with one
    OpenNL CRS matrix (its device mirror is built once, under a lock): every
    result is bitwise the one computed from a single thread."""
    import threading
    L = hip
    A = W.lscm_grid(40, seed=1)
    Bw = W.lscm_grid(50, seed=2)
    SA, SB = capi.System.from_crs(A), capi.System.from_crs(Bw)
    xa0, _ = SA.solve(A.b)
    xb0, _ = SB.solve(Bw.b)
    SA.close()
    SB.close()
    SA, SB = capi.System.from_crs(A), capi.System.from_crs(Bw)   # fresh graphs
    res = {}

    def solve(name, S, b):
        for k in range(3):
            res[(name, k)] = S.solve(b)[0]
    ts = [threading.Thread(target=solve, args=("a", SA, A.b)),
          threading.Thread(target=solve, args=("b", SB, Bw.b))]
    [t.start() for t in ts]
    [t.join() for t in ts]
    for k in range(3):
        assert np.array_equal(res[("a", k)], xa0) and np.array_equal(res[("b", k)], xb0)
    SA.close()
    SB.close()
    M = L.nlCRSMatrixNewFromArrays(A.n, A.n, capi.ptr(A.rowptr), capi.ptr(A.colind),
                                   capi.ptr(A.val))
    x = np.random.default_rng(3).standard_normal(A.n)
    ref = O.spmv(O.CRS(A.rowptr, A.colind, A.val), x)
    ys = [np.empty(A.n), np.empty(A.n)]
    ok = []

    def mult(k):
        for _ in range(5):
            L.nlMultMatrixVector(M, capi.ptr(x), capi.ptr(ys[k]))
            ok.append(np.array_equal(ys[k], ref))
    ts = [threading.Thread(target=mult, args=(k,)) for k in range(2)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    L.nlDeleteMatrix(M)
    assert len(ok) == 10 and all(ok)
