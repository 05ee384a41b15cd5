"""Solve server: budgeted admission, concurrent requests, results that are
bitwise the same alone and in the mix, isolation of a failing request."""
import numpy as np
import pytest

from graphitethree_amd import capi
from graphitethree_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def requests():
    out = []
    for i, N in enumerate((300, 1200, 2500, 700, 4000)):
        S = W.lscm_system(W.chart_mesh(N, seed=100 + i))
        out.append((S, S.b[None, :]))
    for i, g in enumerate((30, 45, 60)):
        S = W.cotan_system(W.grid_mesh(g, seed=200 + i), t=1e-3, nrhs=3, rhs="xyz")
        out.append((S, S.rhs))
    return out


def solve_alone(S, B):
    """The same request on a server of its own."""
    srv = capi.Server(max_active=1)
    X, res = srv.wait(srv.submit(S, B, eps=1e-10, max_iter=5000), timeout=120)
    srv.close()
    assert res.status == 0
    return X


def test_server_mix_equals_alone(hip):
    reqs = requests()
    srv = capi.Server(max_active=4)
    ids = [srv.submit(S, B, eps=1e-10, max_iter=5000) for S, B in reqs]
    results = [srv.wait(i, timeout=120) for i in ids]
    info = srv.info()
    srv.close()
    assert info.completed == len(reqs) and info.failed == 0
    assert info.peak_running >= 2          # requests overlapped on the device
    for (S, B), (X, res) in zip(reqs, results):
        assert res.status == 0 and res.stats.converged == 1
        assert res.t_submit <= res.t_admit <= res.t_done
        assert np.array_equal(X, solve_alone(S, B))
    for i in (1, 4):
        S, B = reqs[i]
        xr, _, _ = O.pcg(O.CRS(S.rowptr, S.colind, S.val), B[0], eps=1e-10, max_iter=5000)
        assert np.linalg.norm(results[i][0][0] - xr) / np.linalg.norm(xr) <= 1e-8
    S, B = reqs[6]  # three right-hand sides, run on its own system
    for j in range(3):
        xr, _, _ = O.pcg(O.CRS(S.rowptr, S.colind, S.val), B[j], eps=1e-10, max_iter=5000)
        assert np.linalg.norm(results[6][0][j] - xr) / np.linalg.norm(xr) <= 1e-8


def test_server_batches_small_requests(hip):
    """Small single-right-hand-side requests waiting together are solved as
    one batch: with one slot, several requests are on the device at once,
    and each result matches the oracle."""
    reqs = [W.lscm_system(W.chart_mesh(N, seed=400 + i))
            for i, N in enumerate((200, 900, 1500, 300, 2200, 600))]
    srv = capi.Server(max_active=1)
    ids = [srv.submit(S, S.b, eps=1e-10, max_iter=5000) for S in reqs]
    out = [srv.wait(i, timeout=120) for i in ids]
    info = srv.info()
    srv.close()
    assert info.completed == len(reqs) and info.failed == 0
    assert info.peak_running >= 2
    for S, (X, res) in zip(reqs, out):
        assert res.status == 0 and res.stats.converged == 1
        xr, _, _ = O.pcg(O.CRS(S.rowptr, S.colind, S.val), S.b, eps=1e-10, max_iter=5000)
        assert np.linalg.norm(X[0] - xr) / np.linalg.norm(xr) <= 1e-8


def test_server_budget_admission(hip):
    # four requests of about the same size; learn their footprint first
    same = [W.lscm_system(W.chart_mesh(2500, seed=300 + i)) for i in range(4)]
    probe = capi.Server(max_active=1)
    _, res = probe.wait(probe.submit(same[0], same[0].b), timeout=120)
    probe.close()
    fp = max(res.device_bytes, 1)
    # budget for one request at a time: all complete, never two at once
    one = capi.Server(hbm_budget_bytes=int(1.5 * fp), max_active=4)
    ids = [one.submit(S, S.b) for S in same]
    for i in ids:
        one.wait(i, timeout=120)
    info = one.info()
    one.close()
    assert info.completed == 4 and info.peak_running == 1
    assert info.peak_in_use_bytes <= info.budget_bytes
    # a request larger than the whole budget is rejected, the others run
    tiny = capi.Server(hbm_budget_bytes=fp // 2, max_active=2)
    small = W.lscm_grid(4, seed=1)
    ok_id = tiny.submit(small, small.b)
    bad_id = tiny.submit(same[1], same[1].b)
    X, res = tiny.wait(ok_id, timeout=60)
    assert res.status == 0
    with pytest.raises(capi.NLHIPError) as e:
        tiny.wait(bad_id, timeout=60)
    assert e.value.code == capi.NLHIP_ERR_BUDGET
    tiny.close()


def test_server_isolates_a_failing_request(hip):
    reqs = requests()[:5]
    srv = capi.Server(max_active=3)
    hip.nlhipInjectFailures(1)
    ids = [srv.submit(S, B) for S, B in reqs]
    codes = []
    for i in ids:
        try:
            srv.wait(i, timeout=120)
            codes.append(0)
        except capi.NLHIPError as e:
            codes.append(e.code)
    srv.close()
    assert codes.count(capi.NLHIP_ERR_INJECTED) == 1
    assert codes.count(0) == len(reqs) - 1


def test_server_resident_system_requests(hip):
    """Requests against systems kept resident on the device
    (nlhipServerSubmitSystem): only right-hand sides travel; each result is
    bitwise nlhipSolve / nlhipSolveMulti on the same system, whether the
    requests on it queue behind each other or run next to uploaded ones."""
    sm = W.cotan_system(W.grid_mesh(80, seed=11), t=1e-3, nrhs=3, rhs="xyz")
    ch = W.lscm_system(W.chart_mesh(3000, seed=12))
    Ssm = capi.System.from_crs(sm)
    Sch = capi.System.from_crs(ch)
    ref_sm, _, _ = Ssm.solve_multi(sm.rhs, eps=1e-10, max_iter=5000)
    ref_ch, _ = Sch.solve(ch.b, eps=1e-10, max_iter=50000)
    # a second pass of smoothing: the first pass's result as right-hand side
    B2 = np.ascontiguousarray(ref_sm)
    ref_sm2, _, _ = Ssm.solve_multi(B2, eps=1e-10, max_iter=5000)
    other = W.lscm_system(W.chart_mesh(1500, seed=13))
    srv = capi.Server(max_active=4)
    ids = [srv.submit_system(Ssm, sm.rhs, eps=1e-10, max_iter=5000),
           srv.submit_system(Sch, ch.b, eps=1e-10, max_iter=50000),
           srv.submit(other, other.b, eps=1e-10, max_iter=50000),
           srv.submit_system(Ssm, B2, eps=1e-10, max_iter=5000),
           srv.submit_system(Sch, ch.b, eps=1e-10, max_iter=50000)]
    out = [srv.wait(i, timeout=120) for i in ids]
    info = srv.info()
    srv.close()
    assert info.completed == 5 and info.failed == 0
    for X, res in out:
        assert res.status == 0 and res.stats.converged == 1
    assert np.array_equal(out[0][0], ref_sm) and np.array_equal(out[3][0], ref_sm2)
    assert np.array_equal(out[1][0][0], ref_ch) and np.array_equal(out[4][0][0], ref_ch)
    # resident requests are charged their workspace only, not the matrix
    assert out[1][1].device_bytes < 12 * ch.nnz
    for j in range(3):
        xr, _, _ = O.pcg(O.CRS(sm.rowptr, sm.colind, sm.val), sm.rhs[j], eps=1e-10,
                         max_iter=5000)
        assert np.linalg.norm(out[0][0][j] - xr) / np.linalg.norm(xr) <= 1e-8
    with pytest.raises(capi.NLHIPError):
        bad = capi.Server(max_active=1)
        try:
            b = sm.rhs.copy()
            b[1, 5] = np.nan
            bad.submit_system(Ssm, b)
        finally:
            bad.close()
    Ssm.close()
    Sch.close()


def test_server_requests_at_serving_scale(hip):
    """Requests of the sizes the serving workload draws: an LSCM chart of
    1e5 vertices (2e5 unknowns, ~16k iterations), a 2e4-vertex chart and a
    three-right-hand-side smoothing solve on a 300 x 300 grid, run together;
    every result reaches a true residual of 1e-10 (recomputed by the
    oracle's SpMV) and matches the oracle's solve within 1e-8."""
    big = W.lscm_system(W.chart_mesh(100_000, seed=31))
    mid = W.lscm_system(W.chart_mesh(20_000, seed=32))
    sm = W.cotan_system(W.grid_mesh(300, seed=33), t=1e-3, nrhs=3, rhs="xyz")
    srv = capi.Server(max_active=4)
    ids = [srv.submit(big, big.b, eps=1e-10, max_iter=200000),
           srv.submit(mid, mid.b, eps=1e-10, max_iter=200000),
           srv.submit(sm, sm.rhs, eps=1e-10, max_iter=5000)]
    out = [srv.wait(i, timeout=300) for i in ids]
    srv.close()
    for S, Bs, (X, res) in zip((big, mid, sm), (big.b[None, :], mid.b[None, :], sm.rhs), out):
        assert res.status == 0 and res.stats.converged == 1
        A = O.CRS(S.rowptr, S.colind, S.val)
        for j in range(Bs.shape[0]):
            r = Bs[j] - O.spmv(A, X[j], parallel=True)
            tr = np.linalg.norm(r) / np.linalg.norm(Bs[j])
            xr, it, _ = O.pcg(A, Bs[j], eps=1e-10, max_iter=200000, variant=1)
            d = np.linalg.norm(X[j] - xr) / np.linalg.norm(xr)
            print(f"n={S.n} rhs {j}: {res.stats.iterations} it (oracle {it}), true residual "
                  f"{tr:.3e}, rel diff {d:.3e}")
            assert tr <= 1e-10
            assert d <= 1e-8
