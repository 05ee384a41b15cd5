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
