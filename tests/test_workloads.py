"""Pins the synthetic system builders (graphitethree_amd/csrc/workloads.cpp).

* The LSCM normal equations built directly in CRS are bitwise equal to the
  oracle's row-by-row least-squares assembly of the same per-triangle rows
  (the nlBegin/nlEnd(NL_ROW) path), with the pinned vertices locked.
* Closed forms: on a flat grid pinned at its own (x, y) the LSCM solution is
  the identity map (Levy et al. 2002: similarity maps have zero conformal
  energy); the rows annihilate every similarity of the flat coordinates.
* Cotangent Laplacian: L 1 = 0, symmetry, flat-grid weights 1 on axis edges
  and 0 on the split diagonals (cot 45 = 1, cot 90 = 0).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from graphitethree_amd import workloads as W
from oracle import oracle as O


def test_grid_sizes_match_formulas():
    for k in (3, 5, 12, 31):
        S = W.lscm_grid(k, seed=3)
        assert (S.n, S.nnz) == W.lscm_grid_size(k)
        assert np.all(np.diff(S.rowptr.astype(np.int64)) > 0)


def _context_assembly(mesh, pins, uv):
    coeffs, ok = mesh.lscm_rows()
    ctx = O.Context(2 * mesh.nv, least_squares=True)
    for p, (u, v) in zip(pins, uv):
        ctx.set_variable(2 * p, u)
        ctx.set_variable(2 * p + 1, v)
        ctx.lock_variable(2 * p)
        ctx.lock_variable(2 * p + 1)
    ctx.begin_matrix()
    for t in range(mesh.nt):
        if not ok[t]:
            continue
        tri = mesh.triangles[t]
        var = [2 * int(tri[e // 2]) + (e & 1) for e in range(6)]
        ctx.row(list(zip(var, coeffs[t, :6])))
        ctx.row(list(zip(var, coeffs[t, 6:])))
    return ctx.system()


@pytest.mark.parametrize("kind", ["grid", "chart"])
def test_lscm_builder_bitwise_equals_row_assembly(kind):
    if kind == "grid":
        mesh = W.grid_mesh(7, seed=5)
        pins = (0, 48)
    else:
        mesh = W.chart_mesh(90, seed=77)
        pins = mesh.extreme_pins()
    uv = [(mesh.points[p, 0], mesh.points[p, 1]) for p in pins]
    S = W.lscm_system(mesh, pins=pins, pin_uv=(uv[0][0], uv[0][1], uv[1][0], uv[1][1]))
    M, b = _context_assembly(mesh, pins, uv)
    assert np.array_equal(S.rowptr, M.rowptr)
    assert np.array_equal(S.colind, M.colind)
    assert np.array_equal(S.val, M.val)
    assert np.array_equal(S.b, b)


def test_flat_grid_lscm_is_identity_map():
    k = 15
    S = W.lscm_grid(k, seed=2, flat=True)
    x = spla.spsolve(S.to_scipy().tocsc(), S.b)
    mesh = W.grid_mesh(k, seed=2, flat=True)
    var = S.var_of
    expected = np.where(var % 2 == 0, mesh.points[var // 2, 0], mesh.points[var // 2, 1])
    np.testing.assert_allclose(x, expected, rtol=0, atol=1e-10)


def test_lscm_rows_annihilate_similarities():
    mesh = W.grid_mesh(6, seed=4, flat=True)
    c, ok = mesh.lscm_rows()
    assert ok.all()
    z = mesh.points[:, 0] + 1j * mesh.points[:, 1]
    for a, t in [(1.0, 0.0), (0.6 - 0.8j, 0.3 + 2j)]:
        U = a * z + t
        uv = np.stack([U.real, U.imag], axis=1)
        for tt in range(mesh.nt):
            tri = mesh.triangles[tt]
            vals = uv[tri].reshape(-1)   # u0 v0 u1 v1 u2 v2
            assert abs(c[tt, :6] @ vals) < 1e-12
            assert abs(c[tt, 6:] @ vals) < 1e-12
    # a reflection is not conformal (orientation): rows do not vanish
    U = np.conj(z)
    uv = np.stack([U.real, U.imag], axis=1)
    tri = mesh.triangles[0]
    assert abs(c[0, :6] @ uv[tri].reshape(-1)) + abs(c[0, 6:] @ uv[tri].reshape(-1)) > 1e-3


def test_lscm_system_symmetric_positive_definite():
    S = W.lscm_grid(9, seed=1)
    A = S.to_scipy().toarray()
    np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-13)
    assert np.linalg.eigvalsh(A).min() > 0


def test_cotan_flat_grid_weights_and_invariants():
    k, t = 6, 1e-3
    mesh = W.grid_mesh(k, seed=1, flat=True)
    S = W.cotan_system(mesh, t=t, nrhs=3)
    A = S.to_scipy().toarray()
    L = (A - np.eye(S.n)) / t
    np.testing.assert_allclose(L.sum(axis=1), 0, atol=1e-9)
    np.testing.assert_allclose(A, A.T, atol=1e-15)
    # interior axis edge: weight 1; split diagonal: weight 0
    v = 2 * k + 2
    np.testing.assert_allclose(-L[v, v + 1], 1.0, atol=1e-9)
    np.testing.assert_allclose(-L[v, v + k], 1.0, atol=1e-9)
    np.testing.assert_allclose(L[v, v + k + 1], 0.0, atol=1e-9)
    # rhs columns are x, y, z
    np.testing.assert_array_equal(S.rhs[0], mesh.points[:, 0])
    np.testing.assert_array_equal(S.rhs[2], mesh.points[:, 2])
    # constant rhs -> constant solution (A 1 = 1)
    np.testing.assert_allclose(A @ np.ones(S.n), np.ones(S.n), atol=1e-12)


def test_cotan_noisy_sizes_and_gaussian_rhs():
    mesh = W.grid_mesh(20, seed=8)
    S = W.cotan_system(mesh, t=1e-3, nrhs=4, rhs="gaussian", seed=3)
    V, E = 400, 19 * 59
    assert S.n == V and S.nnz == V + 2 * E
    assert S.rhs.shape == (4, V)
    assert abs(S.rhs.mean()) < 0.1 and 0.9 < S.rhs.std() < 1.1


def test_chart_mesh_properties():
    for i in range(5):
        N, seed = W.atlas_chart_params(1234, i)
        assert 50 <= N <= 50000
        m = W.chart_mesh(N, seed)
        assert 0.8 * N <= m.nv <= N
        p0, p1 = m.extreme_pins()
        assert p0 != p1
        S = W.lscm_system(m)
        assert S.n == 2 * m.nv - 4


def test_generators_are_deterministic():
    a = W.lscm_grid(10, seed=42)
    b = W.lscm_grid(10, seed=42)
    c = W.lscm_grid(10, seed=43)
    assert np.array_equal(a.val, b.val) and np.array_equal(a.b, b.b)
    assert not np.array_equal(a.val, c.val)


def test_atlas_mesh_is_the_atlas_charts():
    """The concatenated chart meshes are the meshes the packed atlas is built
    from: same vertex counts, triangles inside their charts, pins at the
    extreme vertices, and chart 0's mesh rebuilt alone gives chart 0's system."""
    count, seed = 6, 9
    A = W.atlas(count, seed=seed, nmin=40, nmax=400)
    M = W.atlas_mesh(count, seed=seed, nmin=40, nmax=400)
    assert np.array_equal(np.diff(M.vertex_off), A.chart_nv)
    for c in range(count):
        v0, v1 = int(M.vertex_off[c]), int(M.vertex_off[c + 1])
        t0, t1 = int(M.tri_off[c]), int(M.tri_off[c + 1])
        tri = M.triangles[t0:t1]
        assert tri.min() >= v0 and tri.max() < v1
        N, cs = W.atlas_chart_params(seed, c, 40, 400)
        m = W.chart_mesh(N, cs)
        assert np.array_equal(m.points, M.points[v0:v1])
        assert np.array_equal(m.triangles, tri - v0)
        assert tuple(M.pins[c] - v0) == m.extreme_pins()
        assert np.array_equal(M.pin_uv[c], [m.points[M.pins[c][0] - v0, 0],
                                            m.points[M.pins[c][0] - v0, 1],
                                            m.points[M.pins[c][1] - v0, 0],
                                            m.points[M.pins[c][1] - v0, 1]])
        S = W.lscm_system(m)
        Sc = A.chart(c)
        assert np.array_equal(S.val, Sc.val) and np.array_equal(S.b[0], Sc.b[0])
