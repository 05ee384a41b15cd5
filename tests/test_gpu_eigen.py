"""Generalized symmetric eigenpairs on the device (nlhipEigenSolve,
Chebyshev-filtered subspace iteration), the solve behind spectral LSCM and
manifold harmonics: eigenvalues of the path graph Laplacian against their
closed form, mesh Laplacians with a lumped mass against scipy's
shift-invert Lanczos (ARPACK with a sparse LU, the configuration the
reference uses through its ARPACK extension), residuals, M-orthonormality,
run-to-run bits, and input errors."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from graphitethree_amd import capi
from graphitethree_amd import workloads as W

pytestmark = pytest.mark.gpu


def path_laplacian(n):
    main = np.full(n, 2.0)
    main[0] = main[-1] = 1.0
    A = sp.diags([-np.ones(n - 1), main, -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    A.sort_indices()
    return A


def system_of(A):
    A = A.tocsr()
    A.sort_indices()
    return capi.System(A.indptr.astype(np.uint32), A.indices.astype(np.uint32),
                       A.data.astype(np.float64))


def test_path_laplacian_closed_form(hip):
    n, k = 300, 6
    S = system_of(path_laplacian(n))
    vals, vecs, st = S.eigen(k, shift=-0.01, eps=1e-10)
    exact = 2.0 - 2.0 * np.cos(np.pi * np.arange(k) / n)
    assert st.converged == k
    np.testing.assert_allclose(vals, exact, rtol=1e-9, atol=1e-12)
    # orthonormal, and the first is the constant vector
    np.testing.assert_allclose(vecs @ vecs.T, np.eye(k), atol=1e-9)
    assert abs(abs(vecs[0] @ np.ones(n)) / np.sqrt(n) - 1.0) < 1e-9
    S.close()


@pytest.mark.parametrize("flat,k", [(False, 5), (False, 12), (True, 3)])
def test_mesh_laplacian_with_mass_matches_arpack(hip, flat, k):
    mesh = W.grid_mesh(45, seed=5, flat=flat)
    L, mass = W.laplacian(mesh)
    A = L.to_scipy()
    M = sp.diags(mass)
    # the shift sets the scale of the residual test for the zero eigenvalue,
    # whose residual cannot go below ~1e-16 lambda_max / |shift|
    sigma = -1e-2
    S = capi.System.from_crs(L)
    vals, vecs, st = S.eigen(k, mass=mass, shift=sigma, eps=1e-9)
    ref = np.sort(spla.eigsh(A, k=k, M=M, sigma=sigma, which="LM", tol=1e-13)[0])
    assert st.converged == k
    np.testing.assert_allclose(vals, ref, rtol=1e-8, atol=1e-10)
    # pairs: A x = lambda M x, M-orthonormal
    bound = float(np.max(abs(A).sum(axis=1) / mass))     # Gershgorin, M^-1 A
    for j in range(k):
        r = A @ vecs[j] - vals[j] * (mass * vecs[j])
        mx = np.linalg.norm(mass * vecs[j])
        scale = max(abs(vals[j]), abs(sigma)) * mx
        # eps-relative, or at the operator's rounding level (the zero mode)
        assert np.linalg.norm(r) <= max(1e-8 * scale, 1e3 * 2.2e-16 * bound * mx)
    np.testing.assert_allclose(vecs @ (mass[None, :] * vecs).T, np.eye(k), atol=1e-8)
    S.close()


def test_same_bits_every_run(hip):
    mesh = W.grid_mesh(30, seed=2)
    L, mass = W.laplacian(mesh)
    S = capi.System.from_crs(L)
    v1, x1, _ = S.eigen(4, mass=mass, shift=-1e-3, eps=1e-9)
    v2, x2, _ = S.eigen(4, mass=mass, shift=-1e-3, eps=1e-9)
    assert np.array_equal(v1, v2) and np.array_equal(x1, x2)
    S.close()


def test_rejects_bad_input(hip):
    S = system_of(path_laplacian(50))
    with pytest.raises(capi.NLHIPError):
        S.eigen(0)
    with pytest.raises(capi.NLHIPError):
        S.eigen(25)
    with pytest.raises(capi.NLHIPError):
        S.eigen(3, mass=np.zeros(50))
    S.close()
    T = system_of(path_laplacian(8))          # smaller than the vector block
    with pytest.raises(capi.NLHIPError):
        T.eigen(2, shift=-1.0)
    T.close()
