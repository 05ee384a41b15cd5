"""Compile-level drop-in: tests/dropin/call_shapes.cpp makes the calls of
Graphite's NL wrapper (src/lib/OGF/scene_graph/NL/blas.cpp:65-148,
matrix.cpp:121-398) and of WarpDrive's voxel interpolation
(plugins/OGF/WarpDrive/commands/voxel_grob_transport_commands.cpp:231-287)
in their own call shapes; it must compile against include/nl.h and link to
libnlhip.so unchanged.  Without a GPU only the matrix building runs; with
one, the products, BLAS-1, the CG solve and the context solve are checked
against closed forms and the oracle."""
import os
import subprocess

import numpy as np
import pytest

from graphitethree_amd import _native
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "dropin", "call_shapes.cpp")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    lib = _native.path("libnlhip.so")
    if not os.path.exists(lib):
        pytest.fail("libnlhip.so is not built")
    out = str(tmp_path_factory.mktemp("dropin") / "call_shapes")
    libdir = os.path.dirname(lib)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), SRC, "-L" + libdir, "-lnlhip",
                        "-Wl,-rpath," + libdir, "-o", out],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return out


def run(exe, mode):
    r = subprocess.run([exe, mode], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        key, *vals = line.split()
        out[key] = vals
    return out


def floats(vals):
    return np.array([float.fromhex(v) for v in vals])


def test_dropin_compiles_and_builds_matrices(exe):
    out = run(exe, "host")
    assert out["shape"] == ["2", "2", "Dynamic"]
    assert out["add_out_of_bounds"] == ["0"]
    assert out["factorize_null"] == ["1"]          # no direct solver in this library
    assert out["compressed_format"] == ["CRS"]
    D = O.Dyn(3, 3)
    for i, j, a in ((2, 1, 4.0), (0, 2, -1.0), (0, 0, 2.0), (2, 1, 0.25), (1, 1, 5.0)):
        D.add(i, j, a)
    ref = D.compress()
    assert [int(v) for v in out["crs_rowptr"]] == list(ref.rowptr)
    assert [int(v) for v in out["crs_colind"]] == list(ref.colind)
    assert np.array_equal(floats(out["crs_val"]), ref.val)


@pytest.mark.gpu
def test_dropin_solves_on_the_device(exe):
    out = run(exe, "gpu")
    # D x with x = (1, 2, 3): rows (2, -1 | 5 | 4.25): exact
    assert np.array_equal(floats(out["mult"]), [-1.0, 10.0, 8.5])
    # linear_algebra.lua: [[1,2],[2,3]] x = [5,6] -> [-3, 4] (CG, Jacobi)
    assert int(out["cg_iterations"][0]) <= 3
    np.testing.assert_allclose(floats(out["cg_x"]), [-3.0, 4.0], rtol=0, atol=1e-12)
    u = np.array([1.0, 2.0, 3.0, 4.0, 0.5])
    v = np.array([0.25, -1.0, 2.0, 8.0, 16.0])
    assert np.array_equal(floats(out["blas_copy"]), u)
    assert np.array_equal(floats(out["blas_scal"]), 2 * u)
    assert floats(out["blas_dot"])[0] == 44.25
    assert floats(out["blas_nrm2"])[0] == 13.0
    assert np.array_equal(floats(out["blas_axpy"]), v - 0.5 * u)
    # the voxel context against the oracle's context on the same rows
    nu, nv, nw = 5, 4, 3
    N = nu * nv * nw
    idx = lambda a, b, c: (c * nv + b) * nu + a
    ctx = O.Context(N, least_squares=True)
    for i in range(N):
        ctx.set_variable(i, ((i * 37) % 23) / 8.0)
        if i % 3 == 0:
            ctx.lock_variable(i)
    ctx.begin_matrix()
    for w in range(nw):
        for b in range(nv):
            for a in range(nu):
                ctx.row([(idx(a, b, w), -6.0), (idx((a - 1) % nu, b, w), 1.0),
                         (idx((a + 1) % nu, b, w), 1.0), (idx(a, (b - 1) % nv, w), 1.0),
                         (idx(a, (b + 1) % nv, w), 1.0), (idx(a, b, (w - 1) % nw), 1.0),
                         (idx(a, b, (w + 1) % nw), 1.0)])
    ctx.solve(eps=1e-12, max_iter=10000)
    cpu = np.array([ctx.get_variable(i) for i in range(N)])
    gpu = floats(out["context_x"])
    assert out["context_ok"] == ["1"]
    assert np.linalg.norm(gpu - cpu) <= 1e-8 * np.linalg.norm(cpu)
    locked = np.arange(N) % 3 == 0
    assert np.array_equal(gpu[locked], cpu[locked])
