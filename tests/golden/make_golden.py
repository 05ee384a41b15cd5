"""Writes the committed fixtures in tests/golden/ (run from the repo root).

Every expected value comes from a closed form, exact integer arithmetic, a
library routine (numpy / scipy) or the reference's own example script; none
comes from the code under test.

* lua_kats.json     the 2x2 systems of lib/programming_examples/Lua/
                    linear_algebra.lua:7-68 with their exact solutions.
* spmv_exact.npz    CRS matrices with small-integer values (empty rows, a
                    row longer than 64, duplicates merged) and integer x:
                    y = A x is exact in FP64 whatever the summation order.
* blas_exact.json   dyadic vectors whose dot, axpy, scal results are exact.
* lscm_k12.npz      the seeded k=12 LSCM system and its scipy.sparse direct
                    solution (SuperLU), a forward-error reference.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def lua_kats():
    return {
        "source": "lib/programming_examples/Lua/linear_algebra.lua",
        "general": {"M": [[1.0, 2.0], [3.0, 4.0]], "B": [5.0, 6.0], "X": [-4.0, 4.5],
                    "lines": "7-16,41-44", "solver": "BiCGSTAB", "precond": "None",
                    "max_iter": 1000, "eps": 0.001},
        "symmetric": {"M": [[1.0, 2.0], [2.0, 3.0]], "B": [5.0, 6.0], "X": [-3.0, 4.0],
                      "lines": "49-61", "solver": "CG", "precond": "Jacobi"},
    }


def random_crs(rng, m, n, density, long_row=None, empty_rows=()):
    rows, cols, vals = [], [], []
    for i in range(m):
        if i in empty_rows:
            continue
        k = rng.binomial(n, density)
        if long_row is not None and i == long_row:
            k = min(n, 100)
        c = np.sort(rng.choice(n, size=k, replace=False))
        rows += [i] * len(c)
        cols += list(c)
        vals += list(rng.integers(-8, 9, size=len(c)).astype(np.float64))
    import scipy.sparse as sp
    A = sp.csr_array((vals, (rows, cols)), shape=(m, n))
    A.sort_indices()
    return A


def spmv_exact():
    rng = np.random.default_rng(2024)
    out = {}
    cases = [("tiny", 1, 1, 1.0, None, ()), ("small", 37, 37, 0.2, 5, (0, 7, 36)),
             ("medium", 1000, 1000, 0.014, 17, (3, 500)), ("wide", 300, 300, 0.3, None, ())]
    for name, m, n, d, lr, er in cases:
        A = random_crs(rng, m, n, d, lr, er)
        x = rng.integers(-16, 17, size=n).astype(np.float64)
        y = A @ x
        out[f"{name}_rowptr"] = A.indptr.astype(np.uint32)
        out[f"{name}_colind"] = A.indices.astype(np.uint32)
        out[f"{name}_val"] = A.data.astype(np.float64)
        out[f"{name}_x"] = x
        out[f"{name}_y"] = y
    out["names"] = np.array([c[0] for c in cases])
    return out


def blas_exact():
    rng = np.random.default_rng(7)
    n = 4096
    x = (rng.integers(-1024, 1025, size=n) / 64.0)
    y = (rng.integers(-1024, 1025, size=n) / 32.0)
    a = 0.75
    return {"x": x.tolist(), "y": y.tolist(), "a": a,
            "dot": float(np.dot(x.astype(object), y.astype(object))),
            "axpy": (y + a * x).tolist(), "scal": (a * x).tolist(),
            "nrm2_sq": float(np.dot(x.astype(object), x.astype(object)))}


def lscm_k12():
    import scipy.sparse.linalg as spla
    from graphitethree_amd import workloads as W
    S = W.lscm_grid(12, seed=1)
    x = spla.spsolve(S.to_scipy().tocsc(), S.b)
    return {"rowptr": S.rowptr, "colind": S.colind, "val": S.val, "b": S.b, "x": x}


if __name__ == "__main__":
    with open(os.path.join(HERE, "lua_kats.json"), "w") as f:
        json.dump(lua_kats(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "spmv_exact.npz"), **spmv_exact())
    with open(os.path.join(HERE, "blas_exact.json"), "w") as f:
        json.dump(blas_exact(), f)
    np.savez_compressed(os.path.join(HERE, "lscm_k12.npz"), **lscm_k12())
    print("fixtures written to", HERE)
