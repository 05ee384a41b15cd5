"""Synthetic LSCM / cotangent-Laplacian systems (ctypes over libnlworkloads.so).

These are the systems geogram's ``mesh_compute_LSCM`` and ``mesh_smooth``
build before calling the solver (call sites at
src/lib/OGF/mesh/commands/mesh_grob_surface_commands.cpp:703,724,759), rebuilt
directly in CRS form from seeded meshes.  See ``csrc/workloads.cpp``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

from . import _native

_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = _native.load("libnlworkloads.so")
        sig = {
            "wl_splitmix64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
            "wl_uniform01": (C.c_double, [C.POINTER(C.c_uint64)]),
            "wl_mesh_grid": (C.c_void_p, [C.c_uint32, C.c_uint64, C.c_int]),
            "wl_mesh_chart": (C.c_void_p, [C.c_uint32, C.c_uint64]),
            "wl_mesh_nv": (C.c_uint32, [C.c_void_p]),
            "wl_mesh_nt": (C.c_uint32, [C.c_void_p]),
            "wl_mesh_points": (_f64p, [C.c_void_p]),
            "wl_mesh_triangles": (_u32p, [C.c_void_p]),
            "wl_mesh_free": (None, [C.c_void_p]),
            "wl_mesh_extreme_pins": (None, [C.c_void_p, _u32p, _u32p]),
            "wl_lscm_rows": (None, [C.c_void_p, _f64p, C.POINTER(C.c_uint8)]),
            "wl_system_lscm": (C.c_void_p, [C.c_void_p, C.c_uint32, C.c_double,
                                            C.c_double, C.c_uint32, C.c_double,
                                            C.c_double]),
            "wl_system_cotan": (C.c_void_p, [C.c_void_p, C.c_double, C.c_uint32,
                                             C.c_int, C.c_uint64]),
            "wl_system_laplacian": (C.c_void_p, [C.c_void_p, _f64p]),
            "wl_sys_n": (C.c_uint32, [C.c_void_p]),
            "wl_sys_nnz": (C.c_uint64, [C.c_void_p]),
            "wl_sys_nrhs": (C.c_uint32, [C.c_void_p]),
            "wl_sys_rowptr": (_u32p, [C.c_void_p]),
            "wl_sys_colind": (_u32p, [C.c_void_p]),
            "wl_sys_val": (_f64p, [C.c_void_p]),
            "wl_sys_rhs": (_f64p, [C.c_void_p]),
            "wl_sys_var_of": (_u32p, [C.c_void_p]),
            "wl_sys_free": (None, [C.c_void_p]),
            "wl_atlas_chart": (None, [C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                                      _u32p, C.POINTER(C.c_uint64)]),
            "wl_atlas": (C.c_void_p, [C.c_uint32, C.c_uint64, C.c_double, C.c_double]),
            "wl_atlas_count": (C.c_uint32, [C.c_void_p]),
            "wl_atlas_nnz": (C.c_uint64, [C.c_void_p]),
            "wl_atlas_row_off": (_u32p, [C.c_void_p]),
            "wl_atlas_rowptr": (_u32p, [C.c_void_p]),
            "wl_atlas_colind": (_u32p, [C.c_void_p]),
            "wl_atlas_val": (_f64p, [C.c_void_p]),
            "wl_atlas_rhs": (_f64p, [C.c_void_p]),
            "wl_atlas_chart_nv": (_u32p, [C.c_void_p]),
            "wl_atlas_free": (None, [C.c_void_p]),
            "wl_atlas_mesh": (C.c_void_p, [C.c_uint32, C.c_uint64, C.c_double, C.c_double]),
            "wl_atlas_mesh_nv": (C.c_uint32, [C.c_void_p]),
            "wl_atlas_mesh_nt": (C.c_uint32, [C.c_void_p]),
            "wl_atlas_mesh_vertex_off": (_u32p, [C.c_void_p]),
            "wl_atlas_mesh_tri_off": (_u32p, [C.c_void_p]),
            "wl_atlas_mesh_points": (_f64p, [C.c_void_p]),
            "wl_atlas_mesh_triangles": (_u32p, [C.c_void_p]),
            "wl_atlas_mesh_pins": (_u32p, [C.c_void_p]),
            "wl_atlas_mesh_pin_uv": (_f64p, [C.c_void_p]),
            "wl_atlas_mesh_free": (None, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    if n == 0:
        return np.empty(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


@dataclass
class Mesh:
    points: np.ndarray     # (nv, 3) float64
    triangles: np.ndarray  # (nt, 3) uint32
    _h: int = 0

    @property
    def nv(self) -> int:
        return self.points.shape[0]

    @property
    def nt(self) -> int:
        return self.triangles.shape[0]

    def __del__(self):
        if self._h:
            lib().wl_mesh_free(self._h)
            self._h = 0

    def extreme_pins(self):
        a, b = C.c_uint32(), C.c_uint32()
        lib().wl_mesh_extreme_pins(self._h, C.byref(a), C.byref(b))
        return int(a.value), int(b.value)

    def lscm_rows(self):
        """(nt, 12) coefficients: real row then imaginary row, each ordered
        (u0, v0, u1, v1, u2, v2); and an (nt,) validity mask."""
        c = np.zeros((self.nt, 12), dtype=np.float64)
        ok = np.zeros(self.nt, dtype=np.uint8)
        lib().wl_lscm_rows(self._h, c.ctypes.data_as(_f64p),
                           ok.ctypes.data_as(C.POINTER(C.c_uint8)))
        return c, ok.astype(bool)


def _mesh_from(h) -> Mesh:
    L = lib()
    nv, nt = L.wl_mesh_nv(h), L.wl_mesh_nt(h)
    P = _arr(L.wl_mesh_points(h), 3 * nv, np.float64).reshape(nv, 3)
    T = _arr(L.wl_mesh_triangles(h), 3 * nt, np.uint32).reshape(nt, 3)
    return Mesh(P, T, h)


def grid_mesh(k: int, seed: int = 1, flat: bool = False) -> Mesh:
    """k x k height-field grid over [0,1]^2, two triangles per quad,
    z = sum of 4 sinusoids with amplitudes U(0,0.1), frequencies U(1,8)."""
    return _mesh_from(lib().wl_mesh_grid(k, seed, 1 if flat else 0))


def chart_mesh(N: int, seed: int) -> Mesh:
    """Disk-like patch of about N vertices cut from a noisy grid."""
    return _mesh_from(lib().wl_mesh_chart(N, seed))


@dataclass
class SystemCRS:
    """A square system in CRS form: rowptr/colind (uint32), val (float64),
    rhs (nrhs, n) float64."""
    rowptr: np.ndarray
    colind: np.ndarray
    val: np.ndarray
    rhs: np.ndarray
    var_of: np.ndarray | None = None

    @property
    def n(self) -> int:
        return len(self.rowptr) - 1

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    @property
    def b(self) -> np.ndarray:
        return self.rhs[0]

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_array((self.val, self.colind.astype(np.int64),
                             self.rowptr.astype(np.int64)), shape=(self.n, self.n))


def _system_from(h, with_var_of: bool) -> SystemCRS:
    L = lib()
    if not h:
        raise ValueError("workload system construction failed")
    try:
        n, nnz, nrhs = L.wl_sys_n(h), int(L.wl_sys_nnz(h)), L.wl_sys_nrhs(h)
        rp = _arr(L.wl_sys_rowptr(h), n + 1, np.uint32)
        ci = _arr(L.wl_sys_colind(h), nnz, np.uint32)
        va = _arr(L.wl_sys_val(h), nnz, np.float64)
        rhs = _arr(L.wl_sys_rhs(h), n * nrhs, np.float64).reshape(nrhs, n)
        vo = _arr(L.wl_sys_var_of(h), n, np.uint32) if with_var_of else None
    finally:
        L.wl_sys_free(h)
    return SystemCRS(rp, ci, va, rhs, vo)


def lscm_system(mesh: Mesh, pins=None, pin_uv=None) -> SystemCRS:
    """LSCM normal equations of `mesh` with two pinned vertices.  Defaults:
    the extreme vertices in x+y, pinned at their own (x, y)."""
    if pins is None:
        pins = mesh.extreme_pins()
    p0, p1 = pins
    if pin_uv is None:
        pin_uv = (mesh.points[p0, 0], mesh.points[p0, 1],
                  mesh.points[p1, 0], mesh.points[p1, 1])
    h = lib().wl_system_lscm(mesh._h, p0, pin_uv[0], pin_uv[1], p1, pin_uv[2],
                             pin_uv[3])
    return _system_from(h, True)


def cotan_system(mesh: Mesh, t: float = 1e-3, nrhs: int = 3, rhs: str = "xyz",
                 seed: int = 0) -> SystemCRS:
    """(I + t L) with L the cotangent Laplacian; rhs 'xyz' or 'gaussian'."""
    kind = 0 if rhs == "xyz" else 1
    return _system_from(lib().wl_system_cotan(mesh._h, t, nrhs, kind, seed), False)


def laplacian(mesh: Mesh):
    """(L, mass): the cotangent stiffness matrix (symmetric, rows sum to 0)
    and the lumped mass of each vertex, for L x = lambda M x."""
    mass = np.zeros(mesh.nv, dtype=np.float64)
    S = _system_from(lib().wl_system_laplacian(mesh._h, mass.ctypes.data_as(_f64p)), False)
    return S, mass


def lscm_grid(k: int, seed: int = 1, flat: bool = False) -> SystemCRS:
    """LSCM system of a k x k grid with corners (0,0) and (1,1) pinned."""
    m = grid_mesh(k, seed, flat)
    return lscm_system(m, pins=(0, k * k - 1), pin_uv=(0.0, 0.0, 1.0, 1.0))


def atlas_chart_params(seed: int, idx: int, nmin: float = 50, nmax: float = 50000):
    N, s = C.c_uint32(), C.c_uint64()
    lib().wl_atlas_chart(seed, idx, nmin, nmax, C.byref(N), C.byref(s))
    return int(N.value), int(s.value)


@dataclass
class Atlas:
    """Charts packed block-diagonally: chart c owns rows [row_off[c], row_off[c+1])
    of the global CRS (global column indices)."""
    row_off: np.ndarray
    rowptr: np.ndarray
    colind: np.ndarray
    val: np.ndarray
    b: np.ndarray
    chart_nv: np.ndarray

    @property
    def count(self) -> int:
        return len(self.row_off) - 1

    @property
    def n(self) -> int:
        return int(self.row_off[-1])

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def chart(self, c: int) -> SystemCRS:
        """Chart c as a standalone system (local indices)."""
        r0, r1 = int(self.row_off[c]), int(self.row_off[c + 1])
        z0, z1 = int(self.rowptr[r0]), int(self.rowptr[r1])
        rp = (self.rowptr[r0:r1 + 1].astype(np.int64) - z0).astype(np.uint32)
        ci = (self.colind[z0:z1].astype(np.int64) - r0).astype(np.uint32)
        return SystemCRS(rp, ci, self.val[z0:z1].copy(), self.b[None, r0:r1].copy())


def atlas(count: int, seed: int = 1, nmin: float = 50, nmax: float = 50000) -> Atlas:
    """`count` disk-like LSCM charts, N ~ logU[nmin, nmax] vertices each."""
    L = lib()
    h = L.wl_atlas(count, seed, nmin, nmax)
    if not h:
        raise ValueError("atlas construction failed")
    try:
        m = L.wl_atlas_count(h)
        row_off = _arr(L.wl_atlas_row_off(h), m + 1, np.uint32)
        n = int(row_off[-1])
        nnz = int(L.wl_atlas_nnz(h))
        out = Atlas(row_off, _arr(L.wl_atlas_rowptr(h), n + 1, np.uint32),
                    _arr(L.wl_atlas_colind(h), nnz, np.uint32),
                    _arr(L.wl_atlas_val(h), nnz, np.float64),
                    _arr(L.wl_atlas_rhs(h), n, np.float64),
                    _arr(L.wl_atlas_chart_nv(h), m, np.uint32))
    finally:
        L.wl_atlas_free(h)
    return out


@dataclass
class AtlasMesh:
    """The meshes behind `atlas(count, seed, ...)`, concatenated: chart c owns
    vertices [vertex_off[c], vertex_off[c+1]) and triangles
    [tri_off[c], tri_off[c+1]) (global vertex ids); pins (count, 2) global
    vertex ids with pin_uv (count, 4) = (u0, v0, u1, v1)."""
    vertex_off: np.ndarray
    tri_off: np.ndarray
    points: np.ndarray     # (nv, 3)
    triangles: np.ndarray  # (nt, 3)
    pins: np.ndarray
    pin_uv: np.ndarray

    @property
    def count(self) -> int:
        return len(self.vertex_off) - 1


def atlas_mesh(count: int, seed: int = 1, nmin: float = 50, nmax: float = 50000) -> AtlasMesh:
    L = lib()
    h = L.wl_atlas_mesh(count, seed, nmin, nmax)
    if not h:
        raise ValueError("atlas construction failed")
    try:
        nv, nt = L.wl_atlas_mesh_nv(h), L.wl_atlas_mesh_nt(h)
        out = AtlasMesh(_arr(L.wl_atlas_mesh_vertex_off(h), count + 1, np.uint32),
                        _arr(L.wl_atlas_mesh_tri_off(h), count + 1, np.uint32),
                        _arr(L.wl_atlas_mesh_points(h), 3 * nv, np.float64).reshape(nv, 3),
                        _arr(L.wl_atlas_mesh_triangles(h), 3 * nt, np.uint32).reshape(nt, 3),
                        _arr(L.wl_atlas_mesh_pins(h), 2 * count, np.uint32).reshape(count, 2),
                        _arr(L.wl_atlas_mesh_pin_uv(h), 4 * count, np.float64).reshape(count, 4))
    finally:
        L.wl_atlas_mesh_free(h)
    return out


def lscm_grid_size(k: int):
    """(n, nnz) of the k x k grid LSCM system with two corner pins."""
    V, E = k * k, (k - 1) * (3 * k - 1)
    return 2 * V - 4, 4 * (2 * E + V) - 56


def loguniform_mean(a: float, b: float) -> float:
    return (b - a) / math.log(b / a)
