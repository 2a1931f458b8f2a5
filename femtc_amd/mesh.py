"""1-D tables, structured dofmaps and affine geometry (host tables, device
meshes)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .operators import _dtype_code, _ptr, _stream


def gll_points(degree: int) -> np.ndarray:
    x = np.zeros(degree + 1)
    _lib.check(_lib.lib().femtc_gll_points(
        degree, x.ctypes.data_as(_lib._dp)), "gll_points")
    return x


def gauss_rule(nq: int):
    x, w = np.zeros(nq), np.zeros(nq)
    _lib.check(_lib.lib().femtc_gauss_rule(
        nq, x.ctypes.data_as(_lib._dp), w.ctypes.data_as(_lib._dp)),
        "gauss_rule")
    return x, w


def basis_tables(degree: int, nq: int) -> np.ndarray:
    """[phi | dphi], each [N][nq], GLL Lagrange basis at Gauss points."""
    t = np.zeros(2 * (degree + 1) * nq)
    _lib.check(_lib.lib().femtc_basis_tables(
        degree, nq, t.ctypes.data_as(_lib._dp)), "basis_tables")
    return t


def dofmap_structured(nx: int, ny: int, nz: int, degree: int, out,
                      stream=None) -> None:
    """Fill the int32 device tensor ``out`` [nx*ny*nz*(P+1)^3]."""
    _lib.check(_lib.lib().femtc_dofmap_structured(
        nx, ny, nz, degree, _ptr(out), _stream(stream)), "dofmap_structured")


def geometry_affine(J, nq: int, G=None, detJw=None, stream=None) -> None:
    """Per-cell Jacobians J (FP64 device tensor [C*9]) -> G / detJw."""
    ref = G if G is not None else detJw
    _lib.check(_lib.lib().femtc_geometry_affine(
        J.numel() // 9, nq, _ptr(J), _ptr(G), _ptr(detJw), _dtype_code(ref),
        _stream(stream)), "geometry_affine")


def even_odd_tables(degree: int, nq: int, tables, weights=None):
    """(the even-odd table set of [phi | dphi] and the weights, whether the
    tables pass the symmetry check): what plans and solvers build."""
    import ctypes as C
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights,
                                                          dtype=np.float64)
    size, sym = C.c_int64(), C.c_int32()
    wp = None if w is None else w.ctypes.data_as(_lib._dp)
    _lib.check(_lib.lib().femtc_even_odd_tables(
        degree, nq, tab.ctypes.data_as(_lib._dp), wp, None, C.byref(size),
        C.byref(sym)), "even_odd_tables")
    out = np.zeros(size.value)
    _lib.check(_lib.lib().femtc_even_odd_tables(
        degree, nq, tab.ctypes.data_as(_lib._dp), wp,
        out.ctypes.data_as(_lib._dp), C.byref(size), C.byref(sym)),
        "even_odd_tables")
    return out, bool(sym.value)


def basis_tables_weights(degree: int, nq: int) -> np.ndarray:
    """[phi | dphi | w]: the tables followed by the nq Gauss weights (the
    table layout of the affine-geometry operators)."""
    t = np.zeros(2 * (degree + 1) * nq + nq)
    _lib.check(_lib.lib().femtc_basis_tables_weights(
        degree, nq, t.ctypes.data_as(_lib._dp)), "basis_tables_weights")
    return t


def geometry_cell(J, cellgeo, stream=None) -> None:
    """Per-cell Jacobians J (FP64 device tensor [C*9]) -> cellgeo [C*7]:
    |det J| J^-1 J^-T packed as G, then |det J| (dtype of cellgeo)."""
    _lib.check(_lib.lib().femtc_geometry_cell(
        J.numel() // 9, _ptr(J), _ptr(cellgeo), _dtype_code(cellgeo),
        _stream(stream)), "geometry_cell")
