"""Operators with bfloat16 operands on the matrix cores (MFMA) and FP32
accumulation: the operator of the BF16 inner solve (libfemtc
femtc_mfma_tables / femtc_global_apply_mfma).  Affine cells of a structured
grid, nq = degree + 1 <= 8, FP32 vectors."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .operators import _ptr, _stream


def mfma_tables_bytes(split: int = 1) -> int:
    """Bytes of the operand set (the tables as MFMA operands, the weights)."""
    n = C.c_size_t()
    _lib.check(_lib.lib().femtc_mfma_tables_bytes(int(split), C.byref(n)),
               "mfma_tables_bytes")
    return n.value


def mfma_tables(degree: int, nq: int, tables, out, split: int = 1) -> None:
    """Build the operand set from the host FP64 tables [phi | dphi] into the
    device buffer `out` (mfma_tables_bytes(split) bytes).  Synchronous."""
    t = np.ascontiguousarray(np.asarray(tables, dtype=np.float64))
    if t.size < 2 * (degree + 1) * nq:
        raise ValueError("tables hold 2*(degree+1)*nq values")
    _lib.check(_lib.lib().femtc_mfma_tables(
        degree, nq, int(split), t.ctypes.data_as(C.POINTER(C.c_double)),
        _ptr(out)), "mfma_tables")


def global_apply_mfma(op: int, degree: int, nq: int, operands, x, y, grid,
                      cellgeo, lattice, split: int = 1,
                      accumulate: bool = False, stream=None) -> None:
    """y = A x (or y += A x) on the structured grid (nx, ny, nz); cellgeo
    [C][7] FP32 (femtc_geometry_cell); lattice a workspace of
    C*(degree+1)^3 floats."""
    nx, ny, nz = (int(v) for v in grid)
    _lib.check(_lib.lib().femtc_global_apply_mfma(
        op, degree, nq, int(split), _ptr(operands), _ptr(x), _ptr(y),
        y.numel(), nx, ny, nz, _ptr(cellgeo), _ptr(lattice), int(accumulate),
        _stream(stream)), "global_apply_mfma")


NOALLOC_LATTICE_VALUES = 16 << 20


def global_apply_mfma_noalloc(op: int, degree: int, nq: int, operands, x, y,
                              grid, cellgeo, split: int = 1,
                              accumulate: bool = False, stream=None) -> None:
    """global_apply_mfma without a workspace argument and without allocating:
    the lattice is the library's static 64 MiB device array (grids of at most
    NOALLOC_LATTICE_VALUES cell values); overlapping calls on one GPU race."""
    nx, ny, nz = (int(v) for v in grid)
    _lib.check(_lib.lib().femtc_global_apply_mfma_noalloc(
        op, degree, nq, int(split), _ptr(operands), _ptr(x), _ptr(y),
        y.numel(), nx, ny, nz, _ptr(cellgeo), int(accumulate),
        _stream(stream)), "global_apply_mfma_noalloc")
