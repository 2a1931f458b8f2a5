"""Operator entry points mirroring the reference's ``operators`` namespace.

* ``mass_operator`` / ``stiffness_operator`` keep the reference signatures
  (src/mass.hpp:196-199, src/stiffness.hpp:245-248) on host (numpy) buffers
  and run synchronously on the GPU, including the ladder's behaviour for
  unsupported arguments: print the reference's message and leave W untouched
  (src/mass.hpp:134-136, 165-167, 190-192, 215-217).
* ``mass_apply`` / ``stiffness_apply`` / ``kpm_apply`` / ``global_apply`` are
  the stream-ordered device entry points on torch CUDA tensors (or raw device
  pointers), generalised to Q != P+1.
"""
from __future__ import annotations

import contextlib

import numpy as np

from . import _lib
from .linalg import Order

_LADDER_MSG = {
    _lib.ERR_DEGREE_NOT_SUPPORTED: "degree not supported",
    _lib.ERR_ORDER_NOT_SUPPORTED: "order not supported",
    _lib.ERR_BLOCK_NOT_SUPPORTED: "Block size not supported",
}


def _dtype_code(a) -> int:
    dt = str(getattr(a, "dtype", ""))
    if dt.endswith("float64"):
        return _lib.F64
    if dt.endswith("float32"):
        return _lib.F32
    if dt.endswith("bfloat16"):
        return _lib.BF16
    raise TypeError(f"unsupported dtype {dt}")


def _ptr(a):
    """Device pointer of a torch CUDA tensor, an int address or None."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        if not a.is_cuda:
            raise ValueError("expected a device (cuda) tensor")
        if not a.is_contiguous():
            raise ValueError("expected a contiguous tensor")
        return a.data_ptr()
    raise TypeError(f"expected a device tensor, got {type(a)}")


def _stream(stream):
    if stream is None:
        try:
            import torch
            return torch.cuda.current_stream().cuda_stream
        except Exception:  # torch absent or no device: default stream
            return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _host_ptr(a: np.ndarray):
    return a.ctypes.data_as(_lib.C.c_void_p)


def _host_call(fn, op_name, phi, U, W, geom, num_cells, degree, order, Mb, Nb,
               mode):
    arrs = [np.asarray(x) for x in (phi, U, W, geom)]
    for a in arrs:
        if not a.flags.c_contiguous:
            raise ValueError("buffers must be C-contiguous")
    if not (arrs[0].dtype == arrs[1].dtype == arrs[2].dtype == arrs[3].dtype):
        raise TypeError("phi, U, W and the geometry must share one dtype")
    st = fn(_dtype_code(arrs[0]), _host_ptr(arrs[0]), arrs[0].size,
            _host_ptr(arrs[1]), _host_ptr(arrs[2]), _host_ptr(arrs[3]),
            num_cells, degree, int(order), Mb, Nb, mode)
    if st in _LADDER_MSG:
        print(_LADDER_MSG[st], flush=True)
        return
    _lib.check(st, op_name)


def mass_operator(phi, U, W, detJ, num_cells: int, degree: int,
                  order: Order, Mb: int, Nb: int,
                  mode: int = _lib.MODE_REFERENCE) -> None:
    """operators::mass_operator<T> (src/mass.hpp:196-199) on host buffers.

    W is overwritten with the operator applied to every cell of U.  Q is
    degree+1 and detJ holds (degree+1)^3 values per cell, as in the
    reference.  ``mode`` defaults to the reference's own arithmetic
    (MODE_REFERENCE); MODE_CORRECT applies the true transpose."""
    _host_call(_lib.lib().femtc_mass_operator, "mass_operator", phi, U, W,
               detJ, num_cells, degree, order, Mb, Nb, mode)


def stiffness_operator(phi, U, W, G, num_cells: int, degree: int,
                       order: Order, Mb: int, Nb: int,
                       mode: int = _lib.MODE_REFERENCE) -> None:
    """operators::stiffness_operator<T> (src/stiffness.hpp:245-248) on host
    buffers.  phi must hold [phi | dphi] = 2 (degree+1)^2 values
    (src/stiffness.hpp:44-45 reads the derivative table at offset (P+1)^2)."""
    _host_call(_lib.lib().femtc_stiffness_operator, "stiffness_operator", phi,
               U, W, G, num_cells, degree, order, Mb, Nb, mode)


def mass_apply(tables, U, W, detJw, num_cells: int, degree: int, nq: int,
               mode: int = _lib.MODE_CORRECT, stream=None) -> None:
    """W = M U per cell on device buffers (detJw [C][nq^3])."""
    _lib.check(_lib.lib().femtc_mass_apply(
        _dtype_code(U), degree, nq, mode, _ptr(tables), _ptr(U), _ptr(W),
        _ptr(detJw), num_cells, _stream(stream)), "mass_apply")


def stiffness_apply(tables, U, W, G, num_cells: int, degree: int, nq: int,
                    mode: int = _lib.MODE_CORRECT, stream=None) -> None:
    """W = K U per cell on device buffers (G [C][nq^3][6])."""
    _lib.check(_lib.lib().femtc_stiffness_apply(
        _dtype_code(U), degree, nq, mode, _ptr(tables), _ptr(U), _ptr(W),
        _ptr(G), num_cells, _stream(stream)), "stiffness_apply")


def kpm_apply(tables, U, W, G, detJw, num_cells: int, degree: int, nq: int,
              stream=None) -> None:
    """W = (K + M) U per cell on device buffers."""
    _lib.check(_lib.lib().femtc_kpm_apply(
        _dtype_code(U), degree, nq, _ptr(tables), _ptr(U), _ptr(W), _ptr(G),
        _ptr(detJw), num_cells, _stream(stream)), "kpm_apply")


def set_high_degree_kernel(mode: int) -> int:
    """Kernel choice for degrees 9..25 (femtc_set_high_degree_kernel):
    HIGH_DEGREE_AUTO runs the line kernels where they apply (correct mode,
    mass / stiffness / K+M with per-point geometry, degrees 9..25 where a
    cell fits the registers and LDS: FP64 stiffness and K+M to degree
    20..21), else the general kernel; HIGH_DEGREE_GENERAL the general kernel
    always.
    Process-wide; returns the previous mode."""
    prev = _lib.C.c_int32(0)
    _lib.check(_lib.lib().femtc_set_high_degree_kernel(mode,
                                                       _lib.C.byref(prev)),
               "set_high_degree_kernel")
    return prev.value


def apply_kernel(op: int, dtype: int, degree: int, nq: int,
                 mode: int = 0) -> str:
    """Which kernel family the apply entry points run for these arguments
    under the current setting (femtc_apply_kernel): "fused", "line" or
    "general"."""
    k = _lib.C.c_int32(-1)
    _lib.check(_lib.lib().femtc_apply_kernel(op, dtype, degree, nq, mode,
                                             _lib.C.byref(k)),
               "apply_kernel")
    return {_lib.KERNEL_FUSED: "fused", _lib.KERNEL_GENERAL: "general",
            _lib.KERNEL_LINE: "line"}[k.value]


@contextlib.contextmanager
def high_degree_kernel(mode: int):
    """set_high_degree_kernel(mode) for the duration of a with-block."""
    prev = set_high_degree_kernel(mode)
    try:
        yield
    finally:
        set_high_degree_kernel(prev)


def fill_zero(y, stream=None) -> None:
    """y = 0 with a device kernel."""
    _lib.check(_lib.lib().femtc_fill_zero(_ptr(y), y.numel(), _dtype_code(y),
                                          _stream(stream)), "fill_zero")


def global_apply(op: int, tables, x, y, dofmap, num_cells: int, degree: int,
                 nq: int, G=None, detJw=None, stream=None,
                 accumulate: bool = False) -> None:
    """y = A x through the int32 dofmap (y overwritten, or y += A x with
    ``accumulate``); op is one of OP_MASS, OP_STIFFNESS, OP_KPM."""
    fn = (_lib.lib().femtc_global_apply_add if accumulate
          else _lib.lib().femtc_global_apply)
    _lib.check(fn(
        op, _dtype_code(x), degree, nq, _ptr(tables), _ptr(x), _ptr(y),
        y.numel(), _ptr(dofmap), _ptr(G), _ptr(detJw), num_cells,
        _stream(stream)), "global_apply")


def global_apply_affine(op: int, tables_w, x, y, dofmap, num_cells: int,
                        degree: int, nq: int, cellgeo, stream=None,
                        accumulate: bool = False) -> None:
    """y = A x (or y += A x) on affine cells with the geometry computed in
    the kernel from cellgeo [C][7]; tables_w = [phi | dphi | w]."""
    _lib.check(_lib.lib().femtc_global_apply_affine(
        op, _dtype_code(x), degree, nq, _ptr(tables_w), _ptr(x), _ptr(y),
        y.numel(), _ptr(dofmap), _ptr(cellgeo), num_cells, int(accumulate),
        _stream(stream)), "global_apply_affine")
