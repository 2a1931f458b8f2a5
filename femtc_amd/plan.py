"""Scatter plans: the continuous operator with LDS-assembled scatter over
3-D blocks of a structured cell numbering (libfemtc femtc_plan_*)."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .operators import _dtype_code, _ptr, _stream


class Plan:
    """Plan for a dofmap whose cells are numbered (ci*ny + cj)*nz + ck.

    The dofmap tensor must stay alive while the plan is used.  Results are
    correct for any dofmap; the blocked scatter is fast for the structured
    one."""

    def __init__(self, nx: int, ny: int, nz: int, degree: int, dofmap,
                 ndofs: int, stream=None):
        self._dofmap = dofmap
        self.degree = degree
        self.ndofs = ndofs
        h = C.c_void_p()
        _lib.check(_lib.lib().femtc_plan_create(
            C.byref(h), nx, ny, nz, degree, _ptr(dofmap), ndofs,
            _stream(stream)), "plan_create")
        self._h = h

    def set_block(self, nq: int, dtype: int, op: int, shape) -> None:
        cbx, cby, cbz = shape
        _lib.check(_lib.lib().femtc_plan_set_block(self._h, nq, dtype, op, cbx,
                                                   cby, cbz), "plan_set_block")

    def get_block(self, nq: int, dtype: int, op: int):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _lib.check(_lib.lib().femtc_plan_get_block(
            self._h, nq, dtype, op, C.byref(a), C.byref(b), C.byref(c)),
            "plan_get_block")
        return (a.value, b.value, c.value)

    def set_tables(self, tables, nq: int, stream=None) -> bool:
        """Register a copy of `tables` (its dtype) for applies passing
        tables=None; True when they contract in even-odd form."""
        eo = C.c_int32()
        _lib.check(_lib.lib().femtc_plan_set_tables(
            self._h, _dtype_code(tables), nq, _ptr(tables), _stream(stream),
            C.byref(eo)), "plan_set_tables")
        return bool(eo.value)

    def apply(self, op: int, tables, x, y, nq: int, G=None, detJw=None,
              accumulate: bool = False, stream=None) -> None:
        _lib.check(_lib.lib().femtc_plan_apply(
            self._h, op, _dtype_code(x), nq, _ptr(tables), _ptr(x), _ptr(y),
            _ptr(G), _ptr(detJw), int(accumulate), _stream(stream)),
            "plan_apply")

    def autotune(self, op: int, tables, x, y, nq: int, G=None, detJw=None,
                 reps: int = 5, stream=None):
        """Pick the fastest block shape on these buffers; returns
        (shape, microseconds per apply).  y is left cleared."""
        best = C.c_double()
        _lib.check(_lib.lib().femtc_plan_autotune(
            self._h, op, _dtype_code(x), nq, _ptr(tables), _ptr(x), _ptr(y),
            _ptr(G), _ptr(detJw), reps, _stream(stream), C.byref(best)),
            "plan_autotune")
        return self.get_block(nq, _dtype_code(x), op), best.value

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().femtc_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
