"""Batched applies over many independent problems (libfemtc femtc_batch_*).

Problems sharing (op, dtype, degree, nq, cell-local/continuous) run as one
launch; launches overlap on forked streams; the sequence can be captured
into a hipGraph and replayed."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .operators import _dtype_code, _ptr, _stream


class Batch:
    """items: dicts with keys op, degree, nq, tables, x, y and, as needed,
    G, detJw, dofmap (None: cell-local), num_cells, ndofs (continuous), and
    dims (nx, ny, nz) when the dofmap is the structured numbering of that
    grid (checked; enables the assembled scatter for the item's group).
    Every tensor must stay alive and in place while the batch is used."""

    def __init__(self, items, stream=None):
        arr = (_lib.BatchItem * max(len(items), 1))()
        self._keep = []
        for i, it in enumerate(items):
            x = it["x"]
            dims = it.get("dims") if it.get("dofmap") is not None else None
            grid = (C.c_int32 * 3)(*(dims if dims is not None else (0, 0, 0)))
            arr[i] = _lib.BatchItem(
                it["op"], _dtype_code(x), it["degree"], it["nq"],
                _ptr(it["tables"]), _ptr(x), _ptr(it["y"]), _ptr(it.get("G")),
                _ptr(it.get("detJw")), _ptr(it.get("dofmap")),
                it["num_cells"], it.get("ndofs", 0), grid)
            self._keep.append(it)
        h = C.c_void_p()
        _lib.check(_lib.lib().femtc_batch_create(C.byref(h), arr, len(items),
                                                 _stream(stream)),
                   "batch_create")
        self._h = h

    def run(self) -> None:
        _lib.check(_lib.lib().femtc_batch_run(self._h), "batch_run")

    def graph_launch(self) -> None:
        _lib.check(_lib.lib().femtc_batch_graph_launch(self._h),
                   "batch_graph_launch")

    def autotune(self, reps: int = 3) -> None:
        _lib.check(_lib.lib().femtc_batch_autotune(self._h, reps),
                   "batch_autotune")

    def stats(self):
        g, c, d = C.c_int(), C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().femtc_batch_stats(self._h, C.byref(g),
                                                C.byref(c), C.byref(d)),
                   "batch_stats")
        return {"groups": g.value, "cells": c.value, "dofs": d.value}

    def groups(self):
        out = []
        for g in range(self.stats()["groups"]):
            v = [C.c_int() for _ in range(7)]
            _lib.check(_lib.lib().femtc_batch_group(
                self._h, g, *[C.byref(a) for a in v]), "batch_group")
            out.append(dict(zip(("op", "dtype", "degree", "nq", "gather",
                                 "items", "per_wg"), [a.value for a in v])))
        return out

    def set_group(self, group: int, per_wg: int) -> None:
        """per_wg = 0: one launch per problem; k > 0: fused, k blocks per
        workgroup; -1: the assembled scatter per problem, -2: the z-run
        scatter per problem, -3: one fused z-run launch over the group
        (structured items only; -2 and -3 also nq <= 8)."""
        _lib.check(_lib.lib().femtc_batch_set_group(self._h, group, per_wg),
                   "batch_set_group")

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().femtc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
