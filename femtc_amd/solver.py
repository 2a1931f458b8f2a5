"""Device-resident Jacobi-PCG and FP64 defect correction (libfemtc solver).

The reference has no solver; this is the solve path built on the operator:
A = M, K or K+M applied through the dofmap, a Jacobi preconditioner from the
sum-factorized diagonal, CG scalars kept on the device, and (inner=F32,
BF16 or BF16_ROUNDED) an inner PCG in FP32 vectors wrapped in FP64 iterative
refinement, its operator in FP32 or on the matrix cores in bfloat16.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .operators import _dtype_code, _ptr, _stream


def global_diagonal(op: int, tables, diag, dofmap, num_cells: int,
                    degree: int, nq: int, G=None, detJw=None,
                    stream=None) -> None:
    """diag (FP64, overwritten) = diagonal of the continuous operator."""
    _lib.check(_lib.lib().femtc_global_diagonal(
        op, _dtype_code(tables), degree, nq, _ptr(tables), _ptr(diag),
        diag.numel(), _ptr(dofmap), _ptr(G), _ptr(detJw), num_cells,
        _stream(stream)), "global_diagonal")


class PCGSolver:
    """Solver for A u = f on one problem (device tensors, FP64).

    inner=F64: plain PCG.  inner=F32: FP32 inner PCG (tolerance inner_rtol)
    inside FP64 defect correction.  inner=BF16 / BF16_ROUNDED: the same with
    the inner operator on the matrix cores (bfloat16 operands, FP32
    accumulation; needs J and grid, nq = degree + 1 <= 8).  The tensors
    passed here must stay alive and unchanged while the solver is used."""

    def __init__(self, op: int, tables, dofmap, num_cells: int, degree: int,
                 nq: int, ndofs: int, G=None, detJw=None, inner: int = _lib.F64,
                 stream=None, J=None, grid=None,
                 scatter: int = _lib.SCATTER_AUTO, even_odd=None):
        """J: optional per-cell Jacobians [C*9] (FP64) of affine cells; the
        inner operator then computes its geometry in the kernel.
        grid: optional (nx, ny, nz) of a structured cell numbering; when the
        dofmap is that numbering the CG iterations can use the assembled
        or the z-run scatter (scatter: SCATTER_AUTO times them against the
        atomic one, SCATTER_DIRECT / SCATTER_ASSEMBLED / SCATTER_ZRUN force
        one).
        even_odd: None times the even-odd form (nq = degree + 1 and
        symmetric tables, checked once, here) against the general one and
        keeps the faster; True uses it whenever the check passes; False
        keeps the general form."""
        self._keep = (tables, dofmap, G, detJw, J)
        self.ndofs = ndofs
        self.inner = inner
        g = (0, 0, 0) if grid is None else tuple(int(v) for v in grid)
        self.problem = _lib.Problem(op, degree, nq, num_cells, ndofs,
                                    _ptr(dofmap), _ptr(tables), _ptr(G),
                                    _ptr(detJw), _ptr(J), (C.c_int32 * 3)(*g),
                                    scatter,
                                    {None: 0, False: 1, True: 2}[even_odd])
        h = C.c_void_p()
        _lib.check(_lib.lib().femtc_solver_create(
            C.byref(h), C.byref(self.problem), inner, _stream(stream)),
            "solver_create")
        self._h = h

    def set_batch(self, check_every: int = 16, use_graph: bool = True) -> None:
        """CG iterations queued between convergence checks, and whether the
        batch is replayed from a captured hipGraph."""
        _lib.check(_lib.lib().femtc_solver_set_batch(self._h, check_every,
                                                     int(use_graph)),
                   "solver_set_batch")

    def scatter(self) -> int:
        """Scatter of the CG iterations chosen at creation: SCATTER_DIRECT,
        SCATTER_ASSEMBLED or SCATTER_ZRUN."""
        a = C.c_int32()
        _lib.check(_lib.lib().femtc_solver_scatter(self._h, C.byref(a)),
                   "solver_scatter")
        return a.value

    def bytes(self) -> int:
        """Device bytes the solver holds."""
        n = C.c_size_t()
        _lib.check(_lib.lib().femtc_solver_bytes(self._h, C.byref(n)),
                   "solver_bytes")
        return n.value

    def even_odd(self) -> bool:
        """Whether the solver's operators contract in even-odd form."""
        a = C.c_int32()
        _lib.check(_lib.lib().femtc_solver_even_odd(self._h, C.byref(a)),
                   "solver_even_odd")
        return bool(a.value)

    def solve(self, f, u, rtol: float = 1e-10, inner_rtol: float = 1e-4,
              maxit: int = 20000, check: bool = True) -> dict:
        """u = A^-1 f to rtol.  Returns the solve info as a dict; raises on
        failure unless check=False."""
        if self._h is None:
            raise RuntimeError("solver is closed")
        if f.numel() != self.ndofs or u.numel() != self.ndofs:
            raise ValueError("f and u must have ndofs entries")
        info = _lib.SolveInfo()
        st = _lib.lib().femtc_solver_solve(self._h, _ptr(f), _ptr(u), rtol,
                                           inner_rtol, maxit, C.byref(info))
        out = {k: getattr(info, k) for k, _ in _lib.SolveInfo._fields_}
        out["status_string"] = _lib.status_string(st)
        if check and st != _lib.OK:
            raise _lib.FemtcError(st, "solver_solve", _lib.last_error())
        return out

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().femtc_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
