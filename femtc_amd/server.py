"""Request server over resident problems (libfemtc femtc_server_*).

`slots` host threads with one HIP stream each take apply and solve requests
as they arrive; requests admitted while others run overlap on the device.
Solvers stay resident per (problem, inner precision, slot) and are admitted
against an HBM budget."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .operators import _ptr


class ResidentProblem:
    """A problem kept on the device: op, degree, nq, tables, dofmap, G,
    detJw (FP64 torch tensors), num_cells, ndofs."""

    def __init__(self, op, degree, nq, tables, dofmap, num_cells, ndofs,
                 G=None, detJw=None, J=None, grid=None):
        self._keep = (tables, dofmap, G, detJw, J)
        self.ndofs = ndofs
        g = (0, 0, 0) if grid is None else tuple(int(v) for v in grid)
        self.struct = _lib.Problem(op, degree, nq, num_cells, ndofs,
                                   _ptr(dofmap), _ptr(tables), _ptr(G),
                                   _ptr(detJw), _ptr(J), (C.c_int32 * 3)(*g),
                                   _lib.SCATTER_AUTO)

    @classmethod
    def from_item(cls, it, affine: bool = False, structured: bool = False):
        """From a femtc_amd.workloads.structured_problem dict; affine=True
        also hands the cell Jacobians to the solvers, structured=True the
        grid (the solvers may then use the assembled scatter)."""
        return cls(it["op"], it["degree"], it["nq"], it["tables"],
                   it["dofmap"], it["num_cells"], it["ndofs"], it.get("G"),
                   it.get("detJw"), it.get("J") if affine else None,
                   it.get("dims") if structured else None)


def workspace_bytes(problem: ResidentProblem, inner: int = _lib.F32) -> int:
    """The device bytes the server reserves for a new solver of problem."""
    n = C.c_size_t()
    _lib.check(_lib.lib().femtc_solver_workspace_bytes(
        C.byref(problem.struct), inner, C.byref(n)), "solver_workspace_bytes")
    return n.value


class Server:
    """The slots run on streams of their own: a request's input must be
    complete on the device (e.g. torch.cuda.synchronize()) when it is
    submitted, and its output is ready once its result is returned.

    A problem is registered with the server at its first request (or by
    register()) and stays registered, and alive, until unregister() or
    close(); unregistering destroys its resident solvers."""

    def __init__(self, slots: int = 4, hbm_budget: int = 8 << 30):
        h = C.c_void_p()
        _lib.check(_lib.lib().femtc_server_create(C.byref(h), slots,
                                                  hbm_budget),
                   "server_create")
        self._h = h
        self._inflight = {}  # tag -> tensors kept alive until completion
        self._problems = {}  # id -> registered ResidentProblem

    def register(self, problem: ResidentProblem) -> None:
        if id(problem) in self._problems:
            return
        _lib.check(_lib.lib().femtc_server_register(
            self._h, C.byref(problem.struct)), "server_register")
        self._problems[id(problem)] = problem

    def unregister(self, problem: ResidentProblem) -> None:
        """Forget the problem: its queued requests complete with an error,
        its resident solvers are destroyed (after running ones finish)."""
        if id(problem) not in self._problems:
            return
        _lib.check(_lib.lib().femtc_server_unregister(
            self._h, C.byref(problem.struct)), "server_unregister")
        del self._problems[id(problem)]

    def _submit(self, req, keep):
        self.register(keep[0])
        self._inflight[req.tag] = keep
        st = _lib.lib().femtc_server_submit(self._h, C.byref(req))
        if st != _lib.OK:
            del self._inflight[req.tag]
            _lib.check(st, "server_submit")

    def submit_apply(self, problem: ResidentProblem, x, y, tag: int) -> None:
        req = _lib.Request(_lib.REQUEST_APPLY, C.addressof(problem.struct),
                           _ptr(x), _ptr(y), 0.0, 0.0, 0, _lib.F64, tag)
        self._submit(req, (problem, x, y))

    def submit_solve(self, problem: ResidentProblem, f, u, tag: int,
                     rtol: float = 1e-10, inner: int = _lib.F32,
                     inner_rtol: float = 1e-4, maxit: int = 20000) -> None:
        req = _lib.Request(_lib.REQUEST_SOLVE, C.addressof(problem.struct),
                           _ptr(f), _ptr(u), rtol, inner_rtol, maxit, inner,
                           tag)
        self._submit(req, (problem, f, u))

    def drain(self) -> None:
        _lib.check(_lib.lib().femtc_server_drain(self._h), "server_drain")

    def results(self):
        out = []
        buf = (_lib.RequestResult * 256)()
        while True:
            n = C.c_int()
            _lib.check(_lib.lib().femtc_server_results(self._h, buf, 256,
                                                       C.byref(n)),
                       "server_results")
            for i in range(n.value):
                r = buf[i]
                out.append({k: getattr(r, k) for k, _ in
                            _lib.RequestResult._fields_})
                self._inflight.pop(r.tag, None)
            if n.value < 256:
                return out

    def memory(self):
        used, cached = C.c_size_t(), C.c_int()
        _lib.check(_lib.lib().femtc_server_memory(self._h, C.byref(used),
                                                  C.byref(cached)),
                   "server_memory")
        return used.value, cached.value

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().femtc_server_destroy(self._h)
            self._h = None
            self._inflight.clear()
            self._problems.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def latency_stats(results) -> dict:
    """p50 / p99 / mean latency (arrival to completion, ms) and throughput
    over the span from the first arrival to the last completion."""
    if not results:
        return {"n": 0}
    lat = np.array([r["t_end"] - r["t_arrival"] for r in results]) * 1e3
    span = max(r["t_end"] for r in results) - min(r["t_arrival"]
                                                   for r in results)
    is_solve = np.array([r["kind"] == _lib.REQUEST_SOLVE for r in results])
    out = {"n": len(results), "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()),
           "requests_per_s": len(results) / span if span > 0 else None,
           "solves_per_s": int(is_solve.sum()) / span if span > 0 else None}
    if is_solve.any():
        out["solve_p50_ms"] = float(np.percentile(lat[is_solve], 50))
        out["solve_p99_ms"] = float(np.percentile(lat[is_solve], 99))
    return out


def simulate_admission(requests, slots: int, budget: int, nproblems=None,
                       remove_at=None):
    """Run the server's admission, queueing and eviction code with a fake
    clock and a fake allocator (no device work).  requests: dicts with
    t_arrival, kind (REQUEST_APPLY / REQUEST_SOLVE), problem, inner, bytes,
    actual, service, create.  Returns (results, stats) as dicts."""
    n = len(requests)
    if nproblems is None:
        nproblems = 1 + max((r["problem"] for r in requests), default=0)
    reqs = (_lib.SimRequest * max(n, 1))()
    for i, r in enumerate(requests):
        reqs[i] = _lib.SimRequest(r["t_arrival"], r["kind"], r["problem"],
                                  r.get("inner", _lib.F32), r.get("bytes", 0),
                                  r.get("actual", r.get("bytes", 0)),
                                  r["service"], r.get("create", 0.0))
    out = (_lib.SimResult * max(n, 1))()
    stats = _lib.SimStats()
    rem = None
    if remove_at is not None:
        rem = (C.c_double * nproblems)(*remove_at)
    _lib.check(_lib.lib().femtc_admission_simulate(
        reqs, n, nproblems, rem, slots, budget, out, C.byref(stats)),
        "admission_simulate")
    res = [{k: getattr(out[i], k) for k, _ in _lib.SimResult._fields_}
           for i in range(n)]
    return res, {k: getattr(stats, k) for k, _ in _lib.SimStats._fields_}
