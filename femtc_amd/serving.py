"""The serving workload: a Poisson stream of apply and (K+M)u=f solve
requests on a pool of resident problems, fed to femtc_amd.Server at their
arrival times; reports p50 / p99 latency and throughput.

Pool: `n_problems` structured meshes n^3 with n uniform in [n_lo, n_hi],
degree uniform in `degrees`, nq = degree+1, J = I + 0.1 U(-1,1) per cell.
Requests: arrivals exponential with mean 1/rate; each picks a pool problem
uniformly; a fraction `solve_frac` are solves (FP32 inner PCG to 1e-4 inside
FP64 refinement to rtol), the rest applies; inputs uniform[-1,1]."""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .server import ResidentProblem, Server, latency_stats
from .workloads import structured_problem, uniform


def make_pool(torch, seed=0, n_problems=16, n_lo=4, n_hi=12,
              degrees=(2, 3, 4, 5, 6)):
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n_problems):
        n = int(rng.integers(n_lo, n_hi + 1))
        P = int(rng.choice(np.asarray(degrees)))
        items.append(structured_problem(torch, (n, n, n), P, P + 1,
                                        _lib.OP_KPM, torch.float64,
                                        5000 + 17 * i))
    return items


def run_stream(torch, items, rate=500.0, duration=2.0, solve_frac=0.5,
               slots=4, hbm_budget=16 << 30, rtol=1e-10, seed=1,
               inner=_lib.F32, max_requests=4000, affine=False,
               structured=False):
    """Feed the stream for `duration` seconds of arrivals, drain, and return
    (stats, results).  affine / structured: the resident problems also carry
    their cell Jacobians / grid (ResidentProblem.from_item)."""
    rng = np.random.default_rng(seed)
    probs = [ResidentProblem.from_item(it, affine=affine, structured=structured)
             for it in items]
    # inputs and outputs prepared up front (device-resident)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, max_requests))
    arrivals = arrivals[arrivals < duration]
    picks = rng.integers(len(items), size=len(arrivals))
    kinds = rng.random(len(arrivals)) < solve_frac
    vecs, outs = [], []
    for t, i in enumerate(picks):
        n = items[i]["ndofs"]
        vecs.append(torch.from_numpy(uniform(20000 + t, n)).cuda())
        outs.append(torch.empty(n, device="cuda", dtype=torch.float64))
    warm = [(torch.ones(it["ndofs"], device="cuda", dtype=torch.float64),
             torch.empty(it["ndofs"], device="cuda", dtype=torch.float64))
            for it in items]
    torch.cuda.synchronize()
    with Server(slots=slots, hbm_budget=hbm_budget) as srv:
        # warm-up: solve every problem `slots` times so most solvers are
        # resident when the stream starts; those results are dropped
        for i, p in enumerate(probs):
            for s in range(slots):
                srv.submit_solve(p, warm[i][0], warm[i][1],
                                 tag=-1 - i * slots - s, rtol=rtol,
                                 inner=inner)
        srv.drain()
        srv.results()
        t0 = time.perf_counter()
        for t, (at, i, solve) in enumerate(zip(arrivals, picks, kinds)):
            wait = at - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            if solve:
                srv.submit_solve(probs[i], vecs[t], outs[t], tag=t, rtol=rtol,
                                 inner=inner)
            else:
                srv.submit_apply(probs[i], vecs[t], outs[t], tag=t)
        srv.drain()
        res = [r for r in srv.results() if r["tag"] >= 0]
    st = latency_stats(res)
    st["offered_rate"] = rate
    st["duration_s"] = duration
    st["failed"] = sum(r["status"] != 0 for r in res)
    st["max_relres"] = max((r["relres"] for r in res
                            if r["kind"] == _lib.REQUEST_SOLVE), default=None)
    return st, res
