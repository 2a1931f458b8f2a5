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


def run_sweep(torch, rates=(2.0, 5.0, 10.0, 20.0), duration=4.0,
              n_problems=4, n=48, degree=6, solve_frac=0.5, slots=4,
              hbm_budget=160 << 30, rtol=1e-10, seed=7, inner=_lib.F32,
              burst=16, burst_slots=(1, 2, 4)):
    """The serving stream at full solve sizes: `n_problems` resident meshes
    n^3 (48^3, P=6, Q=7: 24.1 M dofs each, with their Jacobians and grid so
    the solvers take per-cell geometry and the assembled scatter), requests
    arriving as a Poisson stream at each offered rate for `duration` s, a
    fraction `solve_frac` of them (K+M)u=f solves (FP32 inner PCG inside FP64
    refinement to rtol) and the rest applies.  Inputs are one f per mesh,
    outputs one vector per request.  Then `burst` solves submitted at once
    to a server of each slot count in burst_slots give the saturation
    throughput.  Returns a dict: per rate, latency_stats of all requests
    plus the applies' and the solves' p50/p99, and per burst the solves per
    second."""
    rng = np.random.default_rng(seed)
    items = [structured_problem(torch, (n, n, n), degree, degree + 1,
                                _lib.OP_KPM, torch.float64, 7000 + 31 * i)
             for i in range(n_problems)]
    probs = [ResidentProblem.from_item(it, affine=True, structured=True)
             for it in items]
    ng = items[0]["ndofs"]
    fs = [torch.from_numpy(uniform(9000 + i, ng)).cuda()
          for i in range(n_problems)]
    torch.cuda.synchronize()
    out = {"workload": (f"{n_problems} resident meshes {n}^3, P={degree}, "
                        f"Q={degree + 1}, {ng} dofs each; Poisson arrivals, "
                        f"{int(solve_frac * 100)} % (K+M)u=f solves (FP32 "
                        "inner PCG, FP64 refinement to 1e-10), the rest "
                        f"applies; {slots} slots"),
           "rates": {}}
    with Server(slots=slots, hbm_budget=hbm_budget) as srv:
        # every (problem, slot) solver created before the stream starts
        warm = torch.empty(ng, device="cuda", dtype=torch.float64)
        for i, p in enumerate(probs):
            for s in range(slots):
                srv.submit_solve(p, fs[i], warm, tag=-1 - i * slots - s,
                                 rtol=rtol, inner=inner)
        srv.drain()
        srv.results()
        for rate in rates:
            arrivals = np.cumsum(rng.exponential(1.0 / rate,
                                                 int(rate * duration * 3) + 8))
            arrivals = arrivals[arrivals < duration]
            picks = rng.integers(n_problems, size=len(arrivals))
            kinds = rng.random(len(arrivals)) < solve_frac
            outs = [torch.empty(ng, device="cuda", dtype=torch.float64)
                    for _ in range(len(arrivals))]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for t, (at, i, solve) in enumerate(zip(arrivals, picks, kinds)):
                wait = at - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                if solve:
                    srv.submit_solve(probs[i], fs[i], outs[t], tag=t,
                                     rtol=rtol, inner=inner)
                else:
                    srv.submit_apply(probs[i], fs[i], outs[t], tag=t)
            srv.drain()
            res = [r for r in srv.results() if r["tag"] >= 0]
            st = latency_stats(res)
            app = [r for r in res if r["kind"] == _lib.REQUEST_APPLY]
            if app:
                a = latency_stats(app)
                st["apply_p50_ms"], st["apply_p99_ms"] = (a["p50_ms"],
                                                          a["p99_ms"])
            st["failed"] = sum(r["status"] != 0 for r in res)
            st["max_relres"] = max((r["relres"] for r in res
                                    if r["kind"] == _lib.REQUEST_SOLVE),
                                   default=None)
            out["rates"][str(rate)] = st
            del outs
    # saturation: `burst` solves at once, per slot count (a server of its
    # own, its solvers created first)
    out["burst"] = {}
    outs = [torch.empty(ng, device="cuda", dtype=torch.float64)
            for _ in range(burst)]
    for ns in burst_slots:
        with Server(slots=ns, hbm_budget=hbm_budget) as srv:
            for i, p in enumerate(probs):
                for s in range(ns):
                    srv.submit_solve(p, fs[i], outs[0], tag=-1 - i * ns - s,
                                     rtol=rtol, inner=inner)
            srv.drain()
            srv.results()
            torch.cuda.synchronize()
            for t in range(burst):
                srv.submit_solve(probs[t % n_problems], fs[t % n_problems],
                                 outs[t], tag=t, rtol=rtol, inner=inner)
            srv.drain()
            res = [r for r in srv.results() if r["tag"] >= 0]
        b = latency_stats(res)
        out["burst"][str(ns)] = {"n": burst, "slots": ns,
                                 "solves_per_s": b["solves_per_s"],
                                 "p50_ms": b["p50_ms"], "p99_ms": b["p99_ms"],
                                 "failed": sum(r["status"] != 0 for r in res)}
    out["saturation_solves_per_s"] = max(v["solves_per_s"]
                                         for v in out["burst"].values())
    return out
