"""GPU tests of the request server (femtc_server_*) through the C ABI: every
result of a mixed stream of apply and solve requests served concurrently is
the same as the request run alone (apply: 1e-13 relative, the atomic
summation order aside; solve: FP64 residual <= rtol and solution within
1e-8 of the lone solve), under a roomy and under a tight HBM budget."""
import numpy as np
import pytest

import femtc_amd as F
import oracle as O
from femtc_amd import workloads as W

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = a.double(), b.double()
    nb = float(b.norm())
    return float((a - b).norm()) / (nb if nb > 0 else 1.0)


POOL = [((4, 4, 4), 2, F.OP_KPM), ((3, 3, 3), 3, F.OP_KPM),
        ((2, 3, 3), 4, F.OP_KPM), ((3, 3, 2), 2, F.OP_MASS),
        ((2, 2, 2), 6, F.OP_KPM)]


def _pool():
    items = [W.structured_problem(torch, d, P, P + 1, op, torch.float64, 40 + i)
             for i, (d, P, op) in enumerate(POOL)]
    return items, [F.ResidentProblem.from_item(it) for it in items]


def _requests(items, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for t in range(n):
        i = int(rng.integers(len(items)))
        kind = "apply" if rng.random() < 0.4 else "solve"
        inner = F.F32 if rng.random() < 0.6 else F.F64
        v = torch.from_numpy(W.uniform(1000 + t, items[i]["ndofs"])).cuda()
        reqs.append((t, i, kind, inner, v))
    return reqs


def _alone(items, req):
    t, i, kind, inner, v = req
    it = items[i]
    out = torch.empty_like(v)
    if kind == "apply":
        F.global_apply(it["op"], it["tables"], v, out, it["dofmap"],
                       it["num_cells"], it["degree"], it["nq"], G=it["G"],
                       detJw=it["detJw"])
    else:
        with F.PCGSolver(it["op"], it["tables"], it["dofmap"],
                         it["num_cells"], it["degree"], it["nq"],
                         it["ndofs"], G=it["G"], detJw=it["detJw"],
                         inner=inner) as S:
            S.solve(v, out, rtol=1e-10)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("budget", [8 << 30, 0])
def test_results_in_the_mix_equal_results_alone(budget):
    items, probs = _pool()
    reqs = _requests(items, 40, seed=1)
    if budget == 0:
        # room for about one solver: forces waiting and eviction
        from femtc_amd.server import workspace_bytes
        budget = max(workspace_bytes(p, inner) for p in probs
                     for inner in (F.F32, F.F64)) + (1 << 20)
    alone = [_alone(items, r) for r in reqs]
    outs = [torch.full_like(r[4], 9.0) for r in reqs]
    torch.cuda.synchronize()  # the slots run on their own streams
    with F.Server(slots=3, hbm_budget=budget) as srv:
        for (t, i, kind, inner, v), o in zip(reqs, outs):
            if kind == "apply":
                srv.submit_apply(probs[i], v, o, tag=t)
            else:
                srv.submit_solve(probs[i], v, o, tag=t, rtol=1e-10,
                                 inner=inner)
        srv.drain()
        res = {r["tag"]: r for r in srv.results()}
        used, cached = srv.memory()
        assert used <= budget
    assert sorted(res) == list(range(len(reqs)))
    slots = set()
    for (t, i, kind, inner, v), o, a in zip(reqs, outs, alone):
        r = res[t]
        assert r["status"] == 0, r
        assert r["t_arrival"] <= r["t_start"] <= r["t_end"]
        slots.add(r["slot"])
        if kind == "apply":
            assert rel(o, a) <= 1e-13
        else:
            assert r["relres"] <= 1e-10
            assert rel(o, a) <= 1e-8
            # independent FP64 residual with the oracle
            it = items[i]
            f = v.cpu().numpy()
            Au = O.global_operator(it["op"], it["degree"] + 1, it["nq"],
                                   it["tables"].cpu().numpy(),
                                   o.cpu().numpy(),
                                   it["dofmap"].cpu().numpy(),
                                   None if it["G"] is None else it["G"].cpu().numpy(),
                                   it["detJw"].cpu().numpy())
            assert np.linalg.norm(f - Au) / np.linalg.norm(f) <= 1.05e-10
    assert len(slots) > 1  # the requests were actually spread over slots
    st = F.latency_stats(list(res.values()))
    assert st["n"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0


def test_request_larger_than_budget_is_refused():
    items, probs = _pool()
    with F.Server(slots=1, hbm_budget=1024) as srv:
        v = torch.zeros(items[0]["ndofs"], device="cuda", dtype=torch.float64)
        torch.cuda.synchronize()
        with pytest.raises(F.FemtcError, match="budget"):
            srv.submit_solve(probs[0], v, torch.empty_like(v), tag=0)
        # applies need no workspace and are served
        srv.submit_apply(probs[0], v, torch.empty_like(v), tag=1)
        srv.drain()
        assert [r["tag"] for r in srv.results()] == [1]


def test_reregistered_address_gets_a_fresh_solver():
    """A problem unregistered and replaced by another at the same address
    (the same femtc_problem struct, refilled) must not meet the first
    problem's resident solver: the second solve answers the second problem."""
    import ctypes

    items, probs = _pool()
    a, b = items[0], items[1]
    pa, pb = probs[0], probs[1]
    fa = torch.from_numpy(W.uniform(7, a["ndofs"])).cuda()
    fb = torch.from_numpy(W.uniform(8, b["ndofs"])).cuda()
    ua, ub = torch.empty_like(fa), torch.empty_like(fb)
    torch.cuda.synchronize()
    with F.Server(slots=1, hbm_budget=8 << 30) as srv:
        srv.submit_solve(pa, fa, ua, tag=0)
        srv.drain()
        assert srv.memory()[1] == 1  # a's solver is resident
        srv.unregister(pa)
        assert srv.memory()[1] == 0
        # the same struct (same address) now describes problem b
        ctypes.memmove(ctypes.addressof(pa.struct),
                       ctypes.addressof(pb.struct), ctypes.sizeof(pa.struct))
        pa._keep, pa.ndofs = pb._keep, pb.ndofs
        srv.submit_solve(pa, fb, ub, tag=1)
        srv.drain()
        res = {r["tag"]: r for r in srv.results()}
    assert res[0]["status"] == 0 and res[1]["status"] == 0
    Au = O.global_operator(b["op"], b["degree"] + 1, b["nq"],
                           b["tables"].cpu().numpy(), ub.cpu().numpy(),
                           b["dofmap"].cpu().numpy(), b["G"].cpu().numpy(),
                           b["detJw"].cpu().numpy())
    f = fb.cpu().numpy()
    assert np.linalg.norm(f - Au) / np.linalg.norm(f) <= 1.05e-10


def test_unregistered_problem_is_refused_and_queued_requests_fail():
    items, probs = _pool()
    v = torch.zeros(items[0]["ndofs"], device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    with F.Server(slots=1, hbm_budget=8 << 30) as srv:
        from femtc_amd import _lib
        req = _lib.Request(_lib.REQUEST_APPLY,
                           __import__("ctypes").addressof(probs[0].struct),
                           v.data_ptr(), torch.empty_like(v).data_ptr(), 0.0,
                           0.0, 0, _lib.F64, 5)
        st = _lib.lib().femtc_server_submit(srv._h,
                                            __import__("ctypes").byref(req))
        assert st == _lib.ERR_INVALID_ARGUMENT
        assert "not registered" in _lib.last_error()


def test_bf16_inner_solves_through_the_server():
    """Solve requests with the BF16-MFMA inner operator (problems with cell
    Jacobians and the grid) reach 1e-10 and match the FP32 inner solve."""
    items = [W.structured_problem(torch, d, P, P + 1, F.OP_KPM, torch.float64,
                                  60 + i)
             for i, (d, P) in enumerate((((3, 3, 3), 4), ((2, 3, 2), 6)))]
    probs = [F.ResidentProblem.from_item(it, affine=True, structured=True)
             for it in items]
    fs = [torch.from_numpy(W.uniform(9 + i, it["ndofs"])).cuda()
          for i, it in enumerate(items)]
    outs = {k: [torch.empty_like(f) for f in fs] for k in (F.F32, F.BF16)}
    torch.cuda.synchronize()
    with F.Server(slots=2, hbm_budget=8 << 30) as srv:
        tag = 0
        for inner in (F.F32, F.BF16):
            for p, f, u in zip(probs, fs, outs[inner]):
                srv.submit_solve(p, f, u, tag=tag, inner=inner)
                tag += 1
        srv.drain()
        res = srv.results()
    assert len(res) == 4 and all(r["status"] == 0 for r in res)
    assert all(r["relres"] <= 1e-10 for r in res)
    for u32, ubf in zip(outs[F.F32], outs[F.BF16]):
        assert rel(ubf, u32) <= 1e-8
