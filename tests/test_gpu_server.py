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
        from femtc_amd import _lib
        budget = max(int(8 * 8 * it["ndofs"] + 6 * 4 * it["num_cells"] *
                         it["nq"] ** 3 + 4 * it["num_cells"] * it["nq"] ** 3)
                     for it in items) + (1 << 20)
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
