"""GPU parity of batched applies (femtc_batch_*) through the C ABI: every
item of a mixed batch (degrees, nq, operators, continuous and cell-local,
FP64 and FP32) equals the same problem applied alone and the oracle; hipGraph
replays equal the eager run, with eager work and synchronisation between
replays; the autotuner keeps results unchanged."""
import numpy as np
import pytest

import femtc_amd as F
import oracle as O
from femtc_amd import workloads as W

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SPECS = [((2, 3, 4), 1, 2, F.OP_MASS, True), ((3, 3, 3), 2, 3, F.OP_STIFFNESS, True),
         ((1, 1, 5), 3, 5, F.OP_KPM, False), ((2, 2, 2), 4, 5, F.OP_STIFFNESS, False),
         ((4, 1, 2), 5, 6, F.OP_MASS, True), ((2, 2, 2), 6, 7, F.OP_KPM, True),
         ((1, 2, 2), 7, 8, F.OP_STIFFNESS, True), ((1, 1, 3), 8, 9, F.OP_MASS, False),
         ((2, 1, 1), 8, 10, F.OP_KPM, True), ((3, 2, 2), 2, 3, F.OP_STIFFNESS, True),
         ((1, 1, 1), 6, 7, F.OP_KPM, True), ((5, 5, 5), 3, 4, F.OP_STIFFNESS, True)]


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _items(dtype, seed0=100, specs=SPECS):
    return [W.structured_problem(torch, d, P, Q, op, dtype, seed0 + i, cont)
            for i, (d, P, Q, op, cont) in enumerate(specs)]


def _alone(it):
    """The same problem through the single-problem entry points."""
    y = torch.empty_like(it["y"])
    P, Q, C = it["degree"], it["nq"], it["num_cells"]
    if it["dofmap"] is not None:
        F.global_apply(it["op"], it["tables"], it["x"], y, it["dofmap"], C, P,
                       Q, G=it["G"], detJw=it["detJw"])
    elif it["op"] == F.OP_MASS:
        F.mass_apply(it["tables"], it["x"], y, it["detJw"], C, P, Q)
    elif it["op"] == F.OP_STIFFNESS:
        F.stiffness_apply(it["tables"], it["x"], y, it["G"], C, P, Q)
    else:
        F.kpm_apply(it["tables"], it["x"], y, it["G"], it["detJw"], C, P, Q)
    return y


def _oracle(it):
    P, Q = it["degree"], it["nq"]
    tab = host(it["tables"])
    G = None if it["G"] is None else host(it["G"])
    D = None if it["detJw"] is None else host(it["detJw"])
    x = host(it["x"])
    if it["dofmap"] is not None:
        return O.global_operator(it["op"], P + 1, Q, tab, x,
                                 it["dofmap"].cpu().numpy(), G, D)
    return O.cell_operator(it["op"], P + 1, Q, tab, x, G, D, it["num_cells"])


def _check(items, tol_oracle):
    for it in items:
        ya = _alone(it)
        # atomics: the summation order (last bits) may differ between runs
        tol = 1e-13 if it["y"].dtype == torch.float64 else 1e-6
        if it["dofmap"] is None:  # cell-local: same kernel body, same bits
            assert torch.equal(it["y"], ya)
        else:
            assert rel(host(it["y"]), host(ya)) <= tol
        assert rel(host(it["y"]), _oracle(it)) <= tol_oracle


def test_batch_mixed_f64_vs_alone_and_oracle():
    items = _items(torch.float64)
    for it in items:
        it["y"].fill_(7.0)  # outputs are overwritten
    b = F.Batch(items)
    st = b.stats()
    assert st["groups"] == len(SPECS) - 2  # two pairs share a group
    b.run()
    _check(items, 1e-12)


@pytest.mark.parametrize("per_wg", [1, 3, 8])
def test_batch_fused_launch_vs_alone_and_oracle(per_wg):
    items = _items(torch.float64, seed0=200)
    b = F.Batch(items)
    for g in range(b.stats()["groups"]):
        b.set_group(g, per_wg)
    assert all(g["per_wg"] == per_wg for g in b.groups())
    b.run()
    _check(items, 1e-12)
    b.graph_launch()
    torch.cuda.synchronize()
    _check(items, 1e-12)


def test_batch_f32_vs_oracle():
    items = _items(torch.float32, seed0=300)
    b = F.Batch(items)
    b.run()
    _check(items, 1e-5)


def test_batch_graph_replays_match_eager():
    items = _items(torch.float64, seed0=500)
    b = F.Batch(items)
    for rep in range(3):
        # eager work between replays: change the inputs, recompute eagerly
        for it in items:
            it["x"].mul_(-1.5 if rep % 2 else 0.5)
        torch.cuda.synchronize()
        b.run()
        ref = [it["y"].clone() for it in items]
        for it in items:
            it["y"].fill_(3.0)
        torch.cuda.synchronize()
        b.graph_launch()
        torch.cuda.synchronize()
        for it, r in zip(items, ref):
            if it["dofmap"] is None:
                assert torch.equal(it["y"], r), rep
            else:
                assert rel(host(it["y"]), host(r)) <= 1e-13, rep


def test_batch_autotune_keeps_results():
    items = _items(torch.float64, seed0=700)
    b = F.Batch(items)
    b.graph_launch()  # a graph captured before tuning is dropped by it
    b.autotune(reps=2)
    for g in b.groups():
        assert g["per_wg"] in (0, 1, 2, 4, 8) or (
            g["per_wg"] in (-1, -2, -3) and g["gather"] == 1)
    b.graph_launch()
    torch.cuda.synchronize()
    _check(items, 1e-12)


@pytest.mark.parametrize("scatter", [-1, -2, -3])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12),
                                       (torch.float32, 1e-5)])
def test_batch_structured_scatters_vs_alone_and_oracle(dtype, tol, scatter):
    # continuous groups on the assembled (-1) or z-run (-2, per problem;
    # -3, one fused launch per group) scatter (every item carries its
    # verified grid), cell-local groups as they are; eager, graph replay,
    # and a graph replay after eager work
    items = _items(dtype, seed0=800)
    b = F.Batch(items)
    for g, info in enumerate(b.groups()):
        if info["gather"] and (scatter == -1 or info["nq"] <= 8):
            b.set_group(g, scatter)
            assert b.groups()[g]["per_wg"] == scatter
        else:
            with pytest.raises(F.FemtcError, match="structured grid"):
                b.set_group(g, scatter)
    for it in items:
        it["y"].fill_(7.0)
    b.run()
    _check(items, tol)
    for rep in range(2):
        for it in items:
            it["y"].fill_(-3.0)
            it["x"].mul_(0.5)
        torch.cuda.synchronize()
        b.graph_launch()
        torch.cuda.synchronize()
        _check(items, tol)


def test_batch_mode_switches_keep_the_clears():
    # the clears follow each mode change: assembled (no clear: the second
    # pass overwrites y), then per-problem atomic and fused (y cleared
    # first), each run on outputs holding stale values, eager and replayed
    items = _items(torch.float64, seed0=950)
    b = F.Batch(items)
    cont = [g for g, info in enumerate(b.groups()) if info["gather"]]
    assert cont
    for mode in (-1, 0, 2, -1):
        for g in cont:
            b.set_group(g, mode)
        for launch in (b.run, b.graph_launch):
            for it in items:
                it["y"].fill_(5.0)
            torch.cuda.synchronize()
            launch()
            torch.cuda.synchronize()
            _check(items, 1e-12)


def test_batch_assembled_scatter_needs_structured_items():
    # a relabelled dofmap is not the structured numbering: its group keeps
    # the atomic scatter, the others may switch
    items = _items(torch.float64, seed0=900, specs=SPECS[:2])
    it = items[1]
    ng = it["ndofs"]
    perm = torch.randperm(ng, generator=torch.Generator().manual_seed(1))
    it["dofmap"] = perm.to(device="cuda", dtype=torch.int32)[it["dofmap"].long()]
    b = F.Batch(items)
    ops = {g["op"]: i for i, g in enumerate(b.groups())}
    b.set_group(ops[items[0]["op"]], -1)
    with pytest.raises(F.FemtcError, match="structured grid"):
        b.set_group(ops[it["op"]], -1)
    with pytest.raises(F.FemtcError, match="structured grid"):
        b.set_group(ops[it["op"]], -2)
    with pytest.raises(F.FemtcError, match="structured grid"):
        b.set_group(ops[it["op"]], -3)
    b.autotune(reps=1)
    assert b.groups()[ops[it["op"]]]["per_wg"] not in (-1, -2, -3)
    b.run()
    _check(items, 1e-12)


def test_uneven_batch_spec_small():
    spec = W.uneven_batch_spec(seed=3, n_meshes=24, mu=np.log(60), sigma=1.0,
                               lo=5, hi=400)
    items = [W.structured_problem(torch, W.box_dims(c), P, P + 1, op,
                                  torch.float64, 900 + i)
             for i, (c, P, op) in enumerate(spec)]
    b = F.Batch(items)
    b.autotune(reps=1)
    b.graph_launch()
    torch.cuda.synchronize()
    for it in items:
        assert rel(host(it["y"]), host(_alone(it))) <= 1e-13
