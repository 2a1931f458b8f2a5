"""Dofmap validation before any kernel indexes through it: a dofmap entry
outside [0, ndofs) is refused by femtc_solver_create and
femtc_batch_create (FEMTC_ERR_INVALID_ARGUMENT), with no device work done
through the bad entry."""
import numpy as np
import pytest

import femtc_amd as F
import oracle as O
from femtc_amd import workloads as W

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bad", [-1, "ndofs", 1 << 30])
def test_solver_rejects_out_of_range_dofmap(bad):
    it = W.structured_problem(torch, (2, 2, 2), 3, 4, F.OP_KPM, torch.float64, 3)
    dm = it["dofmap"].clone()
    dm[17] = it["ndofs"] if bad == "ndofs" else bad
    for inner in (F.F64, F.F32):
        with pytest.raises(F.FemtcError, match="outside"):
            F.PCGSolver(F.OP_KPM, it["tables"], dm, it["num_cells"], 3, 4,
                        it["ndofs"], G=it["G"], detJw=it["detJw"],
                        inner=inner)
    # the unchanged dofmap is accepted
    with F.PCGSolver(F.OP_KPM, it["tables"], it["dofmap"], it["num_cells"],
                     3, 4, it["ndofs"], G=it["G"], detJw=it["detJw"]) as S:
        u = torch.empty(it["ndofs"], device="cuda", dtype=torch.float64)
        S.solve(it["x"], u, rtol=1e-8)


def test_batch_rejects_out_of_range_dofmap():
    items = [W.structured_problem(torch, (2, 2, 2), P, P + 1, F.OP_MASS,
                                  torch.float64, 10 + P) for P in (2, 3)]
    items[1]["dofmap"][5] = items[1]["ndofs"]
    with pytest.raises(F.FemtcError, match="outside"):
        F.Batch(items)
    items[1]["dofmap"][5] = 0
    b = F.Batch(items)
    b.run()
    b.close()
