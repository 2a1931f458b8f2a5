"""The allocation-free apply (femtc_global_apply_mfma_noalloc) against the
workspace entry point femtc_global_apply_mfma, through the C ABI: two calls
in a row, the second on a smaller grid, each bit-identical to the workspace
apply on the same inputs (same kernels, no atomics), and against the FP64
oracle at the BF16 bar; plus the size limit's error."""
import numpy as np
import pytest

import femtc_amd as F
import oracle as O
from femtc_amd import _lib
from femtc_amd.mfma import NOALLOC_LATTICE_VALUES, global_apply_mfma_noalloc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, P, seed):
    Q = P + 1
    C = nx * ny * nz
    J = O.random_jacobians(seed, C)
    G, Dw = O.geometry_affine(J, Q)
    dm = O.dofmap_structured(nx, ny, nz, P)
    ng = (nx * P + 1) * (ny * P + 1) * (nz * P + 1)
    tab = O.basis_tables(P, Q)
    x = O.uniform(seed + 1, ng)
    cg = torch.empty(C * 7, device="cuda", dtype=torch.float32)
    F.geometry_cell(torch.from_numpy(J).cuda(), cg)
    ops = torch.empty(F.mfma_tables_bytes(1), device="cuda", dtype=torch.uint8)
    F.mfma_tables(P, Q, tab, ops, split=1)
    ref = O.global_operator("kpm", P + 1, Q, tab, x, dm, G, Dw)
    return (nx, ny, nz), P, Q, C, ng, cg, ops, x, ref


@pytest.mark.parametrize("first,second", [((6, 5, 4, 6), (2, 3, 2, 6)),
                                          ((4, 4, 4, 4), (3, 2, 1, 2))])
def test_noalloc_equals_workspace_apply(first, second):
    for k, (nx, ny, nz, P) in enumerate((first, second)):
        grid, P, Q, C, ng, cg, ops, x, ref = _case(nx, ny, nz, P, 80 + k)
        xd = torch.from_numpy(x).float().cuda()
        lat = torch.empty(C * (P + 1) ** 3, device="cuda", dtype=torch.float32)
        y_ws = torch.empty(ng, device="cuda", dtype=torch.float32)
        F.global_apply_mfma(F.OP_KPM, P, Q, ops, xd, y_ws, grid, cg, lat)
        y_na = torch.full((ng,), 7.0, device="cuda", dtype=torch.float32)
        global_apply_mfma_noalloc(F.OP_KPM, P, Q, ops, xd, y_na, grid, cg)
        torch.cuda.synchronize()
        assert torch.equal(y_ws, y_na), (k, grid)
        got = y_na.double().cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= 5e-5, err
        # accumulate mode as well
        y_acc = y_na.clone()
        global_apply_mfma_noalloc(F.OP_KPM, P, Q, ops, xd, y_acc, grid, cg,
                                  accumulate=True)
        torch.cuda.synchronize()
        assert torch.allclose(y_acc, 2 * y_na, rtol=1e-6, atol=0)


def test_noalloc_rejects_grids_beyond_the_scratch():
    n = int(round((NOALLOC_LATTICE_VALUES / 343) ** (1 / 3))) + 1  # P = 6
    ng = (6 * n + 1) ** 3
    y = torch.empty(1, device="cuda", dtype=torch.float32)
    st = _lib.lib().femtc_global_apply_mfma_noalloc(
        F.OP_KPM, 6, 7, 1, y.data_ptr(), y.data_ptr(), y.data_ptr(), ng, n, n,
        n, y.data_ptr(), 0, None)
    assert st == _lib.ERR_INVALID_ARGUMENT
    assert "static scratch" in _lib.last_error()
