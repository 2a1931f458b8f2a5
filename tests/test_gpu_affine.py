"""GPU parity of the affine-geometry operators (geometry computed in the
kernel from 7 values per cell) against the oracle's stored-geometry operator
on the same per-cell Jacobians, through the C ABI; and solvers that use it.
Bars: FP64 <= 1e-12 relative l2, FP32 <= 1e-5; per-cell geometry <= 1e-14
against numpy; solves reach 1e-10 and agree with the stored-geometry solve
to 1e-8."""
import numpy as np
import pytest

import femtc_amd as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda",
                                                         dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _box(nx, ny, nz, P, Q, seed):
    C = nx * ny * nz
    J = O.random_jacobians(seed, C)
    G, Dw = O.geometry_affine(J, Q)
    dm = O.dofmap_structured(nx, ny, nz, P)
    ng = (nx * P + 1) * (ny * P + 1) * (nz * P + 1)
    return C, J, G, Dw, dm, ng


def test_geometry_cell_closed_form():
    C = 300
    J = O.random_jacobians(3, C)
    out = torch.empty(C * 7, device="cuda", dtype=torch.float64)
    F.geometry_cell(dev(J), out)
    got = host(out).reshape(C, 7)
    for c in range(0, C, 37):
        Jc = J[9 * c:9 * c + 9].reshape(3, 3)
        inv = np.linalg.inv(Jc)
        S = abs(np.linalg.det(Jc)) * inv @ inv.T
        packed = [S[2, 2], S[1, 2], S[0, 2], S[1, 1], S[0, 1], S[0, 0],
                  abs(np.linalg.det(Jc))]
        assert rel(got[c], np.array(packed)) <= 1e-14


SHAPES = [(3, 2, 2, 1, 2), (2, 3, 2, 2, 3), (2, 2, 3, 3, 5), (2, 2, 2, 4, 5),
          (2, 1, 2, 5, 8), (2, 1, 2, 6, 7), (1, 2, 2, 7, 8), (1, 1, 2, 8, 9),
          (1, 1, 2, 8, 11)]


@pytest.mark.parametrize("nx,ny,nz,P,Q", SHAPES)
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_affine_operator_vs_oracle(op, nx, ny, nz, P, Q):
    C, J, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 30 + P)
    x = O.uniform(31, ng)
    ref = O.global_operator(op, P + 1, Q, O.basis_tables(P, Q), x, dm, G, Dw)
    tw = F.basis_tables_weights(P, Q)
    ddm, dJ = dev(dm, torch.int32), dev(J)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        cg = torch.empty(C * 7, device="cuda", dtype=dt)
        F.geometry_cell(dJ, cg)
        y = torch.full((ng,), 4.0, device="cuda", dtype=dt)
        F.global_apply_affine(op, dev(tw, dt), dev(x, dt), y, ddm, C, P, Q, cg)
        assert rel(host(y), ref) <= tol, dt
    # accumulate onto a non-zero y (FP64)
    cg = torch.empty(C * 7, device="cuda", dtype=torch.float64)
    F.geometry_cell(dJ, cg)
    y0 = O.uniform(32, ng)
    y = dev(y0)
    F.global_apply_affine(op, dev(tw), dev(x), y, ddm, C, P, Q, cg,
                          accumulate=True)
    assert rel(host(y), ref + y0) <= 1e-12


@pytest.mark.parametrize("inner", [F.F64, F.F32])
@pytest.mark.parametrize("stored", [True, False])
def test_solver_with_affine_cells(inner, stored):
    nx, ny, nz, P, Q = 3, 2, 3, 4, 5
    C, J, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 40)
    tab = O.basis_tables(P, Q)
    f = O.uniform(41, ng)
    u_ref, _, _ = O.pcg("kpm", P + 1, Q, tab, f, dm, G, Dw, rtol=1e-14)
    kw = dict(G=dev(G), detJw=dev(Dw)) if stored else {}
    with F.PCGSolver(F.OP_KPM, dev(tab), dev(dm, torch.int32), C, P, Q, ng,
                     inner=inner, J=dev(J), **kw) as S:
        u = torch.empty(ng, device="cuda", dtype=torch.float64)
        info = S.solve(dev(f), u, rtol=1e-10)
    uh = host(u)
    res = rel(O.global_operator("kpm", P + 1, Q, tab, uh, dm, G, Dw), f)
    assert info["relres"] <= 1.01e-10 and res <= 1.05e-10, (info, res)
    assert rel(uh, u_ref) <= 1e-8
