"""GPU parity of the high-degree kernels (degrees 9..25, the top of the
reference's dispatch ladder) through the C ABI against the CPU oracle:
cell-local mass, stiffness and K+M, the reference's literal mass and
stiffness, the dofmap gather with the atomic scatter, the affine-geometry
variants and the reference-signature host shim.  The operator tests run
twice, with the line kernels where one exists for the shape (degrees 9..25,
femtc_apply_kernel) and with the general kernel forced
(femtc_set_high_degree_kernel).

Bars: FP64 relative l2 <= 1e-12 against the oracle (whose literal operators
match the compiled reference bit for bit at these degrees,
test_oracle_golden.py); FP32 against the FP64 oracle at 1e-5 with GLL/Gauss
tables.  Plans, solvers and batches stay at degrees 1..8 and must say so.
"""
import numpy as np
import pytest

import femtc_amd as F
from femtc_amd import _lib
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


@pytest.fixture(params=[F.HIGH_DEGREE_AUTO, F.HIGH_DEGREE_GENERAL],
                ids=["line", "general"])
def hd_kernel(request):
    with F.high_degree_kernel(request.param):
        yield request.param


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda",
                                                         dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _case(P, Q, ncell, seed, table="gll"):
    N = P + 1
    tab = O.basis_tables(P, Q) if table == "gll" else O.uniform(seed, 2 * N * Q)
    U = O.uniform(seed + 1, ncell * N ** 3)
    G = O.uniform(seed + 2, ncell * Q ** 3 * 6, 0.0, 1.0)
    Dw = O.uniform(seed + 3, ncell * Q ** 3, 0.5, 1.5)
    return N, tab, U, G, Dw


def _apply(op, dt, tab, U, W, G, Dw, ncell, P, Q, mode=F.MODE_CORRECT):
    if op == "mass":
        F.mass_apply(dev(tab, dt), dev(U, dt), W, dev(Dw, dt), ncell, P, Q,
                     mode=mode)
    elif op == "stiffness":
        F.stiffness_apply(dev(tab, dt), dev(U, dt), W, dev(G, dt), ncell, P,
                          Q, mode=mode)
    else:
        F.kpm_apply(dev(tab, dt), dev(U, dt), W, dev(G, dt), dev(Dw, dt),
                    ncell, P, Q)


CELL_SHAPES = [(P, Q) for P in (9, 10, 13, 15, 17, 21, 25)
               for Q in (P + 1, P + 3)]


@pytest.mark.parametrize("P,Q", CELL_SHAPES)
@pytest.mark.parametrize("op", ["mass", "stiffness", "kpm"])
def test_cell_operator_f64_vs_oracle(op, P, Q, hd_kernel):
    ncell = 3
    print(f"general-degree cell {op} P={P} Q={Q}", flush=True)
    N, tab, U, G, Dw = _case(P, Q, ncell, 40 * P + Q, table="random")
    ref = O.cell_operator(op, N, Q, tab, U, G, Dw, ncell)
    W = torch.full((ncell * N ** 3,), 7.0, device="cuda", dtype=torch.float64)
    _apply(op, torch.float64, tab, U, W, G, Dw, ncell, P, Q)
    assert rel(host(W), ref) <= TOL64, (op, P, Q)


@pytest.mark.parametrize("P", [9, 16, 17, 19, 20, 22, 25])
@pytest.mark.parametrize("op", ["mass", "stiffness", "kpm"])
def test_cell_operator_f32_vs_oracle(op, P, hd_kernel):
    ncell, Q = 4, P + 1
    N, tab, U, G, Dw = _case(P, Q, ncell, 3 * P)
    ref = O.cell_operator(op, N, Q, tab, U, G, Dw, ncell)
    W = torch.zeros(ncell * N ** 3, device="cuda", dtype=torch.float32)
    _apply(op, torch.float32, tab, U, W, G, Dw, ncell, P, Q)
    err = rel(host(W), ref)
    print(f"FP32 general-degree {op} P={P}: {err:.2e} vs the FP64 oracle")
    assert err <= TOL32, (op, P)


_OPS = {"mass": F.OP_MASS, "stiffness": F.OP_STIFFNESS, "kpm": F.OP_KPM}
LINE_SHAPES = [(P, Q) for P in range(9, 26) for Q in (P + 1, P + 2, P + 3)]


@pytest.mark.parametrize("P,Q", LINE_SHAPES)
def test_line_kernels_every_shape_vs_oracle(P, Q):
    # every (degree, nq) the line kernels instantiate, all three operators,
    # FP64 against the oracle on random tables and geometry; 3 cells, so a
    # workgroup of two cells also runs with its second cell past the end
    ncell = 3
    print(f"line kernels P={P} Q={Q}", flush=True)
    N, tab, U, G, Dw = _case(P, Q, ncell, 11 * P + Q, table="random")
    for op in ("mass", "stiffness", "kpm"):
        for dt in (F.F64, F.F32):
            # the kernel each dtype runs is the one the query names (the
            # FP32 line runs in test_cell_operator_f32_vs_oracle)
            assert F.apply_kernel(_OPS[op], dt, P, Q) in ("line", "general")
        ref = O.cell_operator(op, N, Q, tab, U, G, Dw, ncell)
        W = torch.full((ncell * N ** 3,), -5.0, device="cuda",
                       dtype=torch.float64)
        _apply(op, torch.float64, tab, U, W, G, Dw, ncell, P, Q)
        assert rel(host(W), ref) <= TOL64, (op, P, Q)


@pytest.mark.parametrize("P,Q,ncell", [(9, 10, 16500), (16, 17, 701),
                                        (15, 16, 33001)])
@pytest.mark.parametrize("op", ["mass", "stiffness"])
def test_line_kernel_matches_general_kernel_at_size(op, P, Q, ncell):
    # many cells (16500 > the 16384-workgroup grid, and 33001 cells in
    # workgroups of two, so workgroups loop over cells; odd counts leave a
    # workgroup's second cell past the end): the line kernel against the
    # general kernel on the GPU, which differ only in the order of the sums
    N, tab, U, G, Dw = _case(P, Q, 1, 5 * P)
    U = O.uniform(91, ncell * N ** 3)
    dU, dtab = dev(U), dev(tab)
    if op == "mass":
        geo = torch.rand(ncell * Q ** 3, device="cuda", dtype=torch.float64)
        geo += 0.5
    else:
        geo = torch.rand(ncell * Q ** 3 * 6, device="cuda",
                         dtype=torch.float64)
    out = {}
    for mode in (F.HIGH_DEGREE_AUTO, F.HIGH_DEGREE_GENERAL):
        W = torch.empty(ncell * N ** 3, device="cuda", dtype=torch.float64)
        with F.high_degree_kernel(mode):
            if op == "mass":
                F.mass_apply(dtab, dU, W, geo, ncell, P, Q)
            else:
                F.stiffness_apply(dtab, dU, W, geo, ncell, P, Q)
        out[mode] = host(W)
    assert rel(out[F.HIGH_DEGREE_AUTO], out[F.HIGH_DEGREE_GENERAL]) <= TOL64


@pytest.mark.parametrize("P", [9, 12, 20, 25])
def test_reference_mode_vs_literal_oracle(P):
    N = P + 1
    ncell = 3
    phi2 = O.uniform(500 + P, 2 * N * N, 0, 1)
    U = O.uniform(501, ncell * N ** 3, 0, 1)
    D = O.uniform(502, ncell * N ** 3, 0, 1)
    G = O.uniform(503, ncell * N ** 3 * 6, 0, 1)
    W = torch.zeros(ncell * N ** 3, device="cuda", dtype=torch.float64)
    F.mass_apply(dev(phi2[: N * N]), dev(U), W, dev(D), ncell, P, N,
                 mode=F.MODE_REFERENCE)
    assert rel(host(W), O.mass_literal(P, phi2[: N * N], U, D, ncell)) <= TOL64
    F.stiffness_apply(dev(phi2), dev(U), W, dev(G), ncell, P, N,
                      mode=F.MODE_REFERENCE)
    assert rel(host(W), O.stiffness_literal(P, phi2, U, G, ncell)) <= TOL64


@pytest.mark.parametrize("P", [10, 25])
def test_host_shim_at_high_degree(P):
    # femtc_mass_operator / femtc_stiffness_operator: the reference
    # signatures (src/mass.hpp:196-199, src/stiffness.hpp:245-248), which
    # now run every degree of the reference's ladder
    N = P + 1
    ncell = 2
    phi2 = O.uniform(700 + P, 2 * N * N, 0, 1)
    U = O.uniform(701, ncell * N ** 3, 0, 1)
    D = O.uniform(702, ncell * N ** 3, 0, 1)
    G = O.uniform(703, ncell * N ** 3 * 6, 0, 1)
    W = np.zeros(ncell * N ** 3)
    F.mass_operator(phi2[: N * N].copy(), U.copy(), W, D.copy(), ncell, P,
                    F.Order(0), 8, 16)
    assert rel(W, O.mass_literal(P, phi2[: N * N], U, D, ncell)) <= TOL64
    W = np.zeros(ncell * N ** 3)
    F.stiffness_operator(phi2.copy(), U.copy(), W, G.copy(), ncell, P,
                         F.Order(3), 0, 0)
    assert rel(W, O.stiffness_literal(P, phi2, U, G, ncell)) <= TOL64


def _box(nx, ny, nz, P, Q, seed):
    C = nx * ny * nz
    J = O.random_jacobians(seed, C)
    G, Dw = O.geometry_affine(J, Q)
    dm = O.dofmap_structured(nx, ny, nz, P)
    ng = (nx * P + 1) * (ny * P + 1) * (nz * P + 1)
    return C, J, G, Dw, dm, ng


@pytest.mark.parametrize("nx,ny,nz,P,Q", [(2, 2, 2, 9, 10), (2, 1, 2, 11, 13),
                                          (1, 2, 1, 16, 17), (1, 1, 2, 24, 25)])
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_global_operator_vs_oracle(op, nx, ny, nz, P, Q, hd_kernel):
    C, J, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 60 + P)
    tab = O.basis_tables(P, Q)
    x = O.uniform(61, ng)
    ref = O.global_operator(op, P + 1, Q, tab, x, dm, G, Dw)
    y = torch.full((ng,), 3.0, device="cuda", dtype=torch.float64)
    ddm = dev(dm, torch.int32)
    F.global_apply(op, dev(tab), dev(x), y, ddm, C, P, Q, G=dev(G),
                   detJw=dev(Dw))
    assert rel(host(y), ref) <= TOL64
    # y += A x onto a non-zero y
    y0 = O.uniform(62, ng)
    y = dev(y0)
    F.global_apply(op, dev(tab), dev(x), y, ddm, C, P, Q, G=dev(G),
                   detJw=dev(Dw), accumulate=True)
    assert rel(host(y), ref + y0) <= TOL64
    # FP32: the gather and the FP32 atomic scatter
    f32 = torch.float32
    y = torch.full((ng,), 2.0, device="cuda", dtype=f32)
    F.global_apply(op, dev(tab, f32), dev(x, f32), y, ddm, C, P, Q,
                   G=dev(G, f32), detJw=dev(Dw, f32))
    assert rel(host(y), ref) <= TOL32


@pytest.mark.parametrize("P,Q", [(9, 10), (12, 14), (16, 17), (17, 18),
                                 (20, 22), (23, 25), (25, 26)])
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_affine_operator_vs_oracle(op, P, Q, hd_kernel):
    nx, ny, nz = 2, 1, 2
    C, J, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 70 + P)
    x = O.uniform(71, ng)
    ref = O.global_operator(op, P + 1, Q, O.basis_tables(P, Q), x, dm, G, Dw)
    tw = F.basis_tables_weights(P, Q)
    ddm, dJ = dev(dm, torch.int32), dev(J)
    for dt, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
        cg = torch.empty(C * 7, device="cuda", dtype=dt)
        F.geometry_cell(dJ, cg)
        y = torch.full((ng,), 4.0, device="cuda", dtype=dt)
        F.global_apply_affine(op, dev(tw, dt), dev(x, dt), y, ddm, C, P, Q, cg)
        assert rel(host(y), ref) <= tol, dt


def test_more_cells_than_workgroups():
    # 9000 cells: the grid (8192 workgroups) walks the cells in a loop
    P, Q, C = 9, 10, 9000
    N = P + 1
    tab = O.basis_tables(P, Q)
    U = O.uniform(81, C * N ** 3)
    Dw = O.uniform(82, C * Q ** 3, 0.5, 1.5)
    ref = O.cell_operator("mass", N, Q, tab, U, None, Dw, C)
    W = torch.empty(C * N ** 3, device="cuda", dtype=torch.float64)
    F.mass_apply(dev(tab), dev(U), W, dev(Dw), C, P, Q)
    assert rel(host(W), ref) <= TOL64


def test_fused_only_paths_reject_high_degrees():
    # plans, solvers and batches cover degrees 1..8 only
    n, P = 1, 9
    dm = O.dofmap_structured(n, n, n, P)
    ng = (n * P + 1) ** 3
    with pytest.raises(RuntimeError, match="degree not supported"):
        F.Plan(n, n, n, P, dev(dm, torch.int32), ng)
    tab = dev(O.basis_tables(P, P + 1))
    Dw = dev(O.uniform(1, (P + 1) ** 3, 0.5, 1.5))
    with pytest.raises(RuntimeError, match="degree not supported"):
        F.PCGSolver(F.OP_MASS, tab, dev(dm, torch.int32), 1, P, P + 1, ng,
                    detJw=Dw)
    assert F.lib().femtc_max_fused_degree() == 8
    assert F.lib().femtc_max_degree() == 25
    # one past the ladder
    W = torch.zeros(27 ** 3, device="cuda", dtype=torch.float64)
    with pytest.raises(RuntimeError, match="degree not supported"):
        F.mass_apply(dev(np.ones(27 * 27)), W, W, W, 1, 26, 27)
    assert _lib.ERR_DEGREE_NOT_SUPPORTED == 2
