"""GPU parity: the HIP kernels through the C ABI against the CPU oracle.

Bar (FP64): relative l2 error <= 1e-12 against the oracle's correct mode,
and against the reference's own outputs (tests/golden) in reference mode.
Dofmap indices are compared bit for bit.  FP32 results are checked against
the FP64 oracle at 1e-5 relative l2.
"""
import os

import numpy as np
import pytest

import femtc_amd as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "ref_goldens.npz"))
TOL64 = 1e-12
TOL32 = 1e-5


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


SHAPES = [(P, Q) for P in range(1, 9) for Q in (P + 1, P + 2, P + 3)]


@pytest.mark.parametrize("P,Q", SHAPES)
@pytest.mark.parametrize("op", ["mass", "stiffness", "kpm"])
def test_cell_operator_f64_vs_oracle(op, P, Q):
    ncell = 37  # not a multiple of any cells-per-block
    N, tab, U, G, Dw = _case(P, Q, ncell, 100 * P + Q, table="random")
    ref = O.cell_operator(op, N, Q, tab, U, G, Dw, ncell)
    dU, dW = dev(U), torch.full((ncell * N ** 3,), 7.0, device="cuda",
                                dtype=torch.float64)
    dt = dev(tab)
    if op == "mass":
        F.mass_apply(dt, dU, dW, dev(Dw), ncell, P, Q)
    elif op == "stiffness":
        F.stiffness_apply(dt, dU, dW, dev(G), ncell, P, Q)
    else:
        F.kpm_apply(dt, dU, dW, dev(G), dev(Dw), ncell, P, Q)
    assert rel(host(dW), ref) <= TOL64, (op, P, Q)


@pytest.mark.parametrize("P,Q", [(1, 2), (3, 4), (3, 5), (4, 5), (6, 7), (8, 9)])
@pytest.mark.parametrize("op", ["mass", "stiffness", "kpm"])
def test_cell_operator_f32_vs_oracle(op, P, Q):
    ncell = 50
    N, tab, U, G, Dw = _case(P, Q, ncell, 7 * P + Q)
    ref = O.cell_operator(op, N, Q, tab, U, G, Dw, ncell)
    f = torch.float32
    dU, dW = dev(U, f), torch.zeros(ncell * N ** 3, device="cuda", dtype=f)
    if op == "mass":
        F.mass_apply(dev(tab, f), dU, dW, dev(Dw, f), ncell, P, Q)
    elif op == "stiffness":
        F.stiffness_apply(dev(tab, f), dU, dW, dev(G, f), ncell, P, Q)
    else:
        F.kpm_apply(dev(tab, f), dU, dW, dev(G, f), dev(Dw, f), ncell, P, Q)
    assert rel(host(dW), ref) <= TOL32, (op, P, Q)


def test_reference_mode_mass_matches_reference_goldens():
    # operators::mass_operator<double> outputs, src/mass.hpp:10-49
    for k in range(int(GOLD["mass_count"][0])):
        P, sym, order, Mb, Nb, ncell = (int(v) for v in GOLD[f"mass_{k}_meta"])
        dW = torch.zeros(ncell * (P + 1) ** 3, device="cuda", dtype=torch.float64)
        F.mass_apply(dev(GOLD[f"mass_{k}_phi"]), dev(GOLD[f"mass_{k}_U"]), dW,
                     dev(GOLD[f"mass_{k}_detJ"]), ncell, P, P + 1,
                     mode=F.MODE_REFERENCE)
        assert rel(host(dW), GOLD[f"mass_{k}_W"]) <= TOL64, (P, sym)


def test_reference_mode_stiffness_matches_reference_goldens():
    # operators::stiffness_operator<double> outputs, src/stiffness.hpp:34-99
    for k in range(int(GOLD["stiff_count"][0])):
        P, order, Mb, Nb, ncell = (int(v) for v in GOLD[f"stiff_{k}_meta"])
        dW = torch.zeros(ncell * (P + 1) ** 3, device="cuda", dtype=torch.float64)
        F.stiffness_apply(dev(GOLD[f"stiff_{k}_phi2"]), dev(GOLD[f"stiff_{k}_U"]),
                          dW, dev(GOLD[f"stiff_{k}_G"]), ncell, P, P + 1,
                          mode=F.MODE_REFERENCE)
        assert rel(host(dW), GOLD[f"stiff_{k}_W"]) <= TOL64, P


@pytest.mark.parametrize("P", [1, 2, 3, 5, 6, 7, 8])
def test_reference_mode_vs_literal_oracle_more_degrees(P):
    N = P + 1
    ncell = 19
    phi2 = O.uniform(500 + P, 2 * N * N, 0, 1)
    U = O.uniform(501, ncell * N ** 3, 0, 1)
    D = O.uniform(502, ncell * N ** 3, 0, 1)
    G = O.uniform(503, ncell * N ** 3 * 6, 0, 1)
    dW = torch.zeros(ncell * N ** 3, device="cuda", dtype=torch.float64)
    F.mass_apply(dev(phi2[: N * N]), dev(U), dW, dev(D), ncell, P, N,
                 mode=F.MODE_REFERENCE)
    assert rel(host(dW), O.mass_literal(P, phi2[: N * N], U, D, ncell)) <= TOL64
    F.stiffness_apply(dev(phi2), dev(U), dW, dev(G), ncell, P, N,
                      mode=F.MODE_REFERENCE)
    assert rel(host(dW), O.stiffness_literal(P, phi2, U, G, ncell)) <= TOL64


def test_host_shim_entry_points_match_goldens():
    # the synchronous reference-signature path (femtc_mass_operator)
    for k in (0, 5, 11, 23):
        P, sym, order, Mb, Nb, ncell = (int(v) for v in GOLD[f"mass_{k}_meta"])
        W = np.zeros(ncell * (P + 1) ** 3)
        F.mass_operator(GOLD[f"mass_{k}_phi"].copy(), GOLD[f"mass_{k}_U"].copy(),
                        W, GOLD[f"mass_{k}_detJ"].copy(), ncell, P,
                        F.Order(order), Mb, Nb)
        assert rel(W, GOLD[f"mass_{k}_W"]) <= TOL64
    for k in range(int(GOLD["stiff_count"][0])):
        P, order, Mb, Nb, ncell = (int(v) for v in GOLD[f"stiff_{k}_meta"])
        W = np.zeros(ncell * (P + 1) ** 3)
        F.stiffness_operator(GOLD[f"stiff_{k}_phi2"].copy(),
                             GOLD[f"stiff_{k}_U"].copy(), W,
                             GOLD[f"stiff_{k}_G"].copy(), ncell, P,
                             F.Order(order), Mb, Nb)
        assert rel(W, GOLD[f"stiff_{k}_W"]) <= TOL64


def test_host_shim_unsupported_arguments_print_and_leave_w(capfd):
    for (deg, order, Mb, Nb), msg in zip(GOLD["ladder_cases"], GOLD["ladder_msgs"]):
        N = max(int(deg), 1) + 1
        W = np.full(N ** 3, 7.0)
        F.mass_operator(np.full(N * N, 0.5), np.ones(N ** 3), W, np.ones(N ** 3),
                        1, int(deg), int(order), int(Mb), int(Nb))
        assert np.all(W == 7.0)
        assert capfd.readouterr().out.strip() == str(msg)


def test_mass_cell_local_32768_cells_p3_q5():
    # 32768 cells, P=3 GLL, Q=5 Gauss, U~U[-1,1], detJw~U[0.5,1.5]
    P, Q, C = 3, 5, 32768
    N = P + 1
    tab = O.basis_tables(P, Q)
    U = O.uniform(1, C * N ** 3)
    Dw = O.uniform(2, C * Q ** 3, 0.5, 1.5)
    ref = O.cell_operator("mass", N, Q, tab, U, None, Dw, C)
    dW = torch.empty(C * N ** 3, device="cuda", dtype=torch.float64)
    F.mass_apply(dev(tab), dev(U), dW, dev(Dw), C, P, Q)
    assert rel(host(dW), ref) <= TOL64


def _mesh(n, P, Q, seed):
    C = n ** 3
    J = O.random_jacobians(seed, C)
    G, Dw = O.geometry_affine(J, Q)
    dm = O.dofmap_structured(n, n, n, P)
    ng = (n * P + 1) ** 3
    return C, J, G, Dw, dm, ng


def test_dofmap_bit_exact():
    for (nx, ny, nz, P) in ((3, 2, 4, 2), (64, 64, 64, 4), (48, 48, 48, 6)):
        N = P + 1
        ref = O.dofmap_structured(nx, ny, nz, P)
        d = torch.empty(ref.size, device="cuda", dtype=torch.int32)
        F.dofmap_structured(nx, ny, nz, P, d)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d.cpu().numpy(), ref)


@pytest.mark.parametrize("Q", [2, 5, 7, 9])
def test_geometry_affine(Q):
    C = 200
    J = O.random_jacobians(11, C)
    G, Dw = O.geometry_affine(J, Q)
    dG = torch.empty(C * Q ** 3 * 6, device="cuda", dtype=torch.float64)
    dD = torch.empty(C * Q ** 3, device="cuda", dtype=torch.float64)
    F.geometry_affine(dev(J), Q, dG, dD)
    assert rel(host(dG), G) <= 1e-14
    assert rel(host(dD), Dw) <= 1e-14


@pytest.mark.parametrize("n,P,Q", [(3, 1, 2), (4, 2, 3), (4, 4, 5), (3, 6, 7),
                                   (2, 8, 9), (3, 3, 5)])
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_global_operator_vs_oracle(op, n, P, Q):
    C, J, G, Dw, dm, ng = _mesh(n, P, Q, 20 + P)
    N = P + 1
    tab = O.basis_tables(P, Q)
    x = O.uniform(21, ng)
    ref = O.global_operator(op, N, Q, tab, x, dm, G, Dw)
    dy = torch.full((ng,), 3.0, device="cuda", dtype=torch.float64)
    F.global_apply(op, dev(tab), dev(x), dy, dev(dm, torch.int32), C, P, Q,
                   G=dev(G), detJw=dev(Dw))
    assert rel(host(dy), ref) <= TOL64


def test_stiffness_global_64cube_p4_q5():
    # 64^3 mesh, P=4, Q=5, J = I + 0.1 U(-1,1) per cell, x ~ U[-1,1]
    n, P, Q = 64, 4, 5
    C, J, G, Dw, dm, ng = _mesh(n, P, Q, 3)
    tab = O.basis_tables(P, Q)
    x = O.uniform(4, ng)
    ref = O.global_operator(F.OP_STIFFNESS, P + 1, Q, tab, x, dm, G, None)
    dy = torch.empty(ng, device="cuda", dtype=torch.float64)
    F.global_apply(F.OP_STIFFNESS, dev(tab), dev(x), dy, dev(dm, torch.int32),
                   C, P, Q, G=dev(G))
    assert rel(host(dy), ref) <= TOL64


def _box(nx, ny, nz, P, Q, seed):
    C = nx * ny * nz
    J = O.random_jacobians(seed, C)
    G, Dw = O.geometry_affine(J, Q)
    dm = O.dofmap_structured(nx, ny, nz, P)
    ng = (nx * P + 1) * (ny * P + 1) * (nz * P + 1)
    return C, G, Dw, dm, ng


BOXES = [(2, 3, 8, 4, 5), (3, 2, 4, 6, 7), (2, 2, 8, 3, 5), (2, 2, 5, 2, 3),
         (2, 2, 4, 7, 8), (2, 1, 4, 8, 9), (1, 1, 32, 1, 2)]


@pytest.mark.parametrize("nx,ny,nz,P,Q", BOXES)
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_global_overwrite_and_accumulate(op, nx, ny, nz, P, Q):
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 40 + P)
    tab = O.basis_tables(P, Q)
    x = O.uniform(41, ng)
    ref = O.global_operator(op, P + 1, Q, tab, x, dm, G, Dw)
    dtab, dx, ddm = dev(tab), dev(x), dev(dm, torch.int32)
    dG, dD = dev(G), dev(Dw)
    # overwrite: prior content of y is ignored
    dy = torch.full((ng,), 5.0, device="cuda", dtype=torch.float64)
    F.global_apply(op, dtab, dx, dy, ddm, C, P, Q, G=dG, detJw=dD)
    assert rel(host(dy), ref) <= TOL64
    # accumulate onto a non-zero y
    y0 = O.uniform(42, ng)
    dy = dev(y0)
    F.global_apply(op, dtab, dx, dy, ddm, C, P, Q, G=dG, detJw=dD,
                   accumulate=True)
    assert rel(host(dy), ref + y0) <= TOL64
    # the library's fill kernel then accumulate
    F.fill_zero(dy)
    F.global_apply(op, dtab, dx, dy, ddm, C, P, Q, G=dG, detJw=dD,
                   accumulate=True)
    assert rel(host(dy), ref) <= TOL64


@pytest.mark.parametrize("nx,ny,nz,P,Q", [(2, 3, 8, 4, 5), (3, 2, 4, 6, 7)])
def test_global_kpm_f32(nx, ny, nz, P, Q):
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 7)
    tab = O.basis_tables(P, Q)
    x = O.uniform(8, ng)
    ref = O.global_operator(F.OP_KPM, P + 1, Q, tab, x, dm, G, Dw)
    f = torch.float32
    dy = torch.empty(ng, device="cuda", dtype=f)
    F.global_apply(F.OP_KPM, dev(tab, f), dev(x, f), dy, dev(dm, torch.int32),
                   C, P, Q, G=dev(G, f), detJw=dev(Dw, f))
    assert rel(host(dy), ref) <= TOL32


def test_stiffness_global_48cube_p6_q7():
    # 48^3 mesh, P=6 GLL, Q=7 Gauss: the benchmark workload
    n, P, Q = 48, 6, 7
    C, G, Dw, dm, ng = _box(n, n, n, P, Q, 3)
    tab = O.basis_tables(P, Q)
    x = O.uniform(4, ng)
    ref = O.global_operator(F.OP_STIFFNESS, P + 1, Q, tab, x, dm, G, None)
    dy = torch.empty(ng, device="cuda", dtype=torch.float64)
    F.global_apply(F.OP_STIFFNESS, dev(tab), dev(x), dy, dev(dm, torch.int32),
                   C, P, Q, G=dev(G))
    assert rel(host(dy), ref) <= TOL64


def test_misaligned_geometry_is_rejected():
    P, Q, C = 2, 3, 4
    N = P + 1
    buf = torch.zeros(C * Q ** 3 * 6 + 1, device="cuda", dtype=torch.float64)
    U = torch.zeros(C * N ** 3, device="cuda", dtype=torch.float64)
    W = torch.zeros_like(U)
    tab = dev(O.basis_tables(P, Q))
    with pytest.raises(F.FemtcError, match="aligned"):
        F.stiffness_apply(tab, U, W, buf[1:].data_ptr(), C, P, Q)
