"""GPU parity of the planned continuous operator (LDS-assembled blocked
scatter, assembled two-pass scatter) against the oracle, through the C ABI:
every candidate block shape, partial blocks, overwrite and accumulate, FP32,
and dofmaps that are not the structured numbering (where the blocked
scatter must fall back to atomics and stay correct, and the assembled
scatter is refused).  Bar: FP64 <= 1e-12 relative l2, FP32 <= 1e-5."""
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
    return C, G, Dw, dm, ng


def _shapes(Q):
    cpw = 64 // (Q * Q)
    base = {1: [(2, 2, 2), (1, 2, 2), (1, 2, 4), (2, 2, 1), (1, 1, 4), (2, 4, 1)],
            2: [(2, 2, 2), (2, 2, 4), (2, 4, 2), (1, 2, 4), (4, 2, 2)],
            4: [(2, 2, 4), (2, 4, 4), (4, 2, 4), (1, 2, 4), (2, 2, 8)],
            7: [(2, 2, 7), (2, 4, 7), (4, 2, 7), (1, 4, 7), (2, 2, 14)]}
    return base.get(cpw, [(2, 2, cpw), (2, 4, cpw), (4, 2, cpw), (1, 2, cpw)])


# (nx, ny, nz, P, Q): dims that are not multiples of the block shapes
MESHES = [(3, 5, 5, 1, 2), (3, 5, 9, 2, 3), (5, 3, 6, 3, 4), (3, 5, 5, 4, 5),
          (3, 3, 4, 5, 6), (3, 3, 3, 6, 7), (3, 2, 3, 7, 8), (3, 3, 5, 3, 5),
          (3, 3, 5, 2, 4)]


@pytest.mark.parametrize("nx,ny,nz,P,Q", MESHES)
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_plan_every_block_shape_vs_oracle(op, nx, ny, nz, P, Q):
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 70 + P)
    tab = O.basis_tables(P, Q)
    x = O.uniform(71, ng)
    ref = O.global_operator(op, P + 1, Q, tab, x, dm, G, Dw)
    ddm = dev(dm, torch.int32)
    plan = F.Plan(nx, ny, nz, P, ddm, ng)
    dtab, dx, dG, dD = dev(tab), dev(x), dev(G), dev(Dw)
    y0 = O.uniform(72, ng)
    for shape in _shapes(Q) + [(0, 0, 0)]:
        plan.set_block(Q, F.F64, op, shape)
        dy = torch.full((ng,), 5.0, device="cuda", dtype=torch.float64)
        plan.apply(op, dtab, dx, dy, Q, G=dG, detJw=dD)
        assert rel(host(dy), ref) <= 1e-12, shape
        dy = dev(y0)
        plan.apply(op, dtab, dx, dy, Q, G=dG, detJw=dD, accumulate=True)
        assert rel(host(dy), ref + y0) <= 1e-12, shape


@pytest.mark.parametrize("P,Q", [(2, 3), (4, 5), (6, 7)])
def test_plan_f32(P, Q):
    nx, ny, nz = 3, 4, 5
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 5)
    tab = O.basis_tables(P, Q)
    x = O.uniform(6, ng)
    ref = O.global_operator(F.OP_KPM, P + 1, Q, tab, x, dm, G, Dw)
    f = torch.float32
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    dy = torch.empty(ng, device="cuda", dtype=f)
    plan.apply(F.OP_KPM, dev(tab, f), dev(x, f), dy, Q, G=dev(G, f),
               detJw=dev(Dw, f))
    assert rel(host(dy), ref) <= 1e-5


@pytest.mark.parametrize("kind", ["relabelled dofs", "shuffled cells"])
@pytest.mark.parametrize("P,Q", [(2, 3), (4, 5), (6, 7)])
def test_plan_on_a_dofmap_that_is_not_structured(kind, P, Q):
    # the blocked scatter may only be an optimisation: a dofmap whose
    # numbering differs from the plan's must still give the right operator
    nx, ny, nz = 3, 4, 4
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 9)
    rng = np.random.default_rng(10)
    N3 = (P + 1) ** 3
    if kind == "relabelled dofs":
        perm = rng.permutation(ng).astype(np.int32)
        dm = perm[dm]
    else:
        cp = rng.permutation(C)
        dm = dm.reshape(C, N3)[cp].ravel()
        G = G.reshape(C, -1)[cp].ravel()
        Dw = Dw.reshape(C, -1)[cp].ravel()
    tab = O.basis_tables(P, Q)
    x = O.uniform(11, ng)
    ref = O.global_operator(F.OP_KPM, P + 1, Q, tab, x, dm, G, Dw)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    dy = torch.empty(ng, device="cuda", dtype=torch.float64)
    plan.apply(F.OP_KPM, dev(tab), dev(x), dy, Q, G=dev(G), detJw=dev(Dw))
    assert rel(host(dy), ref) <= 1e-12


# the block meshes plus Q > 8 (two-wave workgroups), a single cell, single
# rows along each axis and unequal dimensions
ASSEMBLED_MESHES = MESHES + [(2, 3, 3, 7, 9), (2, 2, 3, 8, 10),
                             (3, 2, 2, 8, 11), (4, 4, 4, 1, 4), (1, 1, 1, 3, 4),
                             (1, 1, 5, 2, 3), (5, 1, 1, 2, 4), (1, 4, 1, 3, 4),
                             (2, 3, 7, 4, 6)]


@pytest.mark.parametrize("nx,ny,nz,P,Q", ASSEMBLED_MESHES)
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_plan_assembled_vs_oracle(op, nx, ny, nz, P, Q):
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 80 + P)
    tab = O.basis_tables(P, Q)
    x = O.uniform(81, ng)
    ref = O.global_operator(op, P + 1, Q, tab, x, dm, G, Dw)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    dtab, dx, dG, dD = dev(tab), dev(x), dev(G), dev(Dw)
    plan.set_block(Q, F.F64, op, (-1, -1, -1))
    assert plan.get_block(Q, F.F64, op) == (-1, -1, -1)
    dy = torch.full((ng,), 5.0, device="cuda", dtype=torch.float64)
    plan.apply(op, dtab, dx, dy, Q, G=dG, detJw=dD)
    assert rel(host(dy), ref) <= 1e-12
    y0 = O.uniform(82, ng)
    dy = dev(y0)
    plan.apply(op, dtab, dx, dy, Q, G=dG, detJw=dD, accumulate=True)
    assert rel(host(dy), ref + y0) <= 1e-12


STRUCTURED_CASES = (
    [(m, (-2, -2, -2)) for m in ASSEMBLED_MESHES] +
    [(m, (-3, -3, -3)) for m in ASSEMBLED_MESHES + [(2, 2, 9, 4, 5),
                                                    (2, 3, 7, 6, 7),
                                                    (2, 2, 17, 1, 2),
                                                    (1, 2, 12, 2, 3)]
     if m[4] <= 8])


@pytest.mark.parametrize("mesh,shape", STRUCTURED_CASES)
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_plan_structured_scatters_vs_oracle(op, mesh, shape):
    # -2: the atomic scatter with closed-form gather and scatter indices;
    # -3: closed-form gather, cells summed along z in LDS, atomics on runs
    # of consecutive dofs (segments shorter at the end of a z-line)
    nx, ny, nz, P, Q = mesh
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 85 + P)
    tab = O.basis_tables(P, Q)
    x = O.uniform(86, ng)
    ref = O.global_operator(op, P + 1, Q, tab, x, dm, G, Dw)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    dtab, dx, dG, dD = dev(tab), dev(x), dev(G), dev(Dw)
    plan.set_block(Q, F.F64, op, shape)
    assert plan.get_block(Q, F.F64, op) == shape
    dy = torch.full((ng,), 5.0, device="cuda", dtype=torch.float64)
    plan.apply(op, dtab, dx, dy, Q, G=dG, detJw=dD)
    assert rel(host(dy), ref) <= 1e-12
    y0 = O.uniform(87, ng)
    dy = dev(y0)
    plan.apply(op, dtab, dx, dy, Q, G=dG, detJw=dD, accumulate=True)
    assert rel(host(dy), ref + y0) <= 1e-12
    f = torch.float32
    plan.set_block(Q, F.F32, op, shape)
    dy = torch.empty(ng, device="cuda", dtype=f)
    plan.apply(op, dev(tab, f), dev(x, f), dy, Q, G=dev(G, f),
               detJw=dev(Dw, f))
    assert rel(host(dy), ref) <= 1e-5


@pytest.mark.parametrize("P,Q", [(2, 3), (4, 5), (6, 7), (8, 9)])
def test_plan_assembled_f32(P, Q):
    nx, ny, nz = 3, 4, 5
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 15)
    tab = O.basis_tables(P, Q)
    x = O.uniform(16, ng)
    ref = O.global_operator(F.OP_KPM, P + 1, Q, tab, x, dm, G, Dw)
    f = torch.float32
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    plan.set_block(Q, F.F32, F.OP_KPM, (-1, -1, -1))
    dy = torch.empty(ng, device="cuda", dtype=f)
    plan.apply(F.OP_KPM, dev(tab, f), dev(x, f), dy, Q, G=dev(G, f),
               detJw=dev(Dw, f))
    assert rel(host(dy), ref) <= 1e-5


@pytest.mark.parametrize("kind", ["relabelled dofs", "shuffled cells",
                                  "one entry changed", "extra dof"])
def test_plan_assembled_only_for_the_structured_numbering(kind):
    # any other dofmap is refused for the assembled scatter (and still
    # applies correctly through the other scatters)
    nx, ny, nz, P, Q = 3, 4, 4, 2, 3
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 17)
    rng = np.random.default_rng(18)
    N3 = (P + 1) ** 3
    if kind == "relabelled dofs":
        dm = rng.permutation(ng).astype(np.int32)[dm]
    elif kind == "shuffled cells":
        cp = rng.permutation(C)
        dm = dm.reshape(C, N3)[cp].ravel()
        G = G.reshape(C, -1)[cp].ravel()
        Dw = Dw.reshape(C, -1)[cp].ravel()
    elif kind == "one entry changed":
        dm = dm.copy()
        dm[2 * N3 + N3 // 2] = dm[N3 // 2]
    else:
        ng += 1  # the structured numbering, one unused dof more
    tab = O.basis_tables(P, Q)
    x = O.uniform(19, ng)
    ref = O.global_operator(F.OP_KPM, P + 1, Q, tab, x, dm, G, Dw)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    with pytest.raises(F.FemtcError, match="structured"):
        plan.set_block(Q, F.F64, F.OP_KPM, (-1, -1, -1))
    with pytest.raises(F.FemtcError, match="structured"):
        plan.set_block(Q, F.F64, F.OP_KPM, (-2, -2, -2))
    with pytest.raises(F.FemtcError, match="structured"):
        plan.set_block(Q, F.F64, F.OP_KPM, (-3, -3, -3))
    dy = torch.empty(ng, device="cuda", dtype=torch.float64)
    shape, _ = plan.autotune(F.OP_KPM, dev(tab), dev(x), dy, Q, G=dev(G),
                             detJw=dev(Dw), reps=1)
    assert shape not in ((-1, -1, -1), (-2, -2, -2), (-3, -3, -3))
    plan.apply(F.OP_KPM, dev(tab), dev(x), dy, Q, G=dev(G), detJw=dev(Dw))
    assert rel(host(dy), ref) <= 1e-12


def test_plan_autotune_picks_a_valid_shape_and_clears_y():
    nx, ny, nz, P, Q = 8, 8, 8, 4, 5
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 12)
    tab = O.basis_tables(P, Q)
    x = O.uniform(13, ng)
    ref = O.global_operator(F.OP_STIFFNESS, P + 1, Q, tab, x, dm, G, None)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    dtab, dx, dG = dev(tab), dev(x), dev(G)
    dy = torch.full((ng,), 1.0, device="cuda", dtype=torch.float64)
    shape, us = plan.autotune(F.OP_STIFFNESS, dtab, dx, dy, Q, G=dG, reps=2)
    assert us > 0
    assert float(dy.abs().max()) == 0.0
    # the plan's dofmap is the structured one, so every scatter may win:
    # blocked shapes, direct, assembled, structured direct and z-run
    assert shape in _shapes(Q) + [(0, 0, 0), (-1, -1, -1), (-2, -2, -2),
                                  (-3, -3, -3)]
    plan.apply(F.OP_STIFFNESS, dtab, dx, dy, Q, G=dG)
    assert rel(host(dy), ref) <= 1e-12


def test_plan_rejects_bad_input():
    P, Q = 2, 3
    C, G, Dw, dm, ng = _box(2, 2, 2, P, Q, 1)
    bad = dm.copy()
    bad[5] = ng  # out of range
    with pytest.raises(F.FemtcError, match="outside"):
        F.Plan(2, 2, 2, P, dev(bad, torch.int32), ng)
    plan = F.Plan(2, 2, 2, P, dev(dm, torch.int32), ng)
    with pytest.raises(F.FemtcError, match="multiple"):
        plan.set_block(Q, F.F64, F.OP_MASS, (2, 2, 3))  # cbz % 7 != 0


# (nx, ny, nz, P, Q = P + 1): even and odd N, every degree
EO_MESHES = [(3, 5, 5, 1, 2), (3, 5, 9, 2, 3), (5, 3, 6, 3, 4), (3, 5, 5, 4, 5),
             (3, 3, 4, 5, 6), (3, 3, 3, 6, 7), (3, 2, 3, 7, 8), (2, 2, 3, 8, 9)]


@pytest.mark.parametrize("nx,ny,nz,P,Q", EO_MESHES)
@pytest.mark.parametrize("op", [F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM])
def test_plan_even_odd_vs_oracle(op, nx, ny, nz, P, Q):
    # registered symmetric tables: the even-odd contractions on the direct
    # and the assembled scatter, FP64 and FP32, overwrite and accumulate,
    # against the oracle and against the general form on the same plan
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 150 + P)
    tab = O.basis_tables(P, Q)
    x = O.uniform(151, ng)
    y0 = O.uniform(152, ng)
    ref = O.global_operator(op, P + 1, Q, tab, x, dm, G, Dw)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    for dt, code, tol, tol_gen in ((torch.float64, F.F64, 1e-12, 1e-13),
                                   (torch.float32, F.F32, 1e-5, 2e-6)):
        assert plan.set_tables(dev(tab, dt), Q) is True
        dtab, dx, dG, dD = dev(tab, dt), dev(x, dt), dev(G, dt), dev(Dw, dt)
        for shape in ((0, 0, 0), (-1, -1, -1)):
            plan.set_block(Q, code, op, shape)
            dy = torch.full((ng,), 5.0, device="cuda", dtype=dt)
            plan.apply(op, None, dx, dy, Q, G=dG, detJw=dD)
            assert rel(host(dy), ref) <= tol, (dt, shape)
            dg = torch.full((ng,), 5.0, device="cuda", dtype=dt)
            plan.apply(op, dtab, dx, dg, Q, G=dG, detJw=dD)
            assert rel(host(dy), host(dg)) <= tol_gen, (dt, shape)
            dy = dev(y0, dt)
            plan.apply(op, None, dx, dy, Q, G=dG, detJw=dD, accumulate=True)
            assert rel(host(dy), ref + y0) <= tol, (dt, shape)


def test_plan_registered_tables_without_symmetry_use_the_general_form():
    nx, ny, nz, P = 3, 3, 4, 4
    plan_dm = O.dofmap_structured(nx, ny, nz, P)
    ng = (nx * P + 1) * (ny * P + 1) * (nz * P + 1)
    plan = F.Plan(nx, ny, nz, P, dev(plan_dm, torch.int32), ng)
    x = O.uniform(161, ng)
    # unregistered (dtype, nq): tables=None is refused
    dy = torch.empty(ng, device="cuda", dtype=torch.float64)
    with pytest.raises(F.FemtcError, match="registered"):
        plan.apply(F.OP_MASS, None, dev(x), dy, P + 1, detJw=dev(np.ones(1)))
    for Q, perturb in ((P + 1, True), (P + 2, False)):
        C, G, Dw, dm, _ = _box(nx, ny, nz, P, Q, 160)
        tab = O.basis_tables(P, Q)
        if perturb:
            tab = tab.copy()
            tab[1] *= 1.0 + 1e-9  # breaks phi's symmetry by 1e-9
        ref = O.global_operator(F.OP_KPM, P + 1, Q, tab, x, dm, G, Dw)
        assert plan.set_tables(dev(tab), Q) is False
        for shape in ((0, 0, 0), (-1, -1, -1)):
            plan.set_block(Q, F.F64, F.OP_KPM, shape)
            plan.apply(F.OP_KPM, None, dev(x), dy, Q, G=dev(G), detJw=dev(Dw))
            assert rel(host(dy), ref) <= 1e-12, (Q, shape)
    # registering again replaces the copy (now symmetric: even-odd)
    Q = P + 1
    C, G, Dw, dm, _ = _box(nx, ny, nz, P, Q, 160)
    tab = O.basis_tables(P, Q)
    assert plan.set_tables(dev(tab), Q) is True
    ref = O.global_operator(F.OP_KPM, P + 1, Q, tab, x, dm, G, Dw)
    plan.apply(F.OP_KPM, None, dev(x), dy, Q, G=dev(G), detJw=dev(Dw))
    assert rel(host(dy), ref) <= 1e-12


def test_plan_autotune_with_registered_tables():
    nx, ny, nz, P, Q = 6, 6, 6, 4, 5
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, Q, 170)
    tab = O.basis_tables(P, Q)
    x = O.uniform(171, ng)
    ref = O.global_operator(F.OP_STIFFNESS, P + 1, Q, tab, x, dm, G, None)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    assert plan.set_tables(dev(tab), Q) is True
    dx, dG = dev(x), dev(G)
    dy = torch.full((ng,), 1.0, device="cuda", dtype=torch.float64)
    shape, us = plan.autotune(F.OP_STIFFNESS, None, dx, dy, Q, G=dG, reps=2)
    assert us > 0 and float(dy.abs().max()) == 0.0
    plan.apply(F.OP_STIFFNESS, None, dx, dy, Q, G=dG)
    assert rel(host(dy), ref) <= 1e-12, shape


def test_plan_zrun_scatter_limits():
    # the z-run scatter holds one-wave cell groups: nq <= 8 only
    nx, ny, nz, P = 2, 2, 2, 8
    C, G, Dw, dm, ng = _box(nx, ny, nz, P, 9, 31)
    plan = F.Plan(nx, ny, nz, P, dev(dm, torch.int32), ng)
    with pytest.raises(F.FemtcError, match="nq <= 8"):
        plan.set_block(9, F.F64, F.OP_MASS, (-3, -3, -3))
    plan.set_block(9, F.F64, F.OP_MASS, (-2, -2, -2))
