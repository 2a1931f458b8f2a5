"""CPU-side checks of the C ABI: the library loads, exports every function
include/femtc.h declares, and its host-only logic (tables, string2order,
ladder validation) agrees with the oracle and the reference fixtures.
No GPU compute is called here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import femtc_amd as F
from femtc_amd import _lib
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "ref_goldens.npz"))


def _declared():
    text = open(os.path.join(ROOT, "include", "femtc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(femtc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (femtc_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = F.lib()
    for n in names:
        assert hasattr(L, n)
    # the Python binding table covers the header exactly
    assert sorted(_lib.SIGNATURES) == names


def test_status_strings_and_version():
    L = F.lib()
    assert L.femtc_version() == 1
    assert L.femtc_max_degree() == 25
    assert L.femtc_max_fused_degree() == 8
    assert _lib.status_string(_lib.ERR_DEGREE_NOT_SUPPORTED) == "degree not supported"
    assert _lib.status_string(_lib.ERR_BLOCK_NOT_SUPPORTED) == "Block size not supported"
    assert _lib.status_string(_lib.ERR_INVALID_ORDER_STRING) == "Invalid loop order"


def test_string2order_matches_reference():
    # reference codes from src/linalg.hpp:21-37 (fixture), -1 = throw
    for name, code in zip(GOLD["s2o_names"], GOLD["s2o_codes"]):
        if code < 0:
            with pytest.raises(RuntimeError, match="Invalid loop order"):
                F.string2order(str(name))
        else:
            assert F.string2order(str(name)) == code
            assert F.Order(code).name == str(name)


def test_ladder_validation_matches_reference_messages():
    L = F.lib()
    for (deg, order, Mb, Nb), msg in zip(GOLD["ladder_cases"], GOLD["ladder_msgs"]):
        st = L.femtc_check_ladder(int(deg), int(order), int(Mb), int(Nb))
        assert _lib.status_string(st) == str(msg)
    assert L.femtc_check_ladder(3, 4, 8, 16) == 0
    assert L.femtc_check_ladder(3, 6, 0, 0) == _lib.ERR_ORDER_NOT_SUPPORTED
    # the reference instantiates degrees 1..25 (src/mass.hpp:52-138)
    for deg in (9, 17, 25):
        assert L.femtc_check_ladder(deg, 0, 0, 0) == 0
    assert L.femtc_check_ladder(26, 0, 0, 0) == _lib.ERR_DEGREE_NOT_SUPPORTED
    assert L.femtc_check_ladder(0, 0, 0, 0) == _lib.ERR_DEGREE_NOT_SUPPORTED


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5, 6, 7, 8])
def test_tables_match_oracle(P):
    np.testing.assert_allclose(F.gll_points(P), O.gll_points(P), atol=2e-16)
    for Q in (P + 1, P + 2, P + 3):
        x, w = F.gauss_rule(Q)
        xo, wo = O.gl_rule(Q)
        np.testing.assert_allclose(x, xo, atol=3e-16)
        np.testing.assert_allclose(w, wo, rtol=1e-14)
        np.testing.assert_allclose(F.basis_tables(P, Q), O.basis_tables(P, Q),
                                   rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("P", [9, 12, 17, 25])
def test_tables_match_oracle_high_degree(P):
    # the general-degree kernels' tables: GLL nodes from two independent
    # root finders, Lagrange values and derivatives at Gauss points
    np.testing.assert_allclose(F.gll_points(P), O.gll_points(P), atol=1e-14)
    for Q in (P + 1, P + 3):
        t, o = F.basis_tables(P, Q), O.basis_tables(P, Q)
        N = P + 1
        # partition of unity and zero-sum derivatives (exact for any nodes)
        np.testing.assert_allclose(t[:N * Q].reshape(N, Q).sum(0), 1.0,
                                   atol=1e-12)
        np.testing.assert_allclose(t[N * Q:].reshape(N, Q).sum(0), 0.0,
                                   atol=1e-9)
        scale = np.abs(o).max()
        assert np.abs(t - o).max() <= 1e-11 * scale


def test_argument_errors_without_device_work():
    L = F.lib()
    # unsupported degree / shape are rejected before any HIP call
    assert L.femtc_mass_apply(0, 26, 27, 0, None, None, None, None, 10, None) \
        == _lib.ERR_DEGREE_NOT_SUPPORTED
    assert L.femtc_mass_apply(0, 25, 26, 0, None, None, None, None, 10, None) \
        == _lib.ERR_INVALID_ARGUMENT  # degree 25 runs; the buffers are null
    assert L.femtc_mass_apply(0, 3, 3, 0, None, None, None, None, 10, None) \
        == _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_stiffness_apply(0, 3, 5, 1, None, None, None, None, 10, None) \
        == _lib.ERR_INVALID_ARGUMENT  # reference mode needs nq = P+1
    assert L.femtc_mass_apply(0, 3, 4, 0, None, None, None, None, 0, None) == 0
    # the stiffness shim needs [phi | dphi]; bench_stiffness.cpp:74 passes N^2
    phi = np.zeros(16)
    U = np.zeros(64)
    W = np.zeros(64)
    G = np.zeros(64 * 6)
    st = L.femtc_stiffness_operator(0, phi.ctypes.data, 16, U.ctypes.data,
                                    W.ctypes.data, G.ctypes.data, 1, 3, 0, 0, 0, 1)
    assert st == _lib.ERR_INVALID_ARGUMENT
    assert "dphi" in _lib.last_error()


def test_high_degree_kernel_switch_without_device_work():
    # process-wide choice between the line kernels and the general kernel
    assert F.set_high_degree_kernel(F.HIGH_DEGREE_GENERAL) == F.HIGH_DEGREE_AUTO
    with F.high_degree_kernel(F.HIGH_DEGREE_AUTO):
        assert F.set_high_degree_kernel(F.HIGH_DEGREE_AUTO) == \
            F.HIGH_DEGREE_AUTO
    # restored on leaving the block
    assert F.set_high_degree_kernel(F.HIGH_DEGREE_AUTO) == \
        F.HIGH_DEGREE_GENERAL
    prev = _lib.C.c_int32(-7)
    assert F.lib().femtc_set_high_degree_kernel(2, _lib.C.byref(prev)) == \
        _lib.ERR_INVALID_ARGUMENT
    assert prev.value == -7 and "FEMTC_HIGH_DEGREE_AUTO" in _lib.last_error()
    assert F.lib().femtc_set_high_degree_kernel(F.HIGH_DEGREE_AUTO, None) == 0


def test_apply_kernel_query_without_device_work():
    K = F.apply_kernel
    assert K(F.OP_STIFFNESS, F.F64, 6, 7) == "fused"
    assert K(F.OP_STIFFNESS, F.F64, 16, 17) == "line"
    assert K(F.OP_STIFFNESS, F.F64, 16, 17, F.MODE_REFERENCE) == "general"
    with F.high_degree_kernel(F.HIGH_DEGREE_GENERAL):
        assert K(F.OP_STIFFNESS, F.F64, 16, 17) == "general"
    # FP64 stiffness: N = 21, 22 (Q = N or N+1) fit, larger cells do not
    assert K(F.OP_STIFFNESS, F.F64, 21, 22) == "line"
    assert K(F.OP_STIFFNESS, F.F64, 21, 23) == "general"
    assert K(F.OP_STIFFNESS, F.F64, 25, 26) == "general"
    # mass and FP32 reach the top of the ladder
    assert K(F.OP_MASS, F.F64, 25, 26) == "line"
    assert K(F.OP_STIFFNESS, F.F32, 25, 26) == "line"
    assert K(F.OP_STIFFNESS, F.F32, 25, 28) == "general"
    for P in range(9, 26):
        for dq in (1, 2, 3):
            for op in (F.OP_MASS, F.OP_STIFFNESS, F.OP_KPM):
                for dt in (F.F64, F.F32):
                    assert K(op, dt, P, P + dq) in ("line", "general")
    k = _lib.C.c_int32(-1)
    assert F.lib().femtc_apply_kernel(7, F.F64, 16, 17, 0, _lib.C.byref(k)) \
        == _lib.ERR_INVALID_ARGUMENT
    assert F.lib().femtc_apply_kernel(0, F.F64, 26, 27, 0, _lib.C.byref(k)) \
        == _lib.ERR_DEGREE_NOT_SUPPORTED
    assert F.lib().femtc_apply_kernel(0, F.F64, 16, 17, 0, None) == \
        _lib.ERR_INVALID_ARGUMENT


def test_solver_create_rejects_bad_problems_without_touching_the_device():
    # argument checks run before any HIP call
    import ctypes as C_
    L = F.lib()
    h = C_.c_void_p()
    p = _lib.Problem(F.OP_STIFFNESS, 3, 4, 10, 100, 1, 1, None, None)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "G (or the cell Jacobians J) is required" in _lib.last_error()
    p = _lib.Problem(F.OP_MASS, 9, 10, 10, 100, 1, 1, None, 1)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_DEGREE_NOT_SUPPORTED
    p = _lib.Problem(7, 3, 4, 10, 100, 1, 1, 1, 1)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    p = _lib.Problem(F.OP_KPM, 3, 4, 10, 100, 1, 1, 1, 1)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.BF16, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_solver_create(None, C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    # structured grid and scatter selection
    grid = C_.c_int32 * 3
    p = _lib.Problem(F.OP_KPM, 3, 4, 10, 100, 1, 1, 1, 1, None, grid(2, 2, 2))
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "grid" in _lib.last_error()
    p = _lib.Problem(F.OP_KPM, 3, 4, 10, 100, 1, 1, 1, 1, None, grid(0, 0, 0),
                     F.SCATTER_ASSEMBLED)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "need the structured grid" in _lib.last_error()
    p = _lib.Problem(F.OP_KPM, 3, 4, 10, 100, 1, 1, 1, 1, None, grid(0, 0, 0),
                     F.SCATTER_ZRUN)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "need the structured grid" in _lib.last_error()
    p = _lib.Problem(F.OP_KPM, 8, 9, 10, 100, 1, 1, 1, 1, None, grid(1, 2, 5),
                     F.SCATTER_ZRUN)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "nq <= 8" in _lib.last_error()
    p = _lib.Problem(F.OP_KPM, 3, 4, 10, 100, 1, 1, 1, 1, None, grid(1, 2, 5),
                     7)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "femtc_scatter" in _lib.last_error()
    p = _lib.Problem(F.OP_KPM, 3, 4, 10, 100, 1, 1, 1, 1, None, grid(1, 2, 5),
                     F.SCATTER_AUTO, 3)
    assert L.femtc_solver_create(C_.byref(h), C_.byref(p), F.F64, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "even_odd" in _lib.last_error()
    assert h.value is None
    assert L.femtc_solver_set_batch(None, 16, 1) == _lib.ERR_INVALID_ARGUMENT
    s = C_.c_int32()
    assert L.femtc_solver_scatter(None, C_.byref(s)) == _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_solver_even_odd(None, C_.byref(s)) == \
        _lib.ERR_INVALID_ARGUMENT


def test_plan_create_validates_arguments_without_touching_the_device():
    import ctypes as C_
    L = F.lib()
    h = C_.c_void_p()
    assert L.femtc_plan_create(C_.byref(h), 0, 2, 2, 3, 1, 100, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_plan_create(C_.byref(h), 2, 2, 2, 9, 1, 100, None) == \
        _lib.ERR_DEGREE_NOT_SUPPORTED
    assert L.femtc_plan_create(C_.byref(h), 2, 2, 2, 3, None, 100, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_plan_create(None, 2, 2, 2, 3, 1, 100, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert h.value is None
    eo = C_.c_int32(7)
    assert L.femtc_plan_set_tables(None, F.F64, 4, 1, None, C_.byref(eo)) == \
        _lib.ERR_INVALID_ARGUMENT
    assert eo.value == 0
    assert L.femtc_plan_apply(None, 0, 0, 4, 1, 1, 1, None, 1, 0, None) == \
        _lib.ERR_INVALID_ARGUMENT


def test_batch_create_rejects_bad_items_without_touching_the_device():
    import ctypes as C_
    L = F.lib()
    h = C_.c_void_p()
    items = (_lib.BatchItem * 2)()
    items[0] = _lib.BatchItem(F.OP_MASS, F.F64, 3, 4, 1, 1, 1, None, 1, None, 10, 0)
    items[1] = _lib.BatchItem(F.OP_STIFFNESS, F.F64, 3, 4, 1, 1, 1, None, None,
                              None, 10, 0)  # stiffness without G
    assert L.femtc_batch_create(C_.byref(h), items, 2, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert "batch item 1" in _lib.last_error()
    items[1] = _lib.BatchItem(F.OP_MASS, F.F64, 3, 9, 1, 1, 1, None, 1, None, 10, 0)
    assert L.femtc_batch_create(C_.byref(h), items, 2, None) == \
        _lib.ERR_INVALID_ARGUMENT  # nq > degree + 3
    items[1] = _lib.BatchItem(F.OP_MASS, F.F64, 3, 4, 1, 1, 1, None, 1, 1, 10, 0)
    assert L.femtc_batch_create(C_.byref(h), items, 2, None) == \
        _lib.ERR_INVALID_ARGUMENT  # continuous without ndofs
    assert h.value is None


def test_uneven_batch_spec_is_seeded_and_clipped():
    from femtc_amd import workloads as W
    a = W.uneven_batch_spec(seed=1)
    b = W.uneven_batch_spec(seed=1)
    assert a == b and len(a) == 200
    cells = [c for c, _, _ in a]
    assert min(cells) >= 1000 and max(cells) <= 100000
    assert {p for _, p, _ in a} <= set(range(2, 9))
    assert sum(op == F.OP_MASS for _, _, op in a) == 100
    nx, ny, nz = W.box_dims(12345)
    assert abs(nx * ny * nz - 12345) / 12345 < 0.1
    np.testing.assert_array_equal(W.uniform(5, 1000), O.uniform(5, 1000))


def test_latency_stats():
    res = [{"t_arrival": 0.0, "t_end": 0.010 * (k + 1), "kind": k % 2}
           for k in range(100)]
    st = F.latency_stats(res)
    assert st["n"] == 100
    assert abs(st["p50_ms"] - 505.0) < 1.0
    assert st["p99_ms"] > 985.0
    assert abs(st["solves_per_s"] - 50.0) < 1e-9
    assert F.latency_stats([]) == {"n": 0}


def test_basis_tables_weights_host():
    for P, Q in ((1, 2), (4, 5), (6, 9)):
        t = F.basis_tables_weights(P, Q)
        N = P + 1
        np.testing.assert_allclose(t[:2 * N * Q], O.basis_tables(P, Q),
                                   rtol=0, atol=1e-13)
        _, w = O.gl_rule(Q)
        np.testing.assert_allclose(t[2 * N * Q:], w, rtol=1e-13)
        assert abs(t[2 * N * Q:].sum() - 1.0) < 1e-14
