"""Pin the C restatement (oracle/femtc_oracle.c) to the compiled reference.

The fixtures in tests/golden/ref_goldens.npz were written by
tests/golden/gen_golden.py from the reference headers compiled in place
(oracle/Makefile, target ``ref``).  Both sides are built with -O2 and no
-march, so no a*b+c is contracted into an FMA and the literal restatement
must match bit for bit.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden",
                            "ref_goldens.npz"))


def test_gemm_blocked_matches_reference_bitwise():
    # src/linalg.hpp:101-182, every order and block size in the fixture
    for i in range(int(GOLD["gb_count"][0])):
        P, order, MB, NB = GOLD[f"gb_{i}_meta"]
        N = P + 1
        c = O.gemm_blocked(N, N, N * N, order, MB, NB, GOLD[f"gb_{i}_a"],
                           GOLD[f"gb_{i}_b"])
        np.testing.assert_array_equal(c, GOLD[f"gb_{i}_c"],
                                      err_msg=f"P={P} order={order} {MB}x{NB}")


def test_gemm_micro_matches_reference_bitwise():
    # src/linalg.hpp:45-94 at the tests/test_micro.cpp shape
    for order in range(6):
        c = O.gemm_micro(26, 4, 16, 4, 16, 16, order, GOLD[f"gm_{order}_a"],
                         GOLD[f"gm_{order}_b"], GOLD[f"gm_{order}_c0"])
        np.testing.assert_array_equal(c, GOLD[f"gm_{order}_c"])


def test_tmm_matches_reference():
    # src/tmm.hpp:91-126; the restatement contracts in a different order
    # than the reference's explicit transposes, so compare to rounding.
    for k in range(int(GOLD["tmm_count"][0])):
        Na, Nb, Tr = GOLD[f"tmm_{k}_meta"]
        A = O.tmm_apply(Na, Nb, Tr, GOLD[f"tmm_{k}_I0"], GOLD[f"tmm_{k}_I1"],
                        GOLD[f"tmm_{k}_I2"], GOLD[f"tmm_{k}_B"])
        ref = GOLD[f"tmm_{k}_A"]
        assert np.linalg.norm(A - ref) <= 1e-14 * np.linalg.norm(ref), (Na, Nb, Tr)


def test_mass_literal_matches_reference_bitwise():
    # src/mass.hpp:10-49 through the dispatch ladder (:196-219)
    for k in range(int(GOLD["mass_count"][0])):
        P, sym, order, Mb, Nb, ncell = GOLD[f"mass_{k}_meta"]
        W = O.mass_literal(P, GOLD[f"mass_{k}_phi"], GOLD[f"mass_{k}_U"],
                           GOLD[f"mass_{k}_detJ"], ncell)
        np.testing.assert_array_equal(W, GOLD[f"mass_{k}_W"],
                                      err_msg=f"P={P} order={order}")


def test_stiffness_literal_matches_reference_bitwise():
    # src/stiffness.hpp:34-99 through the dispatch ladder (:245-268)
    for k in range(int(GOLD["stiff_count"][0])):
        P, order, Mb, Nb, ncell = GOLD[f"stiff_{k}_meta"]
        W = O.stiffness_literal(P, GOLD[f"stiff_{k}_phi2"],
                                GOLD[f"stiff_{k}_U"], GOLD[f"stiff_{k}_G"],
                                ncell)
        np.testing.assert_array_equal(W, GOLD[f"stiff_{k}_W"],
                                      err_msg=f"P={P} order={order}")


@pytest.mark.skipif(not O.ref_available("kernels"),
                    reason="compiled reference not present")
def test_gemm_blocked_live_reference_large_degrees():
    # degrees beyond the fixture, straight against the compiled reference
    L = O.ref_lib("kernels")
    for P in (8, 12, 25):
        N = P + 1
        for order in range(6):
            a = O.uniform(77 + P, N * N, 0, 1)
            b = O.uniform(99 + P, N ** 3, 0, 1)
            ref = np.zeros(N ** 3)
            assert L.ref_gemm_blocked(P, order, 8, 16, a, b, ref) == 0
            c = O.gemm_blocked(N, N, N * N, order, 8, 16, a, b)
            np.testing.assert_array_equal(c, ref)


@pytest.mark.skipif(not (O.ref_available("mass") and
                         O.ref_available("stiffness")),
                    reason="compiled reference not present")
@pytest.mark.parametrize("P", [5, 9, 16, 25])
def test_literal_operators_live_reference_high_degrees(P):
    # degrees past the fixture, up to the top of the reference's ladder
    # (src/mass.hpp:52-138, src/stiffness.hpp:102-188), bit for bit
    N = P + 1
    ncell = 2
    phi2 = O.uniform(600 + P, 2 * N * N, 0, 1)
    U = O.uniform(601 + P, ncell * N ** 3, 0, 1)
    D = O.uniform(602 + P, ncell * N ** 3, 0, 1)
    G = O.uniform(603 + P, ncell * N ** 3 * 6, 0, 1)
    ref = np.zeros(ncell * N ** 3)
    O.ref_lib("mass").ref_mass_operator_f64(phi2[:N * N].copy(), U.copy(),
                                            ref, D.copy(), ncell, P, 0, 0, 0)
    np.testing.assert_array_equal(O.mass_literal(P, phi2[:N * N], U, D, ncell),
                                  ref)
    ref = np.zeros(ncell * N ** 3)
    O.ref_lib("stiffness").ref_stiffness_operator_f64(phi2.copy(), U.copy(),
                                                      ref, G.copy(), ncell, P,
                                                      0, 0, 0)
    np.testing.assert_array_equal(O.stiffness_literal(P, phi2, U, G, ncell),
                                  ref)
