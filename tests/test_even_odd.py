"""CPU checks of the even-odd table set (femtc_even_odd_tables), which plans
and solvers build from the caller's tables and which the EO kernel
instances read.  Contracting random lines through the set with the even-odd
recombination must give the direct contraction with the original table, for
every term, degree and quadrature size, so a wrong coefficient, sign, block
order or padding fails.  The symmetry check must accept the GLL/Gauss tables
and refuse perturbed or random ones."""
import numpy as np
import pytest

import femtc_amd as F


def _term_dims(K, M, S):
    Kp, Hk, Mp, Hm = (K + 1) // 2, K // 2, (M + 1) // 2, M // 2
    RA, RB = -(-Mp // 4) * 4, -(-Hm // 4) * 4
    rowsA, rowsB = (Kp, Hk) if S > 0 else (Hk, Kp)
    return Kp, Hk, Mp, Hm, RA, RB, rowsA, rowsB


def _contract(eo, off, K, M, S, v):
    """out[m] = sum_k A(k, m) v[k] through the term at offset off; returns
    (out, term length)."""
    Kp, Hk, Mp, Hm, RA, RB, rowsA, rowsB = _term_dims(K, M, S)
    A = eo[off:off + rowsA * RA].reshape(rowsA, RA)
    B = eo[off + rowsA * RA:off + rowsA * RA + rowsB * RB].reshape(rowsB, RB)
    assert not A[:, Mp:].any() and not B[:, Hm:].any()  # padding is zero
    e = np.array([v[k] + v[K - 1 - k] for k in range(Hk)] +
                 ([v[Hk]] if K % 2 else []))
    o = np.array([v[k] - v[K - 1 - k] for k in range(Hk)])
    ina, inb = (e, o) if S > 0 else (o, e)
    ev = ina @ A[:, :Mp]
    od = inb @ B[:, :Hm]
    out = np.empty(M)
    for m in range(Hm):
        out[m] = ev[m] + od[m]
        out[M - 1 - m] = ev[m] - od[m]
    if M % 2:
        out[Hm] = ev[Hm]
    return out, rowsA * RA + rowsB * RB


@pytest.mark.parametrize("P", range(1, 9))
@pytest.mark.parametrize("dq", [1, 2, 3])
def test_even_odd_set_contracts_like_the_tables(P, dq):
    Q, N = P + dq, P + 1
    tab = F.basis_tables(P, Q)
    _, w = F.gauss_rule(Q)
    eo, sym = F.even_odd_tables(P, Q, tab, w)
    assert sym
    phi = tab[:N * Q].reshape(N, Q)
    dphi = tab[N * Q:].reshape(N, Q)
    rng = np.random.default_rng(100 * P + dq)
    off = 0
    # forward: A(k=dof, m=point) = t[k][m]; backward: A(k=point, m=dof)
    for A, S in ((phi, 1), (dphi, -1), (phi.T, 1), (dphi.T, -1)):
        K, M = A.shape
        for _ in range(3):
            v = rng.uniform(-1, 1, K)
            out, size = _contract(eo, off, K, M, S, v)
            ref = A.T @ v
            assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max() + 1e-15
        off += size
    np.testing.assert_array_equal(eo[off:], w)


def test_symmetry_check_refuses_other_tables():
    P, Q = 4, 5
    tab = F.basis_tables(P, Q)
    assert F.even_odd_tables(P, Q, tab)[1]
    bad = tab.copy()
    bad[3] *= 1.0 + 1e-9  # phi loses its symmetry
    assert not F.even_odd_tables(P, Q, bad)[1]
    bad = tab.copy()
    bad[2 * (P + 1) * Q - 1] += 1e-9  # dphi loses its antisymmetry
    assert not F.even_odd_tables(P, Q, bad)[1]
    rnd = np.random.default_rng(0).uniform(-1, 1, tab.size)
    assert not F.even_odd_tables(P, Q, rnd)[1]
    # weights absent: zeros
    eo, _ = F.even_odd_tables(P, Q, tab)
    assert not eo[-Q:].any()


def test_arguments_are_checked():
    import ctypes as C
    from femtc_amd import _lib
    size, sym = C.c_int64(), C.c_int32()
    t = np.zeros(2 * 10 * 10)
    p = t.ctypes.data_as(_lib._dp)
    L = F.lib()
    assert L.femtc_even_odd_tables(9, 10, p, None, None, C.byref(size),
                                   C.byref(sym)) == _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_even_odd_tables(3, 4, None, None, None, C.byref(size),
                                   C.byref(sym)) == _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_even_odd_tables(3, 4, p, None, None, None,
                                   C.byref(sym)) == _lib.ERR_INVALID_ARGUMENT
