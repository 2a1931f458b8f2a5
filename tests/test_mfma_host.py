"""Host-side checks of the bfloat16-MFMA operand set (femtc_mfma_tables_host,
no device work): every A matrix is rebuilt from its lane layout
(lane l, slot s -> A[l % 16][8 (l / 16) + s]) and compared with the oracle's
tables; and the apply entry point's argument checks."""
import ctypes as C

import numpy as np
import pytest

import femtc_amd as F
import oracle as O
from femtc_amd import _lib


def bf16_to_f64(h):
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def operand_set(P, split):
    tab = np.ascontiguousarray(O.basis_tables(P, P + 1))
    nbytes = C.c_size_t()
    assert _lib.lib().femtc_mfma_tables_bytes(split, C.byref(nbytes)) == 0
    buf = np.zeros(nbytes.value // 2, dtype=np.uint16)
    st = _lib.lib().femtc_mfma_tables_host(
        P, P + 1, split, tab.ctypes.data_as(C.POINTER(C.c_double)),
        buf.ctypes.data_as(C.c_void_p))
    assert st == 0, _lib.last_error()
    nm = 11 if split else 7
    mats = np.zeros((nm, 16, 32))
    A = bf16_to_f64(buf[:nm * 64 * 8]).reshape(nm, 64, 8)
    for lane in range(64):
        mats[:, lane % 16, 8 * (lane // 16):8 * (lane // 16) + 8] = A[:, lane]
    w = buf[nm * 64 * 8:].view(np.float32)
    return tab, mats, w


def tables(P):
    N = Q = P + 1
    t = O.basis_tables(P, Q)
    return t[:N * Q].reshape(N, Q), t[N * Q:].reshape(N, Q)


@pytest.mark.parametrize("P", [1, 3, 6, 7])
def test_unsplit_operand_set(P):
    N = Q = P + 1
    phi, dphi = tables(P)
    _, M, w = operand_set(P, 0)
    scale = np.abs(dphi).max()
    tol = 2.0 ** -16 * scale

    def two(m, rows, g0):  # hi + lo groups (g0, g0+1), slots 0..7
        return M[m][rows][:, 8 * g0:8 * g0 + 8] + M[m][rows][:, 8 * g0 + 8:8 * g0 + 16]

    # F: block 0 rows q = phi[:, q], block 1 = dphi; groups 2, 3 zero
    assert np.abs(two(0, slice(0, Q), 0)[:, :N] - phi.T).max() <= tol
    assert np.abs(two(0, slice(8, 8 + Q), 0)[:, :N] - dphi.T).max() <= tol
    assert not M[0][:, 16:].any()
    assert not M[0][:, N:8].any() and not M[0][:, 8 + N:16].any()
    # the hi part alone is bfloat16-accurate only: the lo part matters
    assert np.abs(M[0][0:Q, :N] - phi.T).max() > 0
    # F3 quads: rows 4 qa + m (u, d_a from groups 0,1; d_b from 2,3)
    for quad, (mx, my) in enumerate(((1, 2), (3, 4))):
        for qa in range(4):
            q = 4 * quad + qa
            for m, (grp, t) in enumerate(((0, phi), (0, dphi), (2, phi))):
                row = two(mx, [4 * qa + m], grp)[0, :N]
                ref = t[:, q] if q < Q else np.zeros(N)
                assert np.abs(row - ref).max() <= tol
            row = two(my, [4 * qa + 3], 0)[0, :N]
            assert np.abs(row - (phi[:, q] if q < Q else 0)).max() <= tol
    # backward: BPG block 0 rows d = (phi[d, :], dphi[d, :]); BPP blocks
    assert np.abs(two(5, slice(0, N), 0)[:, :Q] - phi).max() <= tol
    assert np.abs(two(5, slice(0, N), 2)[:, :Q] - dphi).max() <= tol
    assert not M[5][8:].any()
    assert np.abs(two(6, slice(0, N), 0)[:, :Q] - phi).max() <= tol
    assert np.abs(two(6, slice(8, 8 + N), 2)[:, :Q] - phi).max() <= tol
    assert not M[6][0:8, 16:].any() and not M[6][8:, :16].any()
    xq, wq = O.gl_rule(Q)
    assert np.allclose(w[:Q], wq, rtol=1e-7) and not w[Q:].any()


@pytest.mark.parametrize("P", [2, 6])
def test_split_operand_set(P):
    N = Q = P + 1
    phi, dphi = tables(P)
    _, M, _ = operand_set(P, 1)
    tol = 2.0 ** -16 * np.abs(dphi).max()
    # groups (hi, hi, lo, 0): hi + lo of a row is the table
    def full(m, rows):
        return M[m][rows][:, 0:8] + M[m][rows][:, 16:24]
    assert np.array_equal(M[0][:, 0:8], M[0][:, 8:16])
    assert not M[0][:, 24:].any()
    assert np.abs(full(0, slice(0, Q))[:, :N] - phi.T).max() <= tol
    assert np.abs(full(0, slice(8, 8 + Q))[:, :N] - dphi.T).max() <= tol
    # backward Bp0, Bg0, Bp1, Bpp
    assert np.abs(full(7, slice(0, N))[:, :Q] - phi).max() <= tol
    assert np.abs(full(8, slice(0, N))[:, :Q] - dphi).max() <= tol
    assert not M[7][8:].any() and not M[9][:8].any()
    assert np.abs(full(9, slice(8, 8 + N))[:, :Q] - phi).max() <= tol
    assert np.abs(full(10, slice(8, 8 + N))[:, :Q] - phi).max() <= tol
    # F3: matrix 1 carries u and d_a rows only, 2 d_b, 3 d_c (qa 0..3)
    for qa in range(min(4, Q)):
        assert np.abs(full(1, [4 * qa])[0, :N] - phi[:, qa]).max() <= tol
        assert np.abs(full(1, [4 * qa + 1])[0, :N] - dphi[:, qa]).max() <= tol
        assert np.abs(full(2, [4 * qa + 2])[0, :N] - phi[:, qa]).max() <= tol
        assert np.abs(full(3, [4 * qa + 3])[0, :N] - phi[:, qa]).max() <= tol
        assert not M[1][[4 * qa + 2, 4 * qa + 3]].any()


def test_apply_mfma_argument_checks():
    L = _lib.lib()
    ops = C.c_void_p(8)  # never dereferenced: every call below fails first
    x = y = cg = lat = C.c_void_p(8)
    # degree 8 is beyond the 8 K slots of a group
    assert L.femtc_global_apply_mfma(2, 8, 9, 1, ops, x, y, 17 ** 3, 2, 2, 2,
                                     cg, lat, 0, None) == _lib.ERR_DEGREE_NOT_SUPPORTED
    # nq must be degree + 1
    assert L.femtc_global_apply_mfma(2, 4, 6, 1, ops, x, y, 9 ** 3, 2, 2, 2,
                                     cg, lat, 0, None) == _lib.ERR_DEGREE_NOT_SUPPORTED
    # ndofs must be the structured grid's dof count
    assert L.femtc_global_apply_mfma(2, 4, 5, 1, ops, x, y, 100, 2, 2, 2,
                                     cg, lat, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_global_apply_mfma(2, 4, 5, 2, ops, x, y, 9 ** 3, 2, 2, 2,
                                     cg, lat, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_global_apply_mfma(5, 4, 5, 1, ops, x, y, 9 ** 3, 2, 2, 2,
                                     cg, lat, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.femtc_global_apply_mfma(2, 4, 5, 1, None, x, y, 9 ** 3, 2, 2, 2,
                                     cg, lat, 0, None) == _lib.ERR_INVALID_ARGUMENT
