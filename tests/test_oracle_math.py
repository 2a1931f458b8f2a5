"""Pin the correct-mode oracle with mathematics the reference does not test.

Each check is chosen so a plausible mistake fails it: a transposed table,
a dropped gradient chain, a wrong G slot or a wrong dofmap stride.
Brute-force references are dense einsum/kron constructions (a different
algorithm than the oracle's rotations), closed forms and invariants.
"""
import numpy as np
import pytest

import oracle as O


def _dense_B(ph, gr):
    """Dense [q][dof] interpolation and gradient matrices for one cell:
    B0 = g(x)phi(x)phi, B1 = phi(x)g(x)phi, B2 = phi(x)phi(x)g, Bm = phi^3,
    each of shape (Q^3, N^3); tables are [dof][q]."""
    P_, G_ = ph.T, gr.T  # [q][dof]
    Bm = np.kron(np.kron(P_, P_), P_)
    B0 = np.kron(np.kron(G_, P_), P_)
    B1 = np.kron(np.kron(P_, G_), P_)
    B2 = np.kron(np.kron(P_, P_), G_)
    return Bm, (B0, B1, B2)


def _S_from_G(g):
    # packing (S22, S12, S02, S11, S01, S00), src/stiffness.hpp:25-27
    return np.array([[g[5], g[4], g[2]], [g[4], g[3], g[1]],
                     [g[2], g[1], g[0]]])


def _dense_cell(op, N, Q, tab, Gc=None, Dc=None):
    ph = tab[: N * Q].reshape(N, Q)
    gr = tab[N * Q:].reshape(N, Q) if tab.size > N * Q else np.zeros((N, Q))
    Bm, B = _dense_B(ph, gr)
    A = np.zeros((N ** 3, N ** 3))
    if op in ("mass", "kpm"):
        A += Bm.T @ (Dc[:, None] * Bm)
    if op in ("stiffness", "kpm"):
        Gq = Gc.reshape(Q ** 3, 6)
        for r in range(3):
            for s in range(3):
                Srs = np.array([_S_from_G(g)[r, s] for g in Gq])
                A += B[r].T @ (Srs[:, None] * B[s])
    return A


def test_uniform_matches_c_generator_bitwise():
    for seed in (0, 1, 12345):
        np.testing.assert_array_equal(O.uniform(seed, 1000, -1, 1),
                                      O.fill_uniform(seed, 1000, -1, 1))
        np.testing.assert_array_equal(O.uniform(seed, 333, 0.5, 1.5),
                                      O.fill_uniform(seed, 333, 0.5, 1.5))
    u = O.uniform(3, 100000, -1, 1)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01


def test_gemm_blocked_is_a_plain_product():
    # closed form: C = A B with A col-major, B row-major, C col-major
    for (k, m, n) in ((3, 3, 9), (5, 4, 25), (4, 5, 20), (7, 7, 49)):
        a = O.uniform(k + m, m * k)
        b = O.uniform(k + n, k * n)
        A = a.reshape(k, m).T
        B = b.reshape(k, n)
        for order in range(6):
            for (MB, NB) in ((0, 0), (4, 8), (2, 3), (16, 16)):
                c = O.gemm_blocked(k, m, n, order, MB, NB, a, b)
                np.testing.assert_allclose(c.reshape(n, m).T, A @ B,
                                           rtol=1e-14, atol=1e-14)


def test_gemm_micro_accumulates():
    k, m, nc, lda, ldb, ldc = 6, 3, 5, 4, 7, 9
    a = O.uniform(1, lda * k)
    b = O.uniform(2, k * ldb)
    c0 = O.uniform(3, m * ldc)
    A = np.array([[a[i + p * lda] for p in range(k)] for i in range(m)])
    B = np.array([[b[p * ldb + j] for j in range(nc)] for p in range(k)])
    for order in range(6):
        c = O.gemm_micro(k, m, nc, lda, ldb, ldc, order, a, b, c0)
        C = c.reshape(m, ldc)
        np.testing.assert_allclose(C[:, :nc], c0.reshape(m, ldc)[:, :nc] + A @ B,
                                   rtol=1e-14)
        np.testing.assert_array_equal(C[:, nc:], c0.reshape(m, ldc)[:, nc:])


@pytest.mark.parametrize("Na,Nb", [(2, 2), (5, 4), (4, 5), (3, 6)])
@pytest.mark.parametrize("Tr", [0, 1])
def test_tmm_brute_force(Na, Nb, Tr):
    I = [O.uniform(10 + j, Na * Nb) for j in range(3)]
    B = O.uniform(20, Nb ** 3).reshape(Nb, Nb, Nb)
    if Tr:  # [a][b]
        M = [x.reshape(Na, Nb) for x in I]
    else:   # [b][a]
        M = [x.reshape(Nb, Na).T for x in I]
    ref = np.einsum("ax,by,cz,xyz->abc", M[0], M[1], M[2], B)
    A = O.tmm_apply(Na, Nb, Tr, I[0], I[1], I[2], B.ravel())
    np.testing.assert_allclose(A, ref.ravel(), rtol=1e-13, atol=1e-14)


def test_transform_closed_form():
    nq = 50
    G = O.uniform(5, nq * 6)
    w = [O.uniform(6 + j, nq) for j in range(3)]
    f0, f1, f2 = O.transform(G, *w)
    for q in range(nq):
        S = _S_from_G(G[6 * q: 6 * q + 6])
        v = S @ np.array([w[0][q], w[1][q], w[2][q]])
        np.testing.assert_allclose([f0[q], f1[q], f2[q]], v, rtol=1e-14,
                                   atol=1e-15)


def test_literal_mass_and_stiffness_formulas():
    # Appendix-style restatement of what the reference computes, by einsum.
    P = 3
    N = P + 1
    phi2 = O.uniform(31, 2 * N * N, 0, 1)
    ph, gr = phi2[: N * N].reshape(N, N), phi2[N * N:].reshape(N, N)
    U = O.uniform(32, N ** 3, 0, 1).reshape(N, N, N)
    D = O.uniform(33, N ** 3, 0, 1).reshape(N, N, N)
    G = O.uniform(34, N ** 3 * 6, 0, 1).reshape(N, N, N, 6)

    def chain(A0, A1, A2, X):  # tables [src][dst]
        return np.einsum("ai,bj,ck,abc->ijk", A0, A1, A2, X)

    Wm = chain(ph, ph, ph, D * chain(ph, ph, ph, U))
    np.testing.assert_allclose(O.mass_literal(P, ph.ravel(), U.ravel(),
                                              D.ravel(), 1), Wm.ravel(),
                               rtol=1e-13)
    wx = chain(gr, ph, ph, U)
    wy = chain(ph, gr, ph, U)
    wz = chain(ph, ph, ph, U)  # not differentiated (src/stiffness.hpp:76-77)
    f2 = G[..., 2] * wx + G[..., 1] * wy + G[..., 0] * wz
    Ws = chain(ph, ph, gr, f2)
    np.testing.assert_allclose(O.stiffness_literal(P, phi2, U.ravel(),
                                                    G.ravel(), 1), Ws.ravel(),
                               rtol=1e-13)


@pytest.mark.parametrize("N,Q", [(2, 2), (4, 5), (5, 5), (3, 5), (4, 4)])
@pytest.mark.parametrize("op", ["mass", "stiffness", "kpm"])
def test_correct_cell_operator_dense(op, N, Q):
    tab = O.uniform(40 + N, 2 * N * Q)
    ncell = 2
    U = O.uniform(41, ncell * N ** 3)
    G = O.uniform(42, ncell * Q ** 3 * 6)
    Dw = O.uniform(43, ncell * Q ** 3, 0.5, 1.5)
    W = O.cell_operator(op, N, Q, tab, U, G, Dw, ncell)
    for c in range(ncell):
        A = _dense_cell(op, N, Q, tab, G[c * Q ** 3 * 6:(c + 1) * Q ** 3 * 6],
                        Dw[c * Q ** 3:(c + 1) * Q ** 3])
        ref = A @ U[c * N ** 3:(c + 1) * N ** 3]
        np.testing.assert_allclose(W[c * N ** 3:(c + 1) * N ** 3], ref,
                                   rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_mass_with_identity_table_is_pointwise():
    N = Q = 4
    tab = np.eye(N).ravel()
    U = O.uniform(50, 3 * N ** 3)
    Dw = O.uniform(51, 3 * N ** 3, 0.5, 1.5)
    W = O.cell_operator("mass", N, Q, tab, U, None, Dw, 3)
    np.testing.assert_allclose(W, Dw * U, rtol=1e-15)


@pytest.mark.parametrize("P,Q", [(1, 2), (3, 5), (4, 5), (6, 7)])
def test_mass_volume_and_polynomial_exactness(P, Q):
    N = P + 1
    tab = O.basis_tables(P, Q)
    J = O.random_jacobians(60, 1)
    G, Dw = O.geometry_affine(J, Q)
    one = np.ones(N ** 3)
    W = O.cell_operator("mass", N, Q, tab, one, G, Dw, 1)
    detJ = np.linalg.det(J.reshape(3, 3))
    np.testing.assert_allclose(W.sum(), detJ, rtol=1e-13)
    # u = x^P (degree P along a), v = z: integral over [0,1]^3 of x^P z
    x = O.gll_points(P)
    u = np.einsum("a,b,c->abc", x ** P, np.ones(N), np.ones(N)).ravel()
    v = np.einsum("a,b,c->abc", np.ones(N), np.ones(N), x).ravel()
    Dw1 = np.tile(O.weights3(Q), 1)  # J = I
    Mv = O.cell_operator("mass", N, Q, tab, v, None, Dw1, 1)
    np.testing.assert_allclose(u @ Mv, 1.0 / (P + 1) * 0.5, rtol=1e-13)


@pytest.mark.parametrize("P,Q", [(1, 2), (2, 3), (4, 5), (6, 7), (3, 5)])
def test_stiffness_kernel_and_energy(P, Q):
    N = P + 1
    tab = O.basis_tables(P, Q)
    J = O.random_jacobians(70 + P, 2)
    G, Dw = O.geometry_affine(J, Q)
    # K 1 = 0 (constants are in the kernel)
    W = O.cell_operator("stiffness", N, Q, tab, np.ones(2 * N ** 3), G, Dw, 2)
    assert np.abs(W).max() < 1e-13
    # linear u = c . x(xi) on an affine cell: energy = |c|^2 |det J|
    x = O.gll_points(P)
    Xi = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1).reshape(-1, 3)
    c = np.array([0.3, -1.1, 0.7])
    for cell in range(2):
        Jc = J[9 * cell: 9 * cell + 9].reshape(3, 3)
        u = (Xi @ Jc.T) @ c  # u(xi) = c . (J xi)
        Gc = G[cell * Q ** 3 * 6:(cell + 1) * Q ** 3 * 6]
        Ku = O.cell_operator("stiffness", N, Q, tab, u, Gc, None, 1)
        np.testing.assert_allclose(u @ Ku, (c @ c) * abs(np.linalg.det(Jc)),
                                   rtol=1e-12)
    # symmetry of the assembled element matrix
    A = _dense_cell("stiffness", N, Q, tab, G[: Q ** 3 * 6])
    np.testing.assert_allclose(A, A.T, atol=1e-13)


def test_dofmap_closed_form_and_multiplicity():
    nx, ny, nz, P = 3, 2, 4, 2
    N = P + 1
    dm = O.dofmap_structured(nx, ny, nz, P).reshape(nx, ny, nz, N, N, N)
    Gy, Gz = ny * P + 1, nz * P + 1
    Gx = nx * P + 1
    for (ci, cj, ck, a, b, c) in [(0, 0, 0, 0, 0, 0), (2, 1, 3, 2, 2, 2),
                                  (1, 0, 2, 1, 2, 0)]:
        assert dm[ci, cj, ck, a, b, c] == ((ci * P + a) * Gy + cj * P + b) * Gz + ck * P + c
    counts = np.bincount(dm.ravel(), minlength=Gx * Gy * Gz)
    assert counts.min() == 1 and len(counts) == Gx * Gy * Gz
    g = np.arange(Gx * Gy * Gz)
    I, Jj, K = g // (Gy * Gz), (g // Gz) % Gy, g % Gz

    def mult(i, n):  # cells sharing an index along one axis
        return np.where((i % P == 0) & (i > 0) & (i < n * P), 2, 1)

    np.testing.assert_array_equal(counts, mult(I, nx) * mult(Jj, ny) * mult(K, nz))


def test_geometry_closed_forms():
    Q = 3
    w3 = O.weights3(Q)
    G, D = O.geometry_affine(np.eye(3).ravel(), Q)
    G = G.reshape(-1, 6)
    np.testing.assert_allclose(D, w3, rtol=1e-15)
    np.testing.assert_allclose(G[:, [0, 3, 5]], np.stack([w3] * 3, 1), rtol=1e-15)
    assert np.all(G[:, [1, 2, 4]] == 0)
    # diagonal J: S = det * w * diag(1/d^2); slot 5 is S00 (first axis)
    d = np.array([2.0, 0.5, 4.0])
    G, D = O.geometry_affine(np.diag(d).ravel(), Q)
    G = G.reshape(-1, 6)
    det = d.prod()
    np.testing.assert_allclose(G[:, 5], det * w3 / d[0] ** 2, rtol=1e-15)
    np.testing.assert_allclose(G[:, 3], det * w3 / d[1] ** 2, rtol=1e-15)
    np.testing.assert_allclose(G[:, 0], det * w3 / d[2] ** 2, rtol=1e-15)
    # general J against numpy.linalg.inv
    J = O.random_jacobians(80, 5)
    G, D = O.geometry_affine(J, Q)
    for c in range(5):
        Jc = J[9 * c: 9 * c + 9].reshape(3, 3)
        Ji = np.linalg.inv(Jc)
        S = abs(np.linalg.det(Jc)) * Ji @ Ji.T
        for q in (0, 7, Q ** 3 - 1):
            g = G[(c * Q ** 3 + q) * 6:(c * Q ** 3 + q) * 6 + 6]
            np.testing.assert_allclose(_S_from_G(g), S * w3[q], rtol=1e-13)


@pytest.mark.parametrize("op", ["mass", "stiffness", "kpm"])
def test_global_operator_and_diagonal_dense(op):
    nx, ny, nz, P, Q = 2, 3, 2, 2, 3
    N = P + 1
    tab = O.basis_tables(P, Q)
    C_ = nx * ny * nz
    dm = O.dofmap_structured(nx, ny, nz, P)
    ng = (nx * P + 1) * (ny * P + 1) * (nz * P + 1)
    J = O.random_jacobians(90, C_)
    G, Dw = O.geometry_affine(J, Q)
    A = np.zeros((ng, ng))
    for c in range(C_):
        Ac = _dense_cell(op, N, Q, tab, G[c * Q ** 3 * 6:(c + 1) * Q ** 3 * 6],
                         Dw[c * Q ** 3:(c + 1) * Q ** 3])
        idx = dm[c * N ** 3:(c + 1) * N ** 3]
        A[np.ix_(idx, idx)] += Ac
    x = O.uniform(91, ng)
    y = O.global_operator(op, N, Q, tab, x, dm, G, Dw)
    np.testing.assert_allclose(y, A @ x, rtol=1e-12, atol=1e-13)
    d = O.global_diagonal(op, N, Q, tab, ng, dm, G, Dw)
    np.testing.assert_allclose(d, np.diag(A), rtol=1e-12, atol=1e-14)


def test_pcg_against_dense_solve():
    nx = ny = nz = 2
    P, Q = 2, 3
    N = P + 1
    tab = O.basis_tables(P, Q)
    dm = O.dofmap_structured(nx, ny, nz, P)
    ng = (nx * P + 1) ** 3
    J = O.random_jacobians(95, 8)
    G, Dw = O.geometry_affine(J, Q)
    A = np.zeros((ng, ng))
    for c in range(8):
        Ac = _dense_cell("kpm", N, Q, tab, G[c * Q ** 3 * 6:(c + 1) * Q ** 3 * 6],
                         Dw[c * Q ** 3:(c + 1) * Q ** 3])
        idx = dm[c * N ** 3:(c + 1) * N ** 3]
        A[np.ix_(idx, idx)] += Ac
    f = O.uniform(96, ng)
    u, it, rel = O.pcg("kpm", N, Q, tab, f, dm, G, Dw, rtol=1e-12)
    assert rel <= 1e-12 and it < 200
    np.testing.assert_allclose(u, np.linalg.solve(A, f), rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("P", [1, 2, 3, 4, 6, 8])
def test_tables(P):
    N = P + 1
    x = O.gll_points(P)
    assert x[0] == 0.0 and x[-1] == 1.0 and np.all(np.diff(x) > 0)
    closed = {2: [0, 0.5, 1],
              3: [0, (1 - 1 / np.sqrt(5)) / 2, (1 + 1 / np.sqrt(5)) / 2, 1],
              4: [0, (1 - np.sqrt(3 / 7)) / 2, 0.5, (1 + np.sqrt(3 / 7)) / 2, 1]}
    if P in closed:
        np.testing.assert_allclose(x, closed[P], atol=1e-15)
    V, D = O.lagrange(x, x)
    np.testing.assert_allclose(V, np.eye(N), atol=1e-13)
    for Q in (P + 1, P + 2):
        xq, wq = O.gl_rule(Q)
        assert abs(wq.sum() - 1) < 1e-15
        for k in range(2 * Q):  # exact to degree 2Q-1
            assert abs(wq @ xq ** k - 1 / (k + 1)) < 1e-14
        tab = O.basis_tables(P, Q)
        ph, gr = tab[: N * Q].reshape(N, Q), tab[N * Q:].reshape(N, Q)
        np.testing.assert_allclose(ph.sum(0), 1, atol=1e-13)
        np.testing.assert_allclose(gr.sum(0), 0, atol=1e-11)
        # interpolate x^P exactly and differentiate it
        np.testing.assert_allclose((x ** P) @ ph, xq ** P, atol=1e-13)
        np.testing.assert_allclose((x ** P) @ gr, P * xq ** (P - 1), atol=1e-11)
