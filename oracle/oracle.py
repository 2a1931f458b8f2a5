"""CPU oracle for the femtc sum-factorized operator path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.  The product (``femtc_amd``) never does;
it fails loudly when its HIP library is missing instead of falling back here.

Contents
--------
* ctypes bindings to ``liboracle.so`` (``femtc_oracle.c``): the C restatement of
  the reference's contraction chain, in literal and correct mode, plus dofmap,
  geometry, global operator, diagonal and Jacobi-PCG.
* ctypes bindings to ``_ref/libref_*.so``: the reference's own headers compiled
  from where they lie (see ``Makefile``), used to pin the restatement.
* numpy 1-D tables (``gl_rule``, ``gll_points``, ``basis_tables``), written
  independently of the product's C++ table generator and pinned by
  ``numpy.polynomial.legendre.leggauss``.
* ``uniform``: the counter-based (SplitMix64) generator shared with the C side.

Parity status of each function is stated in its docstring; all are pinned by
at least one check in ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes as C
import os
from functools import lru_cache

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORDERS = ("ijk", "ikj", "jik", "jki", "kij", "kji")  # src/linalg.hpp:11-19

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _np_or_null(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class _NullableDP:
    """ctypes argtype accepting a float64 array or None (NULL)."""

    @classmethod
    def from_param(cls, obj):
        if obj is None:
            return None
        return _dp.from_param(obj)


@lru_cache(maxsize=None)
def lib():
    """Load liboracle.so, building it with make if it is missing."""
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        import subprocess

        subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    L = C.CDLL(path)
    i64, i32, dbl = C.c_int64, C.c_int, C.c_double
    L.orc_fill_uniform.argtypes = [C.c_uint64, i64, dbl, dbl, _dp]
    L.orc_gemm_micro.argtypes = [i32] * 7 + [_dp, _dp, _dp]
    L.orc_gemm_blocked.argtypes = [i32] * 6 + [_dp, _dp, _dp]
    L.orc_rot.argtypes = [i32, i32, i32, _dp, _dp, _dp]
    L.orc_transform.argtypes = [i32, _dp, _dp, _dp, _dp]
    L.orc_mass_literal.argtypes = [i32, _dp, _dp, _dp, _dp, i64]
    L.orc_stiffness_literal.argtypes = [i32, _dp, _dp, _dp, _dp, i64]
    L.orc_tmm_apply.argtypes = [i32, i32, i32, _dp, _dp, _dp, _dp, _dp]
    L.orc_cell_operator.argtypes = [i32, i32, i32, _dp, _dp, _dp, _NullableDP,
                                    _NullableDP, i64]
    L.orc_dofmap_structured.argtypes = [i32, i32, i32, i32, _ip]
    L.orc_geometry_affine.argtypes = [i64, i32, _dp, _dp, _NullableDP,
                                      _NullableDP]
    L.orc_global_operator.argtypes = [i32, i32, i32, _dp, _dp, _dp, i64, _ip,
                                      _NullableDP, _NullableDP, i64]
    L.orc_global_diagonal.argtypes = [i32, i32, i32, _dp, _dp, i64, _ip,
                                      _NullableDP, _NullableDP, i64]
    L.orc_pcg.restype = C.c_int
    L.orc_pcg.argtypes = [i32, i32, i32, _dp, _dp, _dp, i64, _ip, _NullableDP,
                          _NullableDP, i64, dbl, i32, C.POINTER(C.c_double)]
    return L


# --------------------------------------------------------------------------
# Random numbers (identical to orc_fill_uniform bit for bit)
# --------------------------------------------------------------------------
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0,
            offset: int = 0) -> np.ndarray:
    """value_i = lo + (hi-lo)*u_i with u_i = mix64(seed + (i+1)*golden) >> 11
    scaled by 2^-53: the SplitMix64 finaliser on a counter."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


# --------------------------------------------------------------------------
# 1-D tables on the reference interval [0, 1]
# --------------------------------------------------------------------------
def gl_rule(nq: int):
    """Gauss-Legendre points and weights on [0,1] (weights sum to 1)."""
    x, w = np.polynomial.legendre.leggauss(nq)
    return 0.5 * (x + 1.0), 0.5 * w


def gll_points(degree: int) -> np.ndarray:
    """Gauss-Lobatto-Legendre nodes on [0,1]: the end points and the roots
    of P'_degree, in increasing order."""
    if degree < 1:
        raise ValueError("degree must be >= 1")
    inner = np.polynomial.legendre.Legendre.basis(degree).deriv().roots()
    x = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    x = 0.5 * (x + 1.0)
    # symmetrise (the roots are symmetric about 1/2 in exact arithmetic)
    return 0.5 * (x + (1.0 - x[::-1]))


def lagrange(nodes: np.ndarray, x: np.ndarray):
    """Values and derivatives of the Lagrange basis on ``nodes`` at ``x``;
    both returned as [dof][point]."""
    n = len(nodes)
    V = np.ones((n, len(x)))
    D = np.zeros((n, len(x)))
    for d in range(n):
        others = [m for m in range(n) if m != d]
        den = np.prod([nodes[d] - nodes[m] for m in others])
        for m in others:
            V[d] *= (x - nodes[m])
        V[d] /= den
        for skip in others:
            term = np.ones_like(x)
            for m in others:
                if m != skip:
                    term = term * (x - nodes[m])
            D[d] += term
        D[d] /= den
    return V, D


def basis_tables(degree: int, nq: int) -> np.ndarray:
    """[phi | dphi]: Lagrange basis on GLL nodes evaluated (and
    differentiated) at the nq Gauss-Legendre points, each stored [dof][q]."""
    xq, _ = gl_rule(nq)
    V, D = lagrange(gll_points(degree), xq)
    return np.concatenate([V.ravel(), D.ravel()])


def weights3(nq: int) -> np.ndarray:
    _, w = gl_rule(nq)
    return np.einsum("a,b,c->abc", w, w, w).ravel()


# --------------------------------------------------------------------------
# Wrappers over liboracle.so
# --------------------------------------------------------------------------
def fill_uniform(seed, n, lo=-1.0, hi=1.0):
    out = np.empty(n)
    lib().orc_fill_uniform(seed, n, lo, hi, out)
    return out


def gemm_blocked(k, m, n, order, MB, NB, a, b):
    c = np.zeros(m * n)
    lib().orc_gemm_blocked(k, m, n, order, MB, NB, np.ascontiguousarray(a),
                           np.ascontiguousarray(b), c)
    return c


def gemm_micro(k, m, nc, lda, ldb, ldc, order, a, b, c0):
    c = np.array(c0, dtype=np.float64, copy=True)
    lib().orc_gemm_micro(k, m, nc, lda, ldb, ldc, order,
                         np.ascontiguousarray(a), np.ascontiguousarray(b), c)
    return c


def transform(G, w0, w1, w2):
    w0, w1, w2 = (np.array(w, dtype=np.float64, copy=True) for w in (w0, w1, w2))
    lib().orc_transform(len(w0), np.ascontiguousarray(G), w0, w1, w2)
    return w0, w1, w2


def mass_literal(P, phi, U, detJ, num_cells):
    W = np.zeros(num_cells * (P + 1) ** 3)
    lib().orc_mass_literal(P, np.ascontiguousarray(phi), np.ascontiguousarray(U),
                           W, np.ascontiguousarray(detJ), num_cells)
    return W


def stiffness_literal(P, phi2, U, G, num_cells):
    W = np.zeros(num_cells * (P + 1) ** 3)
    lib().orc_stiffness_literal(P, np.ascontiguousarray(phi2),
                                np.ascontiguousarray(U), W,
                                np.ascontiguousarray(G), num_cells)
    return W


def tmm_apply(Na, Nb, Tr, I0, I1, I2, B):
    A = np.zeros(Na ** 3)
    lib().orc_tmm_apply(Na, Nb, int(Tr), *(np.ascontiguousarray(t) for t in
                                           (I0, I1, I2, B)), A)
    return A


OPS = {"mass": 0, "stiffness": 1, "kpm": 2}


def _tab2(op, N, Q, tab):
    tab = np.ascontiguousarray(tab, dtype=np.float64)
    if tab.size == N * Q:
        tab = np.concatenate([tab, np.zeros(N * Q)])
    return tab


def cell_operator(op, N, Q, tab, U, G=None, detJw=None, num_cells=None):
    op = OPS.get(op, op)
    if num_cells is None:
        num_cells = len(U) // N ** 3
    W = np.zeros(num_cells * N ** 3)
    lib().orc_cell_operator(op, N, Q, _tab2(op, N, Q, tab),
                            np.ascontiguousarray(U), W, _np_or_null(G),
                            _np_or_null(detJw), num_cells)
    return W


def dofmap_structured(nx, ny, nz, P):
    N = P + 1
    dm = np.zeros(nx * ny * nz * N ** 3, dtype=np.int32)
    lib().orc_dofmap_structured(nx, ny, nz, P, dm)
    return dm


def geometry_affine(J, nq, want_G=True, want_detJw=True):
    J = np.ascontiguousarray(J, dtype=np.float64)
    C_ = len(J) // 9
    nq3 = nq ** 3
    G = np.zeros(C_ * nq3 * 6) if want_G else None
    D = np.zeros(C_ * nq3) if want_detJw else None
    lib().orc_geometry_affine(C_, nq3, J, weights3(nq), G, D)
    return G, D


def random_jacobians(seed, num_cells, eps=0.1):
    """J = I + eps * U(-1,1)^{3x3} per cell, row-major."""
    J = eps * uniform(seed, 9 * num_cells, -1.0, 1.0)
    J = J.reshape(num_cells, 3, 3)
    J += np.eye(3)
    return J.ravel()


def global_operator(op, N, Q, tab, x, dofmap, G=None, detJw=None,
                    num_cells=None):
    op = OPS.get(op, op)
    if num_cells is None:
        num_cells = len(dofmap) // N ** 3
    y = np.zeros(len(x))
    lib().orc_global_operator(op, N, Q, _tab2(op, N, Q, tab),
                              np.ascontiguousarray(x), y, len(x),
                              np.ascontiguousarray(dofmap, dtype=np.int32),
                              _np_or_null(G), _np_or_null(detJw), num_cells)
    return y


def global_diagonal(op, N, Q, tab, ndofs, dofmap, G=None, detJw=None,
                    num_cells=None):
    op = OPS.get(op, op)
    if num_cells is None:
        num_cells = len(dofmap) // N ** 3
    d = np.zeros(ndofs)
    lib().orc_global_diagonal(op, N, Q, _tab2(op, N, Q, tab), d, ndofs,
                              np.ascontiguousarray(dofmap, dtype=np.int32),
                              _np_or_null(G), _np_or_null(detJw), num_cells)
    return d


def pcg(op, N, Q, tab, f, dofmap, G=None, detJw=None, rtol=1e-10, maxit=10000):
    op = OPS.get(op, op)
    u = np.zeros(len(f))
    rel = C.c_double(0.0)
    it = lib().orc_pcg(op, N, Q, _tab2(op, N, Q, tab), np.ascontiguousarray(f),
                       u, len(f), np.ascontiguousarray(dofmap, dtype=np.int32),
                       _np_or_null(G), _np_or_null(detJw),
                       len(dofmap) // N ** 3, rtol, maxit, C.byref(rel))
    return u, it, rel.value


# --------------------------------------------------------------------------
# The compiled reference (oracle/_ref), when present
# --------------------------------------------------------------------------
REF_DIR = os.path.join(HERE, "_ref")


def ref_available(name: str) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libref_{name}.so"))


@lru_cache(maxsize=None)
def ref_lib(name: str):
    L = C.CDLL(os.path.join(REF_DIR, f"libref_{name}.so"))
    i32, lng = C.c_int, C.c_long
    if name == "mass":
        L.ref_mass_operator_f64.argtypes = [_dp, _dp, _dp, _dp, lng, i32, i32,
                                            i32, i32]
        L.ref_string2order.argtypes = [C.c_char_p]
        L.ref_string2order.restype = i32
    elif name == "stiffness":
        L.ref_stiffness_operator_f64.argtypes = [_dp, _dp, _dp, _dp, i32, i32,
                                                 i32, i32, i32]
    elif name == "kernels":
        L.ref_gemm_blocked.argtypes = [i32, i32, i32, i32, _dp, _dp, _dp]
        L.ref_gemm_micro_26_4_16.argtypes = [i32, _dp, _dp, _dp]
        L.ref_tmm_apply.argtypes = [i32, i32, i32, _dp, _dp, _dp, _dp, _dp]
    elif name.startswith("baseline"):
        L.ref_baseline_stiffness.argtypes = [i32, i32, i32, i32, _dp, _dp, _dp,
                                             _dp, lng]
    return L


def baseline_variant() -> str:
    """Pick the x86-64 level the host supports for the baseline build."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        if " avx512f" in flags and " avx512vl" in flags and " avx512bw" in flags:
            return "baseline_v4"
    except OSError:
        pass
    return "baseline_v3"
