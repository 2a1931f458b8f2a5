"""Synthetic problems for benchmarks and tests: structured hex meshes with
per-cell affine Jacobians J = I + 0.1 U(-1,1), seeded input vectors, and the
uneven batch of many meshes (lognormal cell counts, mixed degrees and
operators).  All data is generated from a counter-based SplitMix64 stream so
the same seed gives the same bits everywhere."""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from .mesh import basis_tables, dofmap_structured, geometry_affine


def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0):
    """n doubles in [lo, hi) from SplitMix64 at counters 1..n of `seed`."""
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * g
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def jacobians(seed: int, num_cells: int, eps: float = 0.1):
    """Per-cell J = I + eps U(-1,1)^{3x3}, row-major [C][9]."""
    J = eps * uniform(seed, 9 * num_cells).reshape(num_cells, 3, 3)
    J += np.eye(3)
    return J.ravel()


def box_dims(cells: int):
    """A near-cubic nx*ny*nz box with about `cells` cells."""
    s = max(1, int(round(cells ** (1.0 / 3.0))))
    nz = max(1, int(math.ceil(cells / (s * s))))
    return s, s, nz


def structured_problem(torch, dims, degree: int, nq: int, op: int, dtype,
                       seed: int, continuous: bool = True,
                       device: str = "cuda"):
    """One problem on the device as a dict (the keys femtc_amd.Batch takes).
    dtype is a torch dtype (float64 or float32); geometry is computed in FP64
    and stored in dtype."""
    nx, ny, nz = dims
    C = nx * ny * nz
    N = degree + 1
    tab = torch.from_numpy(basis_tables(degree, nq)).to(device, dtype)
    J = torch.from_numpy(jacobians(seed, C)).to(device)
    G = D = None
    if op != _lib.OP_MASS:
        G = torch.empty(C * nq ** 3 * 6, device=device, dtype=dtype)
    if op != _lib.OP_STIFFNESS:
        D = torch.empty(C * nq ** 3, device=device, dtype=dtype)
    geometry_affine(J, nq, G=G, detJw=D)
    item = {"op": op, "degree": degree, "nq": nq, "tables": tab, "G": G,
            "detJw": D, "J": J, "num_cells": C, "dims": (nx, ny, nz)}
    if continuous:
        ng = (nx * degree + 1) * (ny * degree + 1) * (nz * degree + 1)
        dm = torch.empty(C * N ** 3, device=device, dtype=torch.int32)
        dofmap_structured(nx, ny, nz, degree, dm)
        item.update(dofmap=dm, ndofs=ng,
                    x=torch.from_numpy(uniform(seed + 1, ng)).to(device, dtype),
                    y=torch.empty(ng, device=device, dtype=dtype))
    else:
        n = C * N ** 3
        item.update(dofmap=None, ndofs=0,
                    x=torch.from_numpy(uniform(seed + 1, n)).to(device, dtype),
                    y=torch.empty(n, device=device, dtype=dtype))
    return item


def uneven_batch_spec(seed: int = 0, n_meshes: int = 200, mu=math.log(1e4),
                      sigma: float = 1.0, lo: int = 1000, hi: int = 100000,
                      degrees=(2, 3, 4, 5, 6, 7, 8)):
    """The uneven batch: cell counts lognormal(mu, sigma) clipped to
    [lo, hi], degree uniform over `degrees` with nq = degree+1, operators
    alternating mass / stiffness.  Returns (cells, degree, op) per mesh."""
    rng = np.random.default_rng(seed)
    cells = np.clip(np.exp(rng.normal(mu, sigma, n_meshes)), lo, hi)
    degs = rng.choice(np.asarray(degrees), n_meshes)
    ops = [_lib.OP_MASS if i % 2 == 0 else _lib.OP_STIFFNESS
           for i in range(n_meshes)]
    return [(int(c), int(p), o) for c, p, o in zip(cells, degs, ops)]


def algorithmic_bytes(item, dtype_size: int) -> int:
    """Compulsory traffic of one apply: x and y once, dofmap and geometry
    once."""
    C, N, Q = item["num_cells"], item["degree"] + 1, item["nq"]
    geo = 0
    if item["op"] != _lib.OP_MASS:
        geo += 6 * C * Q ** 3
    if item["op"] != _lib.OP_STIFFNESS:
        geo += C * Q ** 3
    if item.get("dofmap") is not None:
        return dtype_size * (2 * item["ndofs"] + geo) + 4 * C * N ** 3
    return dtype_size * (2 * C * N ** 3 + geo)
