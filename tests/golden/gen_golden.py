"""Generate golden vectors from the compiled reference (oracle/_ref).

Run in a container where the reference is mounted and ``make -C oracle ref``
has been run.  Writes ``ref_goldens.npz`` beside this script: inputs and the
reference's outputs for small cases of

* ``linalg::gemm_blocked`` (src/linalg.hpp:101-182) for several degrees,
  all six loop orders and block sizes,
* ``linalg::gemm_micro`` at the tests/test_micro.cpp shape (k=26, m=4, n=16),
* ``apply_contractions`` (src/tmm.hpp:91-126) for square and non-square shapes,
* ``operators::mass_operator<double>`` (src/mass.hpp:196-219), literal,
  degrees 1..4, random and symmetric tables,
* ``operators::stiffness_operator<double>`` (src/stiffness.hpp:245-268),
  literal, degrees 1..4, with the full [phi | dphi] table,
* ``linalg::string2order`` (src/linalg.hpp:21-37) and the dispatch-ladder
  behaviour for unsupported degree / block size (src/mass.hpp:134-136,
  190-192, 215-217): W is left untouched and a message is printed.

Inputs are drawn from the counter-based generator in oracle/oracle.py, in
[0, 1) like the reference drivers' ``rand()/RAND_MAX`` (bench_mass.cpp:74-88).
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def _capture_stdout(fn):
    """Run fn() and return what the C side wrote to fd 1."""
    sys.stdout.flush()
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            fn()
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


def main():
    out = {}
    kern = O.ref_lib("kernels")
    seed = 1000

    # gemm_blocked
    gb_cases = []
    for P in (1, 2, 3, 4, 6):
        for order in range(6):
            for (MB, NB) in ((0, 0), (8, 16), (4, 4)):
                gb_cases.append((P, order, MB, NB))
    for P in (8, 12):
        for (MB, NB) in ((0, 0), (8, 16)):
            gb_cases.append((P, 4, MB, NB))
    gb_cases.append((25, 4, 8, 16))  # the tests/test_gemm.cpp:41 blocking
    for idx, (P, order, MB, NB) in enumerate(gb_cases):
        N = P + 1
        a = O.uniform(seed, N * N, 0.0, 1.0); seed += 1
        b = O.uniform(seed, N ** 3, 0.0, 1.0); seed += 1
        c = np.zeros(N ** 3)
        rc = kern.ref_gemm_blocked(P, order, MB, NB, a, b, c)
        assert rc == 0
        out[f"gb_{idx}_meta"] = np.array([P, order, MB, NB])
        out[f"gb_{idx}_a"] = a
        out[f"gb_{idx}_b"] = b
        out[f"gb_{idx}_c"] = c
    out["gb_count"] = np.array([len(gb_cases)])

    # gemm_micro (accumulates into c)
    for order in range(6):
        a = O.uniform(seed, 4 * 26, 0.0, 1.0); seed += 1
        b = O.uniform(seed, 26 * 16, 0.0, 1.0); seed += 1
        c0 = O.uniform(seed, 4 * 16, 0.0, 1.0); seed += 1
        c = c0.copy()
        kern.ref_gemm_micro_26_4_16(order, a, b, c)
        out[f"gm_{order}_a"], out[f"gm_{order}_b"] = a, b
        out[f"gm_{order}_c0"], out[f"gm_{order}_c"] = c0, c

    # apply_contractions
    tmm_shapes = [(2, 2), (3, 2), (2, 3), (4, 4), (5, 4), (4, 5), (6, 5),
                  (5, 6), (7, 7)]
    k = 0
    for (Na, Nb) in tmm_shapes:
        for Tr in (0, 1):
            I = [O.uniform(seed + j, Na * Nb, -1.0, 1.0) for j in range(3)]
            seed += 3
            B = O.uniform(seed, Nb ** 3, -1.0, 1.0); seed += 1
            A = np.zeros(Na ** 3)
            assert kern.ref_tmm_apply(Na, Nb, Tr, I[0], I[1], I[2], B.copy(),
                                      A) == 0
            out[f"tmm_{k}_meta"] = np.array([Na, Nb, Tr])
            for j in range(3):
                out[f"tmm_{k}_I{j}"] = I[j]
            out[f"tmm_{k}_B"], out[f"tmm_{k}_A"] = B, A
            k += 1
    out["tmm_count"] = np.array([k])

    # mass / stiffness literal through the public dispatch ladders
    mass = O.ref_lib("mass")
    stiff = O.ref_lib("stiffness")
    k = 0
    ncell = 3
    for P in (1, 2, 3, 4):
        N = P + 1
        nd = N ** 3
        for sym in (0, 1):
            phi = O.uniform(seed, N * N, 0.0, 1.0); seed += 1
            if sym:
                phi = 0.5 * (phi.reshape(N, N) + phi.reshape(N, N).T).ravel()
            U = O.uniform(seed, ncell * nd, 0.0, 1.0); seed += 1
            detJ = O.uniform(seed, ncell * nd, 0.0, 1.0); seed += 1
            for (order, Mb, Nb) in ((0, 0, 0), (4, 8, 16), (5, 4, 4)):
                W = np.zeros(ncell * nd)
                mass.ref_mass_operator_f64(phi.copy(), U.copy(), W,
                                           detJ.copy(), ncell, P, order, Mb, Nb)
                out[f"mass_{k}_meta"] = np.array([P, sym, order, Mb, Nb, ncell])
                out[f"mass_{k}_phi"], out[f"mass_{k}_U"] = phi, U
                out[f"mass_{k}_detJ"], out[f"mass_{k}_W"] = detJ, W
                k += 1
    out["mass_count"] = np.array([k])

    k = 0
    for P in (1, 2, 3, 4):
        N = P + 1
        nd = N ** 3
        phi2 = O.uniform(seed, 2 * N * N, 0.0, 1.0); seed += 1
        U = O.uniform(seed, ncell * nd, 0.0, 1.0); seed += 1
        G = O.uniform(seed, ncell * nd * 6, 0.0, 1.0); seed += 1
        for (order, Mb, Nb) in ((0, 0, 0), (4, 8, 16)):
            W = np.zeros(ncell * nd)
            stiff.ref_stiffness_operator_f64(phi2.copy(), U.copy(), W,
                                             G.copy(), ncell, P, order, Mb, Nb)
            out[f"stiff_{k}_meta"] = np.array([P, order, Mb, Nb, ncell])
            out[f"stiff_{k}_phi2"], out[f"stiff_{k}_U"] = phi2, U
            out[f"stiff_{k}_G"], out[f"stiff_{k}_W"] = G, W
            k += 1
    out["stiff_count"] = np.array([k])

    # string2order
    names = ["ijk", "ikj", "jik", "jki", "kij", "kji", "xyz", "IJK", ""]
    out["s2o_names"] = np.array(names)
    out["s2o_codes"] = np.array([mass.ref_string2order(s.encode())
                                 for s in names])

    # ladder behaviour for unsupported arguments: message + W untouched
    msgs = []
    cases = [(26, 0, 0, 0), (0, 0, 0, 0), (2, 0, 3, 0), (2, 0, 0, 5)]
    for (deg, order, Mb, Nb) in cases:
        N = max(deg, 1) + 1
        phi = np.full(N * N, 0.5)
        U = np.ones(N ** 3)
        detJ = np.ones(N ** 3)
        W = np.full(N ** 3, 7.0)
        text = _capture_stdout(lambda: mass.ref_mass_operator_f64(
            phi, U, W, detJ, 1, deg, order, Mb, Nb))
        assert np.all(W == 7.0)
        msgs.append(text.strip())
    out["ladder_cases"] = np.array(cases)
    out["ladder_msgs"] = np.array(msgs)

    path = os.path.join(HERE, "ref_goldens.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
