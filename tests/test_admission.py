"""The request server's admission, queueing and eviction (admission.cpp),
driven by femtc_admission_simulate with a fake clock and a fake allocator:
no device work.  The real server runs the same code under its lock."""
import numpy as np
import pytest

from femtc_amd import _lib
from femtc_amd.server import simulate_admission

APPLY, SOLVE = _lib.REQUEST_APPLY, _lib.REQUEST_SOLVE
MB = 1 << 20


def solve(t, p, bytes_=100 * MB, service=1.0, create=0.5, inner=_lib.F32,
          actual=None):
    return dict(t_arrival=t, kind=SOLVE, problem=p, inner=inner, bytes=bytes_,
                actual=bytes_ if actual is None else actual, service=service,
                create=create)


def apply(t, p, service=0.1):
    return dict(t_arrival=t, kind=APPLY, problem=p, service=service)


def no_overlap_per_slot(res):
    by = {}
    for r in res:
        if r["status"] == _lib.OK:
            by.setdefault(r["slot"], []).append((r["t_start"], r["t_end"]))
    for iv in by.values():
        iv.sort()
        for (a0, a1), (b0, b1) in zip(iv, iv[1:]):
            assert b0 >= a1 - 1e-12


def test_resident_workspace_is_reused():
    res, st = simulate_admission([solve(0, 0), solve(0.1, 0)], 1, 1000 * MB)
    assert [r["cached"] for r in res] == [0, 1]
    assert st["created"] == 1 and st["evictions"] == 0
    # the second waits for the slot, then runs without the creation cost
    assert res[1]["t_start"] == pytest.approx(1.5)
    assert res[1]["t_end"] == pytest.approx(2.5)


def test_workspaces_are_per_inner_precision():
    res, st = simulate_admission(
        [solve(0, 0, inner=_lib.F32), solve(0.1, 0, inner=_lib.BF16)], 1,
        1000 * MB)
    assert [r["cached"] for r in res] == [0, 0] and st["created"] == 2


def test_lru_eviction_order():
    # room for two workspaces, one slot: A B A C A B
    seq = [0, 1, 0, 2, 0, 1]
    reqs = [solve(2.0 * i, p) for i, p in enumerate(seq)]
    res, st = simulate_admission(reqs, 1, 200 * MB)
    assert [r["cached"] for r in res] == [0, 0, 1, 0, 1, 0]
    # C evicted B (least recently used), then B evicted C
    assert st["evictions"] == 2
    assert st["high_water"] <= 200 * MB


def test_a_request_that_does_not_fit_waits_while_later_ones_go_ahead():
    # slot 0 runs a long solve on P0 (its workspace busy); P1's solve cannot
    # fit beside it and waits; an apply arriving later goes ahead on slot 1
    reqs = [solve(0.0, 0, service=10.0), solve(0.1, 1), apply(0.2, 2)]
    res, st = simulate_admission(reqs, 2, 150 * MB)
    assert res[2]["t_start"] == pytest.approx(0.2)
    assert res[1]["t_start"] >= res[0]["t_end"] - 1e-12  # after P0 is idle
    assert st["evictions"] == 1  # P0's idle workspace made room for P1
    assert all(r["status"] == _lib.OK for r in res)


def test_oversized_solve_is_refused_without_blocking_others():
    reqs = [solve(0, 0, bytes_=500 * MB), apply(0, 1), solve(0.1, 2)]
    res, st = simulate_admission(reqs, 1, 200 * MB)
    assert res[0]["status"] == _lib.ERR_OUT_OF_MEMORY
    assert res[1]["status"] == _lib.OK and res[2]["status"] == _lib.OK
    assert st["left_in_queue"] == 0


def test_actual_bytes_replace_the_estimate():
    # estimates of 100 MB, workspaces of 60 MB: two stay resident in 160 MB
    reqs = [solve(0, 0, actual=60 * MB), solve(2, 1, actual=60 * MB),
            solve(4, 0, actual=60 * MB)]
    res, st = simulate_admission(reqs, 1, 160 * MB)
    assert st["evictions"] == 0 and res[2]["cached"] == 1
    # with the estimate charged for good the third would have evicted
    reqs = [solve(0, 0), solve(2, 1), solve(4, 0)]
    res, st = simulate_admission(reqs, 1, 160 * MB)
    assert st["evictions"] >= 1


def test_removed_problem_retires_its_workspaces():
    # P0 solving until t=1.5 when it is removed at t=1.0; its queued request
    # fails, its workspace is destroyed once the running request ends, and
    # other problems are unaffected
    reqs = [solve(0.0, 0), solve(0.2, 0), solve(0.3, 1), solve(2.0, 0)]
    res, st = simulate_admission(reqs, 1, 1000 * MB, nproblems=2,
                                 remove_at=[1.0, -1.0])
    assert res[0]["status"] == _lib.OK
    assert res[1]["status"] == _lib.ERR_INVALID_ARGUMENT  # dropped queued
    assert res[2]["status"] == _lib.OK and res[2]["cached"] == 0
    assert res[3]["status"] == _lib.ERR_INVALID_ARGUMENT  # after removal
    assert st["destroyed"] == 1 and st["created"] == 2


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_poisson_stream_invariants(seed):
    """A Poisson stream (fake clock): the budget is never exceeded, no slot
    runs two requests at once, nothing starts before it arrives, every
    request is served, and arrivals during busy periods are admitted as
    slots and memory free up (continuous batching)."""
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    for i in range(400):
        t += rng.exponential(1 / 500.0)
        p = int(rng.integers(0, 12))
        if rng.random() < 0.5:
            reqs.append(apply(t, p, service=float(rng.uniform(1e-4, 1e-3))))
        else:
            b = int(rng.integers(20, 200)) * MB
            reqs.append(solve(t, p, bytes_=b, actual=int(b * 0.8),
                              service=float(rng.uniform(1e-3, 5e-3)),
                              create=2e-3))
    budget = 600 * MB
    res, st = simulate_admission(reqs, 4, budget, nproblems=12)
    assert all(r["status"] == _lib.OK for r in res)
    assert st["high_water"] <= budget and st["left_in_queue"] == 0
    assert all(r["t_start"] >= q["t_arrival"] - 1e-12
               for r, q in zip(res, reqs))
    no_overlap_per_slot(res)
    # overlap: at some instant several slots run at once
    ev = sorted([(r["t_start"], 1) for r in res] +
                [(r["t_end"], -1) for r in res])
    level, top = 0, 0
    for _, d in ev:
        level += d
        top = max(top, level)
    assert top >= 2
    assert sum(r["cached"] for r in res) > 0


def test_simulator_rejects_bad_requests():
    with pytest.raises(_lib.FemtcError):
        simulate_admission([solve(0, 3)], 1, MB, nproblems=2)
    with pytest.raises(_lib.FemtcError):
        simulate_admission([solve(-1.0, 0)], 1, MB)
