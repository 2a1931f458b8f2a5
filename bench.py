"""Benchmark: FP64 stiffness operator apply y = K x through the dofmap.

Workload (the metric's): structured 48^3 hexahedral mesh (110592 cells),
degree P=6 Lagrange on GLL nodes, Q=7 Gauss-Legendre points per direction,
int32 dofmap with shared faces (289^3 = 24,137,569 global dofs), per-cell
affine Jacobians J = I + 0.1 U(-1,1) with G = detJ w J^-1 J^-T per point in
FP64, x ~ U[-1,1].

One step = one continuous operator application: clear y (fill kernel), then
the fused gather -> contraction chain -> geometry -> transposed chain ->
atomic scatter kernel.  Inputs are resident in HBM before the timed region
starts.  `value` is global dofs per second of the timed steps (GDoF/s).

Beside the metric line's required fields it reports, on the same GPU:
  roofline   the apply kernel against the MI355X HBM peak with algorithmic
             bytes (x + y + dofmap + G once each) over the kernel's own time
             from HIP events on its stream; `traffic` = HBM bytes per launch
             from the committed FETCH_SIZE / WRITE_SIZE counter profile
  cpu_baseline  the reference's own stiffness kernel (src/stiffness.hpp:34-99)
             built -Ofast from the reference sources (oracle/_ref) on the host
             cores, on a bounded sample of cells of the same shape
  solve      (K+M)u = f on the same mesh to 1e-10: FP64 Jacobi-PCG vs FP32
             inner PCG with FP64 defect correction (solves/s, error), each
             also with per-cell geometry and with the structured grid given
             (assembled scatter)
  apply_p4_64cube  the P=4 64^3 stiffness apply (GDoF/s, roofline fraction)
  mass_c1    the cell-local P=3 mass apply on 32768 cells, cache-resident
             and streamed over 16 rotating buffer sets
  batch      200 uneven meshes in one batched, graph-captured launch set
             (batch_groups, when named: each group of it alone)
  serving    a 500 req/s Poisson stream of applies and solves (p50/p99)

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--apply-only]
                  [--only PART] [--no-cpu-baseline] [--copy-calibration MiB]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GDoF/s stiffness apply P=6 FP64 vs BF16-MFMA+refine @1e-10; p99 solve ms"
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles",
                               "pmc_apply_stiffness_p6_q7_48cube_f64.txt")


def _uniform(seed, n, lo, hi):
    """Counter-based SplitMix64 uniforms (the workloads' generator)."""
    from femtc_amd.workloads import uniform
    return uniform(seed, n, lo, hi)


def measured_traffic():
    """HBM bytes per apply launch from the committed counter profile."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*(\d+)", text)
    return int(m.group(1)) if m else None


def cpu_baseline(P: int, ncells_global: int, ndofs_global: int,
                 budget_s: float = 15.0) -> dict:
    """Reference stiffness kernel on the host cores, one thread per core on
    disjoint cell ranges (ctypes releases the GIL): applies repeated until
    about budget_s of CPU time (at least 10), mean of the second half
    (bench_stiffness.cpp:95-106 times 10 and keeps the last 5).  Reported as
    global-dof throughput of the same mesh: cells/s * (ndofs_global /
    ncells_global)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # checker / baseline infrastructure only

    variant = O.baseline_variant()
    if not O.ref_available(variant):
        return {"value": None, "unit": "GDoF/s", "cores": 0,
                "kind": "reference", "sample": "oracle/_ref not built"}
    L = O.ref_lib(variant)
    N = P + 1
    nd = N ** 3
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    phi2 = np.concatenate([O.basis_tables(P, N)])  # Q = P+1 as the reference
    # pick the fastest loop order / block on one core (run.sh:13-23 sweep)
    probe = 256
    U = _uniform(5, probe * nd, -1, 1)
    G = _uniform(6, probe * nd * 6, 0, 1)
    W = np.zeros(probe * nd)
    best = None
    for order in range(6):
        for (mb, nb) in ((0, 0), (4, 16), (8, 8), (0, 16)):
            t0 = time.perf_counter()
            rc = L.ref_baseline_stiffness(P, order, mb, nb, phi2, U, W, G, probe)
            dt = time.perf_counter() - t0
            if rc == 0 and (best is None or dt < best[0]):
                best = (dt, order, mb, nb)
    _, order, mb, nb = best
    per_cell = best[0] / probe
    # the whole mesh split evenly over the threads, unless 10 applies would
    # exceed about budget_s of CPU time
    per_thread = -(-ncells_global // cores)
    per_thread = int(max(64, min(per_thread, budget_s / (10 * per_cell))))
    bufs = []
    for t in range(cores):
        bufs.append((_uniform(10 + t, per_thread * nd, -1, 1),
                     np.zeros(per_thread * nd),
                     _uniform(50 + t, per_thread * nd * 6, 0, 1)))
    times = []
    while len(times) < 10 or sum(times) * cores < budget_s:
        barrier = threading.Barrier(cores + 1)
        done = threading.Barrier(cores + 1)

        def work(t):
            u, w, g = bufs[t]
            barrier.wait()
            L.ref_baseline_stiffness(P, order, mb, nb, phi2, u, w, g, per_thread)
            done.wait()

        ths = [threading.Thread(target=work, args=(t,)) for t in range(cores)]
        for th in ths:
            th.start()
        barrier.wait()
        t0 = time.perf_counter()
        done.wait()
        times.append(time.perf_counter() - t0)
        for th in ths:
            th.join()
    elapsed = float(np.mean(times[len(times) // 2:]))
    cells_per_s = cores * per_thread / elapsed
    names = ("ijk", "ikj", "jik", "jki", "kij", "kji")
    return {
        "value": cells_per_s * ndofs_global / ncells_global / 1e9,
        "unit": "GDoF/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"reference stiffness_operator<double,P={P}> "
                   f"(src/stiffness.hpp:34-99, -Ofast {variant}), "
                   f"{cores} threads x {per_thread} cells, order={names[order]} "
                   f"Mb={mb} Nb={nb}, {len(times)} applies "
                   f"({sum(times) * cores:.1f} s of CPU time), mean of the "
                   f"last {len(times) - len(times) // 2} = "
                   f"{elapsed * 1e3:.1f} ms; {cells_per_s:.4g} cells/s scaled "
                   f"to the mesh's global dofs per cell"),
        "cell_local_gdofs": cells_per_s * nd / 1e9,
    }


def structured(torch, F, n, P, want_detJw=False):
    """Resident FP64 problem on an n^3 mesh at degree P, Q = P+1."""
    dev = torch.device("cuda:0")
    Q, N, C = P + 1, P + 1, n ** 3
    ng = (n * P + 1) ** 3
    tab = torch.from_numpy(F.basis_tables(P, Q)).to(dev)
    J = _uniform(3, 9 * C, -1, 1) * 0.1
    J = (J.reshape(C, 3, 3) + np.eye(3)).ravel()
    dJ = torch.from_numpy(J).to(dev)
    G = torch.empty(C * Q ** 3 * 6, device=dev, dtype=torch.float64)
    D = (torch.empty(C * Q ** 3, device=dev, dtype=torch.float64)
         if want_detJw else None)
    F.geometry_affine(dJ, Q, G=G, detJw=D)
    dm = torch.empty(C * N ** 3, device=dev, dtype=torch.int32)
    F.dofmap_structured(n, n, n, P, dm)
    return {"tab": tab, "G": G, "D": D, "dm": dm, "C": C, "ng": ng, "P": P,
            "Q": Q, "N": N, "J": dJ, "n": n}


# Applies run at setup, before the warm-up steps.  On a freshly started GPU
# the first ~40 applies of this HBM-heavy kernel run 5-25 % slow (measured:
# launches 1-3 at ~500 us, 4-20 at up to ~640 us, settled at ~505 us from
# about the 40th), the same after 400 ms of unrelated GPU work or a read of
# every input buffer: the device's power management settling under this
# load, not first-touch of the buffers.  Settling here lets a short --warmup
# measure the steady state.
SETTLE_APPLIES = 64


def time_apply(torch, F, pr, steps, warmup, x, y, settle=SETTLE_APPLIES):
    """Per step: fill y, y += K x.  Returns (wall ms/step, kernel ms)."""
    stream = torch.cuda.current_stream()

    def apply():
        F.global_apply(F.OP_STIFFNESS, pr["tab"], x, y, pr["dm"], pr["C"],
                       pr["P"], pr["Q"], G=pr["G"], accumulate=True)

    for _ in range(settle):
        F.fill_zero(y)
        apply()
    torch.cuda.synchronize()
    for _ in range(warmup):
        F.fill_zero(y)
        apply()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        F.fill_zero(y)
        ev[i][0].record(stream)
        apply()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in ev]))
    return wall * 1e3 / steps, kern_ms


def stream_bandwidth(torch, F, n=64 << 20, reps=10):
    """The practical HBM ceiling from the library's hand-written streaming
    probe (femtc_stream_probe, 16-byte non-temporal loads), on arrays of n
    doubles (512 MiB each), mean over reps after two warm-up launches, the
    best over grids of 1, 2, 4 and 8 workgroups per CU:
      read5_write1: five arrays read, one written (the apply's mix, 3 GiB)
      copy:         one read, one written (1 GiB)
      read_only:    six arrays read (3 GiB)"""
    from femtc_amd import _lib
    src = torch.ones(6 * n, device="cuda", dtype=torch.float64)
    dst = torch.empty(n, device="cuda", dtype=torch.float64)
    sink = torch.empty(1 << 16, device="cuda", dtype=torch.float64)
    stream = torch.cuda.current_stream()
    L = _lib.lib()
    out = {}
    for name, nr, nw in (("read5_write1", 5, 1), ("copy", 1, 1),
                         ("read_only", 6, 0)):
        best, best_wg = 0.0, 0
        for wg in (1, 2, 4, 8):
            def go():
                _lib.check(L.femtc_stream_probe(nr, nw, src.data_ptr(),
                                                dst.data_ptr(), n,
                                                sink.data_ptr(), sink.numel(),
                                                wg, stream.cuda_stream),
                           "stream_probe")
            go()
            go()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                go()
            e1.record(stream)
            torch.cuda.synchronize()
            sec = e0.elapsed_time(e1) * 1e-3 / reps
            gbs = 8 * n * (nr + nw) / sec / 1e9
            if gbs > best:
                best, best_wg = gbs, wg
        out[name + "_GBs"] = best
        out[name + "_wg_per_cu"] = best_wg
    del src, dst, sink
    torch.cuda.empty_cache()
    return out

def alg_bytes_stiffness(pr):
    return 8 * pr["ng"] * 2 + 4 * pr["C"] * pr["N"] ** 3 + 48 * pr["C"] * pr["Q"] ** 3


def bench_solve(torch, F, n=48, P=6):
    """(K+M)u = f to 1e-10: the plain FP64 PCG (the reference's data: per-
    point geometry, dofmap, atomic scatter) against FP32 inner PCG + FP64
    refinement, each also with per-cell geometry (affine cells) and with the
    structured grid given (the assembled or z-run scatter when it times
    faster)."""
    pr = structured(torch, F, n, P, want_detJw=True)
    f = torch.from_numpy(_uniform(7, pr["ng"], -1, 1)).cuda()
    out = {"workload": f"(K+M)u=f, {n}^3 mesh, P={P}, Q={P + 1}, f~U[-1,1], "
                       "rtol 1e-10 (FP64 residual)"}
    sols = {}
    grid = (n, n, n)
    names = {F.SCATTER_DIRECT: "atomic", F.SCATTER_ASSEMBLED: "assembled",
             F.SCATTER_ZRUN: "z-run"}
    inner_names = {F.F64: "fp64", F.F32: "fp32", F.BF16: "bf16_mfma_split",
                   F.BF16_ROUNDED: "bf16_mfma_rounded"}
    for name, inner, J, g, irtol in (
            ("fp64_pcg", F.F64, None, None, 1e-4),
            ("fp32_inner_fp64_refine", F.F32, None, None, 1e-4),
            ("fp64_pcg_cell_geometry", F.F64, pr["J"], None, 1e-4),
            ("fp32_inner_cell_geometry_fp64_refine", F.F32, pr["J"], None,
             1e-4),
            ("fp64_pcg_structured", F.F64, None, grid, 1e-4),
            ("fp64_pcg_cell_geometry_structured", F.F64, pr["J"], grid, 1e-4),
            ("fp32_inner_cell_geometry_structured_fp64_refine", F.F32,
             pr["J"], grid, 1e-4),
            ("bf16_mfma_split_inner_fp64_refine", F.BF16, pr["J"], grid,
             1e-4),
            ("bf16_mfma_rounded_inner_fp64_refine", F.BF16_ROUNDED, pr["J"],
             grid, 1e-4),
            ("bf16_mfma_rounded_inner_rtol3e-3_fp64_refine", F.BF16_ROUNDED,
             pr["J"], grid, 3e-3)):
        with F.PCGSolver(F.OP_KPM, pr["tab"], pr["dm"], pr["C"], P, P + 1,
                         pr["ng"], G=pr["G"], detJw=pr["D"], inner=inner,
                         J=J, grid=g) as S:
            u = torch.empty(pr["ng"], device="cuda", dtype=torch.float64)
            # warm-up (captures the batch graph)
            S.solve(f, u, rtol=1e-10, inner_rtol=irtol)
            infos = [S.solve(f, u, rtol=1e-10, inner_rtol=irtol)
                     for _ in range(2)]
            scatter = names[S.scatter()]
            even_odd = S.even_odd()
        ms = float(np.mean([i["seconds"] for i in infos])) * 1e3
        out[name] = {"ms_per_solve": ms, "solves_per_s": 1e3 / ms,
                     "inner": inner_names[inner], "inner_rtol": irtol,
                     "iterations": infos[-1]["inner_iterations"],
                     "outer": infos[-1]["outer_iterations"],
                     "relres": infos[-1]["relres"], "scatter": scatter,
                     "even_odd": even_odd}
        sols[name] = u
    u64 = sols["fp64_pcg"]
    base = out["fp64_pcg"]["ms_per_solve"]
    best64 = min(v["ms_per_solve"] for k, v in out.items()
                 if k.startswith("fp64_pcg"))
    for name, u in sols.items():
        if name != "fp64_pcg":
            out[name]["error_vs_fp64_solution"] = float((u - u64).norm() /
                                                        u64.norm())
            out[name]["speedup_vs_fp64_pcg"] = base / out[name]["ms_per_solve"]
    # the headline refined solve: the fastest one that reached 1e-10
    refined = [k for k, v in out.items() if isinstance(v, dict)
               and v.get("inner", "fp64") != "fp64" and v["relres"] <= 1e-10]
    best = min(refined, key=lambda k: out[k]["ms_per_solve"])
    out["best_refined"] = best
    out["speedup_refined_vs_fp64"] = out[best]["speedup_vs_fp64_pcg"]
    out["speedup_refined_vs_fastest_fp64"] = best64 / out[best]["ms_per_solve"]
    bf = [k for k in refined if out[k]["inner"].startswith("bf16")]
    if bf:
        kb = min(bf, key=lambda k: out[k]["ms_per_solve"])
        out["bf16_mfma_inner"] = {
            "best": kb, "ms_per_solve": out[kb]["ms_per_solve"],
            "outer": out[kb]["outer"], "iterations": out[kb]["iterations"],
            "error_vs_fp64_solution": out[kb]["error_vs_fp64_solution"],
            "speedup_vs_fp64_pcg": out[kb]["speedup_vs_fp64_pcg"],
            "speedup_vs_fastest_fp64": best64 / out[kb]["ms_per_solve"]}
    return out


def bench_mfma_apply(torch, F, n=48, P=6, reps=20):
    """The BF16-MFMA (K+M) apply on the solve's mesh (48^3, P=6, Q=7, affine
    cells, FP32 vectors): cell pass on the matrix cores + assembly, split =
    1 (tables and values hi + lo) and 0 (values rounded).  MFMA flops issued
    count every v_mfma_f32_16x16x32_bf16 at 16*16*32*2; the useful flops are
    the sum-factorized operator's (92,267 per cell at P=6)."""
    Q, C = P + 1, n ** 3
    ng = (n * P + 1) ** 3
    J = torch.from_numpy((_uniform(3, 9 * C, -1, 1) * 0.1).reshape(C, 3, 3)
                         + np.eye(3)).reshape(-1).cuda()
    cg = torch.empty(C * 7, device="cuda", dtype=torch.float32)
    F.geometry_cell(J, cg)
    x = torch.from_numpy(_uniform(4, ng, -1, 1)).float().cuda()
    y = torch.empty(ng, device="cuda", dtype=torch.float32)
    lat = torch.empty(C * (P + 1) ** 3, device="cuda", dtype=torch.float32)
    tab = F.basis_tables(P, Q)
    out = {"workload": f"(K+M) apply, {n}^3 mesh, P={P}, Q={Q}, affine cells, "
                       "bfloat16 MFMA contractions, FP32 accumulate"}
    useful = 92267.0 * C
    for split, mfmas in ((1, 72), (0, 48)):
        ops = torch.empty(F.mfma_tables_bytes(split), device="cuda",
                          dtype=torch.uint8)
        F.mfma_tables(P, Q, tab, ops, split=split)
        for _ in range(5):
            F.global_apply_mfma(F.OP_KPM, P, Q, ops, x, y, (n, n, n), cg, lat,
                                split=split)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            F.global_apply_mfma(F.OP_KPM, P, Q, ops, x, y, (n, n, n), cg, lat,
                                split=split)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        issued = mfmas * 16 * 16 * 32 * 2 * C
        out[f"split{split}"] = {
            "us_per_apply": us, "gdofs": ng / (us * 1e-6) / 1e9,
            "mfma_per_cell": mfmas,
            "mfma_tflops_issued": issued / (us * 1e-6) / 1e12,
            "useful_tflops": useful / (us * 1e-6) / 1e12}
    return out


def plan_scatters(torch, F, pr, x, y, steps=30, warmup=60):
    """The apply through a plan on the structured dofmap: every scatter
    timed on its own (ms per overwriting apply, clearing included), then the
    autotuned choice."""
    n, P, Q = pr["n"], pr["P"], pr["Q"]
    ab = alg_bytes_stiffness(pr)
    plan = F.Plan(n, n, n, P, pr["dm"], pr["ng"])
    stream = torch.cuda.current_stream()
    names = {(0, 0, 0): "atomic", (-1, -1, -1): "assembled",
             (-2, -2, -2): "structured atomic",
             (-3, -3, -3): "structured z-run"}

    def timed():
        for _ in range(warmup):
            plan.apply(F.OP_STIFFNESS, pr["tab"], x, y, Q, G=pr["G"])
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            plan.apply(F.OP_STIFFNESS, pr["tab"], x, y, Q, G=pr["G"])
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    by_scatter = {}
    for shape, name in names.items():
        plan.set_block(Q, F.F64, F.OP_STIFFNESS, shape)
        t = timed()
        by_scatter[name] = {"ms_per_apply": t, "roofline_frac":
                            ab / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
    shape, _ = plan.autotune(F.OP_STIFFNESS, pr["tab"], x, y, Q, G=pr["G"],
                             reps=5)
    pms = timed()
    # the chosen scatter's operator kernel alone: accumulating applies
    # (y += Kx) need no clear of y
    for _ in range(warmup):
        plan.apply(F.OP_STIFFNESS, pr["tab"], x, y, Q, G=pr["G"],
                   accumulate=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        plan.apply(F.OP_STIFFNESS, pr["tab"], x, y, Q, G=pr["G"],
                   accumulate=True)
    e1.record(stream)
    torch.cuda.synchronize()
    kms = e0.elapsed_time(e1) / steps
    plan.close()
    return {"scatter": names.get(tuple(shape), f"blocked {tuple(shape)}"),
            "ms_per_apply": pms,
            "gdofs": pr["ng"] / (pms * 1e-3) / 1e9,
            "roofline_frac": ab / (pms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernel_ms": kms,
            "kernel_roofline_frac": ab / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "by_scatter": by_scatter}


def bench_apply_p4(torch, F, steps=30, warmup=60):
    """The P=4 64^3 stiffness apply through a plan (the dofmap verified to
    be the structured numbering once, at setup; the autotuned scatter), and
    beside it the one-shot entry point on the same dofmap (direct atomic
    kernel).  The top-level kernel figures are the plan's operator kernel
    (accumulating applies); apply_* include the plan's clear of y."""
    pr = structured(torch, F, 64, 4)
    x = torch.from_numpy(_uniform(4, pr["ng"], -1, 1)).cuda()
    y = torch.empty(pr["ng"], device="cuda", dtype=torch.float64)
    ms, kern = time_apply(torch, F, pr, steps, warmup, x, y)
    ab = alg_bytes_stiffness(pr)
    plan = plan_scatters(torch, F, pr, x, y, steps, warmup)
    return {"workload": "stiffness apply y=Kx, 64^3 mesh, P=4 GLL, Q=5, "
                        "int32 dofmap, FP64",
            "path": f"plan, {plan['scatter']} scatter",
            "algorithmic_bytes": ab,
            "kernel_ms": plan["kernel_ms"],
            "roofline_frac": plan["kernel_roofline_frac"],
            "apply_ms": plan["ms_per_apply"],
            "apply_roofline_frac": plan["roofline_frac"],
            "gdofs": plan["gdofs"],
            "plan": plan,
            "one_shot_direct": {
                "gdofs": pr["ng"] / (ms * 1e-3) / 1e9, "kernel_ms": kern,
                "roofline_frac": ab / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS}}

def _sumfact_madds(op, N, Q):
    """Multiply-adds per cell of the sum-factorized chain both high-degree
    kernels run: mass N^3Q + N^2Q^2 + NQ^3 each way plus Q^3; stiffness
    2N^3Q + 3N^2Q^2 + 3NQ^3 each way plus 9Q^3 for S(q)."""
    if op == "mass":
        return 2 * (N ** 3 * Q + N ** 2 * Q ** 2 + N * Q ** 3) + Q ** 3
    return 2 * (2 * N ** 3 * Q + 3 * N ** 2 * Q ** 2 + 3 * N * Q ** 3) + \
        9 * Q ** 3


def bench_high_degree(torch, F, reps=10):
    """Cell-local mass and stiffness applies at degrees 9..25, Q = P+1, about
    2e7 cell-local dofs (at least 2048 cells), U, G and detJw ~ U[0, 1) from
    torch's seeded generator; FP64 and FP32.  Each case runs with the
    default kernel choice (the line kernels where one exists for the shape,
    femtc_apply_kernel) and with the general kernel forced.  Reports the time per
    apply, cell-local GDoF/s, algorithmic GB/s (U + W + geometry) and the
    multiply-add rate in TFLOP/s, for both.  Run with --only high_degree."""
    out = {}
    for P in (9, 12, 16, 19, 20, 21, 25):
        N = Q = P + 1
        C = max(2048, int(2.0e7 / N ** 3))
        gen = torch.Generator(device="cuda").manual_seed(P)
        tab = torch.from_numpy(F.basis_tables(P, Q)).cuda()
        U = torch.rand(C * N ** 3, device="cuda", dtype=torch.float64,
                       generator=gen)
        G = torch.rand(C * Q ** 3 * 6, device="cuda", dtype=torch.float64,
                       generator=gen)
        D = torch.rand(C * Q ** 3, device="cuda", dtype=torch.float64,
                       generator=gen)
        for dt, name in ((torch.float64, "f64"), (torch.float32, "f32")):
            t, u, g, d = tab.to(dt), U.to(dt), G.to(dt), D.to(dt)
            w = torch.empty_like(u)
            for op in ("mass", "stiffness"):
                def run():
                    if op == "mass":
                        F.mass_apply(t, u, w, d, C, P, Q)
                    else:
                        F.stiffness_apply(t, u, w, g, C, P, Q)

                def time_ms():
                    for _ in range(3):
                        run()
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    for _ in range(reps):
                        run()
                    e1.record()
                    torch.cuda.synchronize()
                    return e0.elapsed_time(e1) / reps
                ms = time_ms()
                with F.high_degree_kernel(F.HIGH_DEGREE_GENERAL):
                    ms_gen = time_ms()
                es = 8 if dt == torch.float64 else 4
                geo = Q ** 3 * (1 if op == "mass" else 6)
                flops = 2 * _sumfact_madds(op, N, Q) * C
                kop = F.OP_MASS if op == "mass" else F.OP_STIFFNESS
                kdt = F.F64 if dt == torch.float64 else F.F32
                out[f"P{P}_{op}_{name}"] = {
                    "cells": C,
                    "kernel": F.apply_kernel(kop, kdt, P, Q),
                    "us": ms * 1e3,
                    "gdofs": C * N ** 3 / (ms * 1e-3) / 1e9,
                    "GBs": es * C * (2 * N ** 3 + geo) / (ms * 1e-3) / 1e9,
                    "TFLOPs": flops / (ms * 1e-3) / 1e12,
                    "general_us": ms_gen * 1e3,
                    "general_gdofs": C * N ** 3 / (ms_gen * 1e-3) / 1e9,
                    "speedup_vs_general": ms_gen / ms}
            del t, u, g, d, w
        del U, G, D
        torch.cuda.empty_cache()
    # affine cells (7 geometry values per cell, computed into G(q) in the
    # kernel): the continuous stiffness apply on an n^3 mesh of about 2e7
    # global dofs, FP64
    for P in (9, 16, 25):
        Q = P + 1
        n = max(2, round((2.0e7 ** (1 / 3) - 1) / P))
        C = n ** 3
        ng = (n * P + 1) ** 3
        dm = torch.empty(C * (P + 1) ** 3, device="cuda", dtype=torch.int32)
        F.dofmap_structured(n, n, n, P, dm)
        J = _uniform(11 + P, 9 * C, -1, 1) * 0.1
        J = torch.from_numpy((J.reshape(C, 3, 3) + np.eye(3)).ravel()).cuda()
        cg = torch.empty(C * 7, device="cuda", dtype=torch.float64)
        F.geometry_cell(J, cg)
        tw = torch.from_numpy(F.basis_tables_weights(P, Q)).cuda()
        x = torch.from_numpy(_uniform(12 + P, ng, -1, 1)).cuda()
        y = torch.empty_like(x)

        def run_aff():
            F.global_apply_affine(F.OP_STIFFNESS, tw, x, y, dm, C, P, Q, cg)

        def time_aff():
            for _ in range(3):
                run_aff()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                run_aff()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps
        ms = time_aff()
        with F.high_degree_kernel(F.HIGH_DEGREE_GENERAL):
            ms_gen = time_aff()
        out[f"P{P}_stiffness_affine_global_f64"] = {
            "mesh": f"{n}^3", "cells": C, "global_dofs": ng,
            "us": ms * 1e3, "gdofs": ng / (ms * 1e-3) / 1e9,
            "general_us": ms_gen * 1e3,
            "general_gdofs": ng / (ms_gen * 1e-3) / 1e9,
            "speedup_vs_general": ms_gen / ms}
        del dm, J, cg, tw, x, y
        torch.cuda.empty_cache()
    return out


def bench_mass_c1(torch, F, sets=16, reps=8):
    """Cell-local mass apply W = M U: 32768 cells, P=3 GLL, Q=5 Gauss, U and
    W [C][64], detJw [C][125] ~ U[0.5, 1.5].  One set of buffers (66 MB)
    stays in the 256 MB Infinity Cache, so the kernel is timed both on one
    set (cache-resident) and rotating over `sets` sets (1.06 GB, streamed
    from HBM).  The applies are replayed from a captured graph, so the time
    per apply is the kernel's, without host launch gaps."""
    dev = torch.device("cuda:0")
    P, Q, C = 3, 5, 32768
    N = P + 1
    tab = torch.from_numpy(F.basis_tables(P, Q)).to(dev)
    bufs = []
    for s in range(sets):
        U = torch.from_numpy(_uniform(100 + s, C * N ** 3, -1, 1)).to(dev)
        D = torch.from_numpy(_uniform(200 + s, C * Q ** 3, 0.5, 1.5)).to(dev)
        bufs.append((U, torch.empty_like(U), D))

    def per_apply_us(order):
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):  # warm-up outside the graph
            for k in order:
                U, W, D = bufs[k]
                F.mass_apply(tab, U, W, D, C, P, Q)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for k in order:
                U, W, D = bufs[k]
                F.mass_apply(tab, U, W, D, C, P, Q)
        g.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e3 / (reps * len(order))

    rot = per_apply_us([i % sets for i in range(4 * sets)])
    res = per_apply_us([0] * (4 * sets))
    ab = 8 * C * (2 * N ** 3 + Q ** 3)
    nd = C * N ** 3
    return {"workload": "cell-local mass apply W = M U, 32768 cells, P=3 GLL, "
                        "Q=5 Gauss, FP64",
            "algorithmic_bytes": ab,
            "us_per_apply_rotated": rot,
            "gdofs_rotated": nd / (rot * 1e-6) / 1e9,
            "roofline_frac_rotated": ab / (rot * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "us_per_apply_cache_resident": res,
            "gdofs_cache_resident": nd / (res * 1e-6) / 1e9}


def bench_batch(torch, F):
    from femtc_amd import workloads as Wk
    out = {}
    for dname, dt, es in (("f64", torch.float64, 8), ("f32", torch.float32, 4)):
        spec = Wk.uneven_batch_spec(seed=0)
        items = [Wk.structured_problem(torch, Wk.box_dims(c), P, P + 1, op, dt,
                                       1000 + i) for i, (c, P, op) in enumerate(spec)]
        nbytes = sum(Wk.algorithmic_bytes(it, es) for it in items)
        b = F.Batch(items)
        st = b.stats()
        b.autotune(5)
        for _ in range(3):
            b.graph_launch()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            b.graph_launch()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        out[dname] = {"ms": ms, "gdofs": st["dofs"] / (ms * 1e-3) / 1e9,
                      "achieved_GBs": nbytes / (ms * 1e-3) / 1e9,
                      "frac_of_hbm_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "groups": st["groups"], "cells": st["cells"]}
        b.close()
        del items
        torch.cuda.empty_cache()
    out["workload"] = ("200 meshes, cells lognormal(ln 1e4, 1) clipped to "
                       "[1e3, 1e5], P uniform in 2..8, Q=P+1, half mass / half "
                       "stiffness, continuous, hipGraph replay, autotuned")
    return out


def bench_batch_groups(torch, F):
    """Each (op, degree) group of the uneven FP64 batch run as a batch of its
    own (autotuned, graph replays): where the batch's time goes without the
    overlap of the side streams."""
    import collections
    from femtc_amd import workloads as Wk
    dt, es = torch.float64, 8
    spec = Wk.uneven_batch_spec(seed=0)
    items = [Wk.structured_problem(torch, Wk.box_dims(c), P, P + 1, op, dt,
                                   1000 + i) for i, (c, P, op) in enumerate(spec)]
    groups = collections.defaultdict(list)
    for it in items:
        groups[(it["op"], it["degree"])].append(it)

    def replay_ms(b):
        b.autotune(5)
        for _ in range(3):
            b.graph_launch()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            b.graph_launch()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 10

    rows, tot_ms, tot_b = [], 0.0, 0
    for key in sorted(groups, key=lambda k: -sum(
            Wk.algorithmic_bytes(i, es) for i in groups[k])):
        sub = groups[key]
        nb = sum(Wk.algorithmic_bytes(i, es) for i in sub)
        b = F.Batch(sub)
        ms = replay_ms(b)
        rows.append({"op": "mass" if key[0] == F.OP_MASS else "stiffness",
                     "degree": key[1], "items": len(sub), "GB": nb / 1e9,
                     "ms": ms, "frac": nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "per_wg": b.groups()[0]["per_wg"]})
        b.close()
        tot_ms += ms
        tot_b += nb
    b = F.Batch(items)
    whole = replay_ms(b)
    b.close()
    return {"groups": rows, "sum_alone_ms": tot_ms,
            "sum_alone_frac": tot_b / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "whole_ms": whole,
            "whole_frac": tot_b / (whole * 1e-3) / 1e9 / HBM_PEAK_GBS}


def bench_serving(torch, F):
    """The request server: the small-mesh stream at 500 req/s, and the
    stream at full solve sizes (48^3, P=6) over a sweep of offered rates,
    with the saturation throughput of a burst."""
    from femtc_amd import serving as Sv
    items = Sv.make_pool(torch)
    st, _ = Sv.run_stream(torch, items, rate=500.0, duration=2.0, slots=4)
    st["workload"] = ("Poisson 500 req/s for 2 s, half applies / half (K+M)u=f "
                      "solves (FP32 inner + FP64 refinement to 1e-10), 16 "
                      "resident meshes n^3, n in 4..12, P in 2..6, 4 slots")
    del items
    torch.cuda.empty_cache()
    st["full_size"] = Sv.run_sweep(torch)
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=48, help="cells per direction")
    ap.add_argument("--degree", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--apply-only", action="store_true",
                    help="skip the solve, P=4 apply, batch and serving parts")
    ap.add_argument("--only", choices=("solve", "apply_p4_64cube",
                                       "mass_c1", "batch", "batch_groups",
                                       "serving", "high_degree", "mfma_apply"),
                    help="run the metric apply and this part only "
                         "(high_degree and batch_groups run only when "
                         "named)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only copy MiB three times (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported "
              "(one process per GPU would own a contiguous range of cells)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import femtc_amd as F

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.copy_calibration:
        # the hand-written copy (femtc_stream_probe, one read, one write) of
        # a.copy_calibration MiB, three times: the counter calibration run
        from femtc_amd import _lib
        m = a.copy_calibration * (1 << 20) // 8
        src = torch.ones(m, device=dev, dtype=torch.float64)
        dst = torch.empty_like(src)
        for _ in range(3):
            _lib.check(_lib.lib().femtc_stream_probe(
                1, 1, src.data_ptr(), dst.data_ptr(), m, None, 0, 4,
                torch.cuda.current_stream().cuda_stream), "stream_probe")
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 8 * m}), flush=True)
        return

    n, P = a.n, a.degree
    pr = structured(torch, F, n, P)
    C, ng, N, Q = pr["C"], pr["ng"], pr["N"], pr["Q"]
    print(f"[bench] mesh {n}^3 P={P} Q={Q} cells={C} dofs={ng}", file=sys.stderr,
          flush=True)
    x = torch.from_numpy(_uniform(4, ng, -1, 1)).to(dev)
    y = torch.empty(ng, device=dev, dtype=torch.float64)
    torch.cuda.synchronize()
    ms_per_step, kern_ms = time_apply(torch, F, pr, a.steps, a.warmup, x, y)
    value = ng / (ms_per_step * 1e-3) / 1e9
    alg_bytes = alg_bytes_stiffness(pr)
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    # a sanity check of the result: y^T 1 = x^T K 1 = 0 up to rounding
    s = float(y.sum().item())
    traffic = measured_traffic() if (n, P) == (48, 6) else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "settle_applies": SETTLE_APPLIES,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"stiffness apply y=Kx, {n}^3 hex mesh, P={P} GLL, "
                         f"Q={Q} Gauss, int32 dofmap gather/scatter, FP64"),
            "cells": C, "global_dofs": ng, "degree": P, "nq": Q,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": f"apply_cells<double,{N},{Q},OP_STIFF,gather>",
            "kernel_ms": kern_ms,
            "algorithmic_bytes": alg_bytes,
        },
        "cell_local_gdofs": C * N ** 3 / (ms_per_step * 1e-3) / 1e9,
        "check_sum_y": s,
    }
    if not a.apply_only:
        # the same apply through a plan (dofmap verified structured once at
        # setup), every scatter and the autotuned pick
        out["plan"] = plan_scatters(torch, F, pr, x, y)
    del x, y, pr
    torch.cuda.empty_cache()
    if not a.apply_only:
        for key, fn in (("solve", bench_solve), ("apply_p4_64cube", bench_apply_p4),
                        ("mass_c1", bench_mass_c1), ("batch", bench_batch),
                        ("batch_groups", bench_batch_groups),
                        ("serving", bench_serving),
                        ("mfma_apply", bench_mfma_apply),
                        ("high_degree", bench_high_degree)):
            if (a.only and key != a.only) or (not a.only and key in (
                    "high_degree", "batch_groups")):
                continue
            print(f"[bench] {key}", file=sys.stderr, flush=True)
            out[key] = fn(torch, F)
            torch.cuda.empty_cache()
        if "serving" in out:
            # p99 of the solves at full size (48^3, P=6) at 5 requests/s
            # offered (2.5 solves/s, under saturation); the small-mesh
            # stream's p99 beside it
            full = out["serving"].get("full_size", {}).get("rates", {})
            out["p99_solve_ms"] = full.get("5.0", {}).get("solve_p99_ms")
            out["p99_solve_ms_small_meshes"] = out["serving"].get(
                "solve_p99_ms")
        if "solve" in out:
            out["solve_speedup_refined_vs_fp64"] = out["solve"][
                "speedup_refined_vs_fp64"]
        if not a.only or a.only == "apply_p4_64cube":
            st = stream_bandwidth(torch, F)
            out["hbm_stream"] = st
            out["hbm_copy_GBs"] = st["copy_GBs"]
    if not a.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(P, C, ng)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
