"""Build libfemtc.so (HIP, gfx950) in-tree.

    python -m femtc_amd.build [-j JOBS]

Compiles every source under femtc_amd/csrc with hipcc for gfx950 into
femtc_amd/_build/*.o (in parallel) and links femtc_amd/libfemtc.so, then the
C++ drivers (bench_mass, bench_stiffness) against it.  Always recompiles.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfemtc.so")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: keep hipcc from emitting packed FP32 VALU ops.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
]


# Per-file flags.  The MFMA cell pass keeps its accumulators in VGPRs: with
# AGPR accumulators every D tile costs four v_accvgpr_read before its values
# can be converted and stored.
EXTRA = {"mfma_cell.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _compile(src: str) -> tuple[str, float]:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    t0 = time.time()
    cmd = [HIPCC, *COMMON, *EXTRA.get(os.path.basename(src), []), "-c", src,
           "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj, time.time() - t0


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    t0 = time.time()
    # largest units first
    srcs.sort(key=lambda s: -os.path.getsize(s) if s.endswith("apply_n9.hip") else 0)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_drivers()
    if verbose:
        slow = sorted(results, key=lambda t: -t[1])[:3]
        print(f"[femtc build] {len(objs)} units in {time.time() - t0:.1f}s "
              f"(slowest: {', '.join(f'{os.path.basename(o)} {t:.0f}s' for o, t in slow)})",
              flush=True)
    return LIB


def _build_drivers() -> None:
    """The reference-CLI drivers, linked against libfemtc.so."""
    ddir = os.path.join(PKG, "drivers")
    if not os.path.isdir(ddir):
        return
    for src in sorted(os.listdir(ddir)):
        if not src.endswith(".cpp"):
            continue
        exe = os.path.join(PKG, "bin", src[:-4])
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra",
               "-I", os.path.join(ROOT, "include"),
               os.path.join(ddir, src), "-o", exe,
               "-L", PKG, "-lfemtc", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {src}\n{r.stderr}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.j)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
