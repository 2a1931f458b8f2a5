"""Source-level drop-in: the reference's own, unmodified bench drivers
(bench_mass.cpp, bench_stiffness.cpp) compiled against the operators::
shims of include/femtc and the image's MPICH, linked to libfemtc.so
(oracle/Makefile target `dropin`, binaries in oracle/_ref/).  Needs the
reference sources and MPICH, so it runs where they are mounted; the GPU run
of the mass driver is tests/test_gpu_dropin.py.

What the unmodified stiffness driver does: it allocates its table with
(P+1)^2 values (bench_stiffness.cpp:74) while the operator reads
[phi | dphi], 2 (P+1)^2 (src/stiffness.hpp:45).  The reference reads past
the end; the shim refuses the short table before any device work, and the
driver, which catches nothing, terminates with that message."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
MPI_HOME = "/opt/conda"
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.skipif(
    not (os.path.isfile(os.path.join(REF, "bench_mass.cpp")) and
         os.path.isfile(os.path.join(MPI_HOME, "include", "mpi.h"))),
    reason="needs the reference sources and MPICH")


@pytest.fixture(scope="module")
def drivers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"),
                        "dropin", f"REF={REF}", f"MPI_HOME={MPI_HOME}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = {k: os.path.join(REF_DIR, f"bench_{k}_dropin")
           for k in ("mass", "stiffness")}
    assert all(os.access(p, os.X_OK) for p in out.values())
    return out


def test_reference_drivers_build_unmodified_against_the_shims(drivers):
    # the shims, not the reference's headers, supplied the operators: the
    # binaries depend on libfemtc.so
    for exe in drivers.values():
        r = subprocess.run(["ldd", exe], capture_output=True, text=True)
        assert "libfemtc.so" in r.stdout and "libmpi" in r.stdout


def test_short_stiffness_table_is_refused_before_device_work(drivers):
    r = subprocess.run([drivers["stiffness"], "20000", "3", "ijk", "0", "0",
                        "double"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "phi must hold [phi | dphi] = 2 (P+1)^2 values" in r.stderr


def test_reference_cli_errors_are_kept(drivers):
    # an invalid order string ends in the reference's own message
    # (linalg::string2order, src/linalg.hpp:36) from the shim's linalg
    r = subprocess.run([drivers["mass"], "20000", "3", "abc", "0", "0",
                        "double"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "Invalid loop order" in r.stderr
