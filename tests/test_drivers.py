"""The reference-CLI drivers (femtc_amd/bin/bench_mass, bench_stiffness) and
the shim example: command line, ladder messages and output line of
bench_mass.cpp:142-145,170-185.  The argument-error paths run on the CPU;
the timed runs need the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "femtc_amd", "bin")


def _run(name, *args, timeout=120):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    return subprocess.run([exe, *map(str, args)], capture_output=True,
                          text=True, timeout=timeout)


@pytest.mark.parametrize("name", ["bench_mass", "bench_stiffness"])
def test_usage_line(name):
    r = _run(name, 100, 3)
    assert r.returncode == 1
    assert "Usage:" in r.stdout and "<Ndofs> <degree> <order>" in r.stdout


@pytest.mark.parametrize("name", ["bench_mass", "bench_stiffness"])
def test_ladder_messages_exit_cleanly_before_any_device_work(name):
    # src/mass.hpp:134-136, 190-192: print and return
    r = _run(name, 1000000, 30, "ijk", 0, 0, "double")
    assert r.returncode == 0 and r.stdout.strip() == "degree not supported"
    r = _run(name, 1000000, 3, "ijk", 3, 0, "double")
    assert r.returncode == 0 and r.stdout.strip() == "Block size not supported"


@pytest.mark.parametrize("name", ["bench_mass", "bench_stiffness"])
def test_invalid_order_string_throws_like_the_reference(name):
    # linalg::string2order throws std::runtime_error (src/linalg.hpp:36),
    # uncaught in the reference drivers as here
    r = _run(name, 1000000, 3, "abc", 0, 0, "double")
    assert r.returncode != 0
    assert "Invalid loop order" in r.stderr


def _fields(line):
    f = line.split()
    assert len(f) == 11, line
    return f


@pytest.mark.gpu
@pytest.mark.parametrize("name,args", [
    ("bench_mass", (2097152, 3, "ijk", 0, 0, "double")),
    ("bench_mass", (1000000, 5, "kij", 8, 16, "float", "correct")),
    ("bench_stiffness", (1000000, 4, "kij", 8, 16, "double")),
    ("bench_stiffness", (500000, 6, "ikj", 4, 4, "double", "correct")),
    # degrees past the fused kernels, up to the top of the reference's ladder
    ("bench_mass", (1000000, 12, "ijk", 0, 0, "double")),
    ("bench_stiffness", (1000000, 25, "kji", 16, 8, "double"))])
def test_driver_output_line_on_gpu(name, args):
    r = _run(name, *args)
    assert r.returncode == 0, r.stderr
    f = _fields(r.stdout.strip().splitlines()[-1])
    # degree nranks type Ndofs elapsed order Mb Nb GFLOP/s GB/s GDoF/s
    assert f[0] == str(args[1]) and f[1] == "1" and f[2] == args[5]
    assert f[3] == str(args[0]) and f[5] == args[2]
    assert f[6] == str(args[3]) and f[7] == str(args[4])
    assert float(f[4]) > 0 and float(f[10]) > 0


@pytest.mark.gpu
def test_shim_example_runs_reference_style_code():
    r = _run("shim_example", 3)
    assert r.returncode == 0, r.stdout + r.stderr
