"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code that
needs no device: the CPU oracle (oracle/femtc_oracle.c, every entry point
on small inputs) and the request server's admission core
(femtc_amd/csrc/admission.cpp, randomised fake-clock runs).  Built with the
host compilers only; any report fails the run (halt on error, leaks
included)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if "asan" in r.stderr.lower() or "sanitize" in r.stderr.lower():
            pytest.skip(f"sanitizer runtime unavailable: {r.stderr[:200]}")
        raise AssertionError(r.stderr)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_san")
    _build(["gcc", "-std=c11", *SAN,
            os.path.join(ROOT, "oracle", "femtc_oracle.c"),
            os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"),
            "-lm", "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "oracle driver done" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_admission_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "admission_san")
    _build(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
            "-I", os.path.join(ROOT, "femtc_amd", "csrc"),
            os.path.join(ROOT, "femtc_amd", "csrc", "admission.cpp"),
            os.path.join(ROOT, "tests", "sanitize", "admission_driver.cpp"),
            "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "admission driver done" in r.stdout
