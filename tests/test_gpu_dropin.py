"""The reference's own mass driver (bench_mass.cpp, unmodified) running on
the GPU through the operators:: shims: the binary built by oracle/Makefile
`dropin` (see tests/test_dropin.py) with the reference CLI
<Ndofs> <degree> <order> <Mb> <Nb> <type>; it checks W >= 0 itself and
prints its result line (bench_mass.cpp:140-146)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "bench_mass_dropin")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.access(EXE, os.X_OK),
                                 reason="drop-in driver not built here")]


@pytest.mark.parametrize("degree,order,mb,nb,typ", [(3, "ijk", 0, 0, "double"),
                                                   (4, "kij", 8, 16, "float"),
                                                   (6, "jik", 4, 4, "double")])
def test_reference_mass_driver_runs_through_the_shim(degree, order, mb, nb,
                                                     typ):
    ndofs = 2_000_000
    r = subprocess.run([EXE, str(ndofs), str(degree), order, str(mb), str(nb),
                        typ], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    line = r.stdout.strip().splitlines()[-1].split()
    # degree nranks type Ndofs elapsed order Mb Nb GFLOP/s GB/s GDoF/s
    assert len(line) == 11, line
    assert int(line[0]) == degree and int(line[1]) == 1
    assert line[2] == typ and int(line[3]) == ndofs and line[5] == order
    assert int(line[6]) == mb and int(line[7]) == nb
    elapsed, gdofs = float(line[4]), float(line[10])
    assert elapsed > 0 and gdofs > 0
    print(" ".join(line))
