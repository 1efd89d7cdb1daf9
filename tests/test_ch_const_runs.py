"""The constant runs of the factored CH matrix (chConstantRuns,
custen_amd/csrc/ch_setup.h) and the fused column solve's split of each sweep
into steps that read the factors and steps that take the run's constants
(custen_amd/csrc/ch_kernels.hip, colSolveKernel), restated in C++ over the
same header and compared bit for bit with the plain pentadiagonal solve
(serialCahnADI.c:161-241).  The GPU parity tests check the kernel itself
against the oracle."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _run(tmp_path, *args):
    exe = tmp_path / "const_runs_check"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17",
                    os.path.join(HERE, "cpu", "const_runs_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe), *args], capture_output=True, text=True, check=True,
                         timeout=120).stdout.splitlines()
    last = out[-1].split()
    return out[:-1], int(last[1]), int(last[3])


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_phased_sweep_matches_plain_solve(tmp_path):
    lines, cases, mismatches = _run(tmp_path)
    assert cases == 98
    assert mismatches == 0, "\n".join(l for l in lines if not l.endswith(" 0"))
    # the factors do settle into long runs for most coefficients (n = 2048,
    # s = 0.37: rows 21 on), so the constant phase is exercised
    runs = {tuple(l.split()[1:4:2]): int(l.split()[5]) for l in lines}
    assert runs[("2048", "0.37")] == 21


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_check_detects_a_constant_phase_one_step_early(tmp_path):
    """The check can fail: starting the constant phase one step before the
    run begins changes the solution's bits."""
    _, cases, mismatches = _run(tmp_path, "early")
    assert cases == 98
    assert mismatches > 1000
