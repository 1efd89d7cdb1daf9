"""The reciprocal-based division of the fused CH column solve
(custen_amd/csrc/ch_kernels.hip, divByRecip) restated in C and compared with
IEEE division on random operands over its stated range.  The GPU parity tests
check the kernel itself against the oracle, which divides with IEEE division."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_reciprocal_division_is_correctly_rounded(tmp_path):
    exe = tmp_path / "division_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-std=c99",
                    os.path.join(HERE, "cpu", "division_check.c"), "-o", str(exe), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "4000000", "12345"], capture_output=True, text=True,
                         check=True, timeout=120).stdout.split()
    assert int(out[0]) == 4000000
    assert int(out[1]) == 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_check_detects_an_uncorrected_quotient(tmp_path):
    """The check can fail: the plain product q = RN(n y), without the
    corrections, differs from the IEEE quotient in about a quarter of the
    samples."""
    src = open(os.path.join(HERE, "cpu", "division_check.c")).read()
    plain = src.replace("    q = fma(fma(-d, q, n), y, q);\n    return fma(fma(-d, q, n), y, q);",
                        "    return q;")
    assert plain != src
    path = tmp_path / "plain.c"
    path.write_text(plain)
    exe = tmp_path / "plain"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-std=c99", str(path), "-o", str(exe),
                    "-lm"], check=True)
    out = subprocess.run([str(exe), "100000", "7"], capture_output=True, text=True, check=True,
                         timeout=60).stdout.split()
    assert int(out[1]) > 1000
