"""The divisor check of the two-operation division used by the fused CH
column solve inside the factors' constant run (chDivCertified,
custen_amd/csrc/ch_setup.h; the kernel divides as fma(x, zh, RN(x zl)) only
for a certified pivot).  In single precision every numerator significand is
tried, so the check's answer is compared with the truth in both
directions; in double precision certified divisors (the CH pivots at eight
grid sizes and random ones) are tried on random numerators of every allowed
exponent.  The GPU parity tests check the kernel against the oracle, which
divides with IEEE division."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _run(tmp_path, *args):
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17",
                    os.path.join(HERE, "cpu", "fastdiv_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), *args], capture_output=True, text=True, check=True,
                         timeout=300).stdout.splitlines()
    res = {}
    for line in out:
        f = line.split()
        if f[0] in ("float", "double"):
            res[f[0]] = (int(f[1]), int(f[3]), int(f[5]))
        elif f[0] == "pivot":
            res[("pivot", int(f[2]))] = int(f[6])
    return res, out


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_check_agrees_with_brute_force(tmp_path):
    res, out = _run(tmp_path, "400", "2000", "20000", "11")
    nf, fcert, disagree = res["float"]
    assert nf == 400
    assert disagree == 0, "\n".join(out)
    # some divisors really fail (and are refused), most pass
    assert 380 <= fcert < 400
    nd, dcert, mism = res["double"]
    assert mism == 0
    assert dcert > 0.9 * nd
    # the bench's pivots are all certified, so the fast path is what runs
    for n in (64, 128, 130, 256, 512, 1024, 2048, 4096):
        assert res[("pivot", n)] == 1


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_check_without_candidates_certifies_failing_divisors(tmp_path):
    """The comparison can fail: a check that tries no numerators certifies
    the divisors for which some numerator is rounded wrongly."""
    res, _ = _run(tmp_path, "400", "0", "0", "11", "narrow")
    nf, fcert, disagree = res["float"]
    assert fcert == 400
    assert disagree >= 1
