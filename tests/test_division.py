"""The reciprocal-based division of the step kernels (dvr: Markstein's
correction, used for constant and per-row divisors) equals IEEE a / b bit for
bit: a C restatement with the C library's correctly rounded fma, checked on
random operands, quotients next to rounding midpoints, divisors with nearly
all-ones significands, negative divisors, signed zeros and infinities.  The
GPU parity tests then cover the kernels that use it."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reciprocal_division_is_ieee_division(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "dvr_check")
    # -O1 without contraction or target-specific FMA folding: fma() is the libm call
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "csrc", "dvr_check.c"), "-lm"],
                   check=True)
    r = subprocess.run([exe, "20000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "0 differ" in r.stdout
