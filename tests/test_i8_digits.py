"""CPU check of the split-INT8 digit formulations (csrc/common/i8_digits.hpp).

The device pre-pass computes each element's INT8 digits in fixed point (one
rounding of the remainder to 8 (S - 1) bits, a byte bias, carry into the
leading digit); the stepwise form rounds digit by digit with carries.  The
balanced base-256 representation is unique, so both must give the same
digits for every input: random elements, exact ties of the last digit's
rounding, half-integer leading digits and carry boundaries, S = 2..5.  The
header is compiled here with g++ (tests/i8_digits_check.cpp).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_fixed_point_digits_equal_stepwise_digits(tmp_path):
    exe = tmp_path / "i8_digits_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", str(exe),
                    os.path.join(HERE, "i8_digits_check.cpp")], check=True)
    r = subprocess.run([str(exe), "200000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("S=")]
    assert len(lines) == 4
    for l in lines:
        f = l.split()
        assert f[f.index("mismatches") + 1] == "0"
        assert int(f[f.index("ties") + 1]) > 0  # the tie cases were exercised
