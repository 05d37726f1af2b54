"""The reciprocal-corrected division of the sweeps (armon::div_by_recip, device_math.hpp), which
replaces x / dx for cell sizes that are not powers of two, against IEEE division on the host,
bit for bit: random operands over the whole accepted range, quotients constructed next to a
rounding midpoint (the cases a correction step can get wrong), the corner of the proof near
|x| = |d| = 2^k, signed zeros, and the guard's range. tests/div/recip_div_check.cpp; the same
program ran 2.4e9 comparisons (24 M of them next to a midpoint) with no mismatch."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "div", "bin", "recip_div_check")


def test_recip_division_is_the_ieee_quotient():
    out = subprocess.run([EXE, "2000000", "50000", "7"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0," in out.stdout and "signed zeros ok" in out.stdout and "guard ok" in out.stdout
