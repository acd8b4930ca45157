"""The line sweeps' division (am-cfd_amd/csrc/div_pos.h) against IEEE division.

The recurrence qr = (d + aw qr) / den (CFDSolverManager.cpp:2442, 2514) is
the sweep's dependency chain, so the device divides with a first quotient
x*RN(1/den) and one exact-remainder correction (three dependent operations)
and confirms every row after the pass with Markstein's second correction; a
line with an unconfirmed row is solved again with IEEE division.  The test
checks on the host, with the same header, that the confirmation never
passes a quotient that differs from x / den, and that the short form is
itself correctly rounded in practice (so the fallback stays rare), including
for dividends placed next to b times a rounding midpoint and for signed
zeros."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "div_pos_check.cpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def div_pos_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("div_pos") / "div_pos_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off", "-o", exe, SRC],
                   check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("mode,n", [(0, 4000000), (1, 4000000), (2, 1000)])
def test_div_pos_is_ieee_division(div_pos_check, mode, n):
    out = subprocess.run([div_pos_check, str(mode), str(n)], check=True, capture_output=True, text=True).stdout
    m = re.search(r"wrong (\d+) flagged (\d+) missed (\d+) of (\d+)", out)
    wrong, flagged, missed = int(m.group(1)), int(m.group(2)), int(m.group(3))
    assert missed == 0, out
    assert flagged == wrong, out
    assert wrong <= n // 1000000, out



def test_confirmation_flags_a_wrong_quotient(div_pos_check):
    """A quotient one ulp off is never confirmed, so the check is not vacuous."""
    n = 1000000
    out = subprocess.run([div_pos_check, "3", str(n)], check=True, capture_output=True, text=True).stdout
    m = re.search(r"wrong (\d+) flagged (\d+) missed (\d+) of (\d+)", out)
    assert int(m.group(1)) == n and int(m.group(2)) == n and int(m.group(3)) == 0, out
