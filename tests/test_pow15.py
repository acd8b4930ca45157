"""pow(v, 1.5) of the free-surface energy balance (am-cfd_amd/csrc/pow15.h)
against the C library's pow.

The device evaluates v*sqrt(v) with the rounding errors of the square root
and the product carried exactly (FMA residuals), i.e. the correctly rounded
v^1.5.  The C library's pow is accurate to about 0.52 ulp and so differs by
one ulp about once in 1300 calls; the test bounds that rate and checks, with
60-digit decimal arithmetic, that in the sampled disagreements the device
value is the correctly rounded one."""
import os
import re
import shutil
import subprocess
from decimal import Decimal, getcontext

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "pow15_check.cpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def pow15_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("pow15") / "pow15_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off", "-o", exe, SRC],
                   check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("lo,hi", [(1.0, 1.0000001), (1.0, 1.01), (1.0, 2.0), (1.0, 100.0)])
def test_pow15_against_the_c_library(pow15_check, lo, hi):
    n = 400000
    out = subprocess.run([pow15_check, repr(lo), repr(hi), str(n)], check=True, capture_output=True,
                         text=True).stdout
    m = re.search(r"mismatches (\d+) of (\d+), max ulp (\d+)", out)
    bad, worst = int(m.group(1)), int(m.group(3))
    assert worst <= 1 and bad <= n // 500, out
    getcontext().prec = 60
    for v, mine, libm in re.findall(r"v=(\S+) mine=(\S+) libm=(\S+)", out):
        v, mine, libm = (float.fromhex(t) for t in (v, mine, libm))
        exact = Decimal(v) * Decimal(v).sqrt()
        assert abs(Decimal(mine) - exact) < abs(Decimal(libm) - exact), (v, mine, libm)
