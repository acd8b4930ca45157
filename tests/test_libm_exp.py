"""The device exp (am-cfd_amd/csrc/libm_exp.h) against the C library's exp.

The reference evaluates exp() on the host (laser map, evaporation, the
surfactant model, the volumetric source, the exponential scheme); the HIP
path evaluates the same expressions on the device with libm_exp, which must
return the C library's bits exactly.  Compiled here for the host with hipcc
(host-only) and compared over seeded arguments in the ranges the solver
uses, the subnormal and overflow edges, and special values."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "exp_check.cpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def exp_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("libm") / "exp_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off", "-o", exe, SRC],
                   check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("lo,hi", [(-20.0, 0.0), (0.0, 20.0), (-1.0, 1.0), (-0.01, 0.01), (-1e-15, 1e-15),
                                   (-745.2, -700.0), (-708.5, -512.0), (512.0, 709.8), (-1100.0, 1100.0)])
def test_libm_exp_is_bitwise_the_c_library(exp_check, lo, hi):
    out = subprocess.run([exp_check, repr(lo), repr(hi), "400000"], check=True, capture_output=True,
                         text=True).stdout
    assert out.strip().splitlines()[-1] == "mismatches 0 of 400000", out
