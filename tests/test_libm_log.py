"""The device log (am-cfd_amd/csrc/libm_log.h) against the C library's log.

The reference's Fe-O surfactant model calls log(1 + K a0)
(BoundCondManager.cpp:730); the HIP path evaluates the same expression on the
device with libm_log, which must return the C library's bits exactly.
- Host: the header compiled for the host with hipcc (host-only) and compared
  over seeded arguments in the ranges the model uses, the near-1 polynomial's
  range and its edges, every binary exponent, subnormals and special values.
- Table: the committed 128-entry table is what gen_log_table.c's selection
  procedure produces (entries at both ends of the reduced range regenerated
  here; the whole table takes half a minute on 8 cores).
- Device (gpu): the kernels' exp and log on the MI355X against Python's
  math.exp / math.log, which call the C library."""
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "am-cfd_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "native", "log_check.cpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def log_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("libm") / "log_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off", "-o", exe, SRC],
                   check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("lo,hi", [(1.0, 2.0), (1.0, 1e3), (1.0, 1e8), (0.9375, 1.0647), (0.999, 1.001),
                                   (0.5, 4.0), (1e-3, 1.0), (1e-300, 1e-290), (1e300, 1.7e308)])
def test_libm_log_is_bitwise_the_c_library(log_check, lo, hi):
    out = subprocess.run([log_check, repr(lo), repr(hi), "400000"], check=True, capture_output=True,
                         text=True).stdout
    assert out.strip().splitlines()[-1] == "mismatches 0 of 400000", out


@pytest.mark.parametrize("e0,e1", [(-1074, -1023), (-1022, 1023), (-2, 2)])
def test_libm_log_every_exponent(log_check, e0, e1):
    out = subprocess.run([log_check, "bits", str(e0), str(e1), "400000"], check=True, capture_output=True,
                         text=True).stdout
    assert out.strip().splitlines()[-1] == "mismatches 0 of 400000", out


def test_log_table_is_the_selection_procedure(tmp_path):
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "gen_log_table")
    r = subprocess.run([gcc, "-O2", "-fopenmp", os.path.join(CSRC, "gen_log_table.c"), "-lquadmath", "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("generator does not build here: " + r.stderr[-200:])
    rows = re.findall(r"^    0x[0-9a-f]{16}ull, 0x[0-9a-f]{16}ull,$",
                      open(os.path.join(CSRC, "log_table.h")).read(), flags=re.M)
    assert len(rows) == 128
    for first, end in ((0, 3), (78, 82), (125, 128)):
        out = subprocess.run([exe, str(first), str(end)], check=True, capture_output=True, text=True).stdout
        assert out.splitlines() == rows[first:end], (first, end)


def _c_library(fn, v):
    """The C library's exp/log of one value (math.exp and math.log call it),
    with IEEE results where Python raises instead."""
    if fn == "log":
        if v > 0:
            return math.log(v)
        return -math.inf if v == 0 else math.nan
    try:
        return math.exp(v)
    except OverflowError:
        return math.inf


@pytest.mark.gpu
def test_device_exp_and_log_are_the_c_library(amcfd):
    rng = np.random.default_rng(20261204)
    n = 1 << 20
    edges = [float.fromhex("0x1.6p-1"), float.fromhex("0x1.6p0"), 1.0 - 2.0 ** -4, 1.0 + float.fromhex("0x1.09p-4")]
    cases = {
        "log": np.concatenate([1.0 + rng.random(n) * 1e3, 0.9375 + rng.random(n) * 0.1272,
                               np.ldexp(1.0 + rng.random(n), rng.integers(-1074, 1024, n)),
                               [1.0, 0.0, -1.0, np.inf, 5e-324] + edges]),
        "exp": np.concatenate([rng.uniform(-745.2, 709.8, n), rng.uniform(-1.0, 1.0, n),
                               rng.uniform(-30.0, 0.0, n), [0.0, -0.0, 709.78, -745.1, 1024.0, -np.inf]]),
    }
    for fn, x in cases.items():
        y = amcfd.device_libm(fn, x)
        want = np.array([_c_library(fn, float(v)) for v in x])
        same = (y.view(np.uint64) == want.view(np.uint64)) | (np.isnan(y) & np.isnan(want))
        bad = np.flatnonzero(~same)
        print(fn, x.size, "values, mismatches", bad.size)
        assert bad.size == 0, (fn, bad.size, [(x[i].hex(), y[i].hex(), want[i].hex()) for i in bad[:5]])
