"""Closed-form and brute-force checks of pieces of the C restatement, chosen
so that a dropped term, a wrong sign or a transposed index fails one."""
import ctypes
import csv
import os

import numpy as np
import pytest

from conftest import golden_path


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_power_law_closed_form(oracle_mod):
    lib = oracle_mod.load()
    for pe in [0.0, 0.3, 1.0, 2.5, 5.0, 9.99, 10.0, 12.0, 50.0]:
        want = max(0.0, (1.0 - 0.1 * pe) ** 5)
        assert abs(lib.orc_disc_power(pe) - want) <= 1e-15 * max(1.0, want)
    assert lib.orc_disc_power(10.0) == 0.0
    assert lib.orc_disc_power(11.0) == 0.0       # negative base clamps to zero
    # exponential scheme is P/(e^P - 1): 1 - P/2 + P^2/12 near zero
    for pe in [1e-3, 1e-2]:
        assert abs(lib.orc_disc_exponential(pe) - (1 - pe / 2 + pe * pe / 12)) < 1e-9


def test_enthalpy_temperature_inverse(amcfd, oracle_mod, case_dir):
    p = amcfd.Deck(case_dir("tiny")).params
    lib = oracle_mod.load()
    fl = ctypes.c_double()
    for T in [300.0, 800.0, 1500.0, p.tsolid - 1e-3]:
        h = 0.5 * p.acpa * T * T + p.acpb * T
        t = lib.orc_h_to_t(ctypes.byref(p), h, ctypes.byref(fl))
        assert abs(t - T) <= 1e-9 * T and fl.value == 0.0
    # mushy zone: linear in h between hsmelt and hlcal
    for f in [0.1, 0.5, 0.9]:
        h = p.hsmelt + f * (p.hlcal - p.hsmelt)
        t = lib.orc_h_to_t(ctypes.byref(p), h, ctypes.byref(fl))
        assert abs(fl.value - f) < 1e-12
        assert abs(t - (p.tsolid + f * (p.tliquid - p.tsolid))) < 1e-9
    # liquid: T = Tl + (h - hlcal)/acpl
    h = p.hlcal + 7.0e4
    t = lib.orc_h_to_t(ctypes.byref(p), h, ctypes.byref(fl))
    assert fl.value == 1.0 and abs(t - (p.tliquid + 7.0e4 / p.acpl)) < 1e-9


@pytest.mark.parametrize("n", [1, 2, 5, 17])
def test_thomas_against_dense_solve(oracle_mod, n):
    rng = np.random.default_rng(n)
    aw = rng.uniform(0, 1, n + 2)
    ae = rng.uniform(0, 1, n + 2)
    ap = aw + ae + rng.uniform(0.1, 1, n + 2)
    d = rng.normal(size=n + 2)
    x = rng.normal(size=n + 2)       # x[0], x[n+1] are the fixed ends
    x0, xn1 = x[0], x[n + 1]
    lib = oracle_mod.load()
    lib.orc_thomas(n, _dp(aw), _dp(ap), _dp(ae), _dp(d), _dp(x))
    A = np.zeros((n, n))
    rhs = d[1:n + 1].copy()
    for r in range(n):
        i = r + 1
        A[r, r] = ap[i]
        if r > 0:
            A[r, r - 1] = -aw[i]
        else:
            rhs[r] += aw[i] * x0
        if r < n - 1:
            A[r, r + 1] = -ae[i]
        else:
            rhs[r] += ae[i] * xn1
    want = np.linalg.solve(A, rhs)
    assert np.allclose(x[1:n + 1], want, rtol=1e-12, atol=1e-12)
    assert x[0] == x0 and x[n + 1] == xn1


def test_grid_closed_form(amcfd, oracle_mod, case_dir):
    p = amcfd.Deck(case_dir("tiny")).params
    o = oracle_mod.Oracle(p)
    ni = p.ni
    xu = o.field("xu", shape=(ni,))
    x = o.field("x", shape=(ni,))
    # zone 1: 4 cells over 0.1 mm, exponent -1.2 (coarse -> fine)
    s = np.arange(1, 5) / 4.0
    assert np.allclose(xu[2:6], 0.1e-3 * (1 - (1 - s) ** 1.2), rtol=0, atol=1e-18)
    # zone 2: 12 uniform cells over 0.3 mm
    assert np.allclose(np.diff(xu[5:18]), 0.3e-3 / 12, rtol=1e-12)
    # zone 3: exponent 1.2 (fine -> coarse) and the domain ends at 0.5 mm
    s = np.arange(1, 5) / 4.0
    assert np.allclose(xu[18:22], 0.4e-3 + 0.1e-3 * s ** 1.2, rtol=1e-14)
    assert xu[0] == 0.0 and xu[1] == 0.0 and abs(xu[ni - 1] - 0.5e-3) < 1e-18
    assert np.allclose(x[:-1], 0.5 * (xu[1:] + xu[:-1]), rtol=0, atol=0)
    assert x[-1] == xu[-1]


def test_plate120_first_steps_bitwise(amcfd, oracle_mod, case_dir):
    """The 120x48x32 plate: the first 6 steps reproduce the reference's
    per-step scalars exactly (iteration counts, residuals, pool extents)."""
    deck = amcfd.Deck(case_dir("plate120"))
    o = oracle_mod.Oracle(deck.params)
    rows = list(csv.DictReader(open(golden_path("plate120_scalars.csv"))))
    for step in range(1, 7):
        o.step()
        s = o.scalars()
        r = rows[step - 1]
        assert int(r["niter"]) == s["niter"], step
        for key in ("resorh", "resorm", "resoru", "tpeak", "alen", "depth", "width", "ratio"):
            assert float(r[key]) == s[key], (step, key, r[key], s[key])
        if s["tpeak"] > deck.params.tsolid:
            for key in ("istat", "iend", "jstat", "jend", "kstat"):
                assert int(r[key]) == s[key], (step, key)


@pytest.mark.slow
def test_plate120_full_run_bitwise(amcfd, oracle_mod, case_dir):
    """All 400 steps (about 15 min on one core): scalars every step, field
    digests every 50 steps."""
    import hashlib
    import json
    deck = amcfd.Deck(case_dir("plate120"))
    o = oracle_mod.Oracle(deck.params)
    rows = list(csv.DictReader(open(golden_path("plate120_scalars.csv"))))
    dig = json.load(open(golden_path("plate120_digests.json")))
    for step in range(1, len(rows) + 1):
        o.step()
        s = o.scalars()
        assert int(rows[step - 1]["niter"]) == s["niter"], step
        assert float(rows[step - 1]["tpeak"]) == s["tpeak"], step
        for fld, rec in dig["steps"].get(str(step), {}).items():
            a = o.field(fld)
            assert hashlib.sha256(a.astype("<f8").tobytes()).hexdigest() == rec["sha256"], (step, fld)
