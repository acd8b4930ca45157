"""The C restatement against the golden fixtures written by the compiled
reference (tests/golden/make_fixtures.py): same deck -> same trajectory, to
the last bit, for every field at every step."""
import csv
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import golden_path

CASES = ["tiny", "tiny_cond", "tiny_evap", "tiny_vap", "tiny_vol", "tiny_tool", "tiny_powder",
         "tiny_move", "tiny_move_powder", "tiny_species", "tiny_move_species", "tiny_fs", "tiny_dtvar",
         "tiny_heattable", "tiny_steady", "multilayer_small"]
AUX_FILES = [("_tool.crs", "@TOOL@", "tool"), ("_dt.k", "@DT@", "dt"), ("_heat.k", "@HEAT@", "heat")]
MOVING = ["tiny_move", "tiny_move_powder", "tiny_move_species"]
INT_KEYS = ["niter", "itertot", "powerindicator", "istat", "iend", "jstat", "jend", "kstat", "jstart",
            "istart", "nk"]
REAL_KEYS = ["timet", "resorh", "resorm", "resoru", "resorv", "resorw", "tpeak", "alen", "depth",
             "width", "ratio", "heatinlaser", "heatout", "accul", "heatvol", "ahtoploss", "beamposx",
             "beamposy"]


def _same(a, b):
    if np.isnan(a) and np.isnan(b):
        return True
    return a == b


@pytest.mark.parametrize("case", CASES)
def test_deck_text_matches_fixture(case, case_dir):
    path = case_dir(case)
    mesh = os.path.join(os.path.dirname(path), case + "_mesh.k")
    with open(path) as f:
        text = f.read().replace(os.path.abspath(mesh), "@MESH@")
    for suffix, mark, kind in AUX_FILES:
        aux = os.path.join(os.path.dirname(path), case + suffix)
        text = text.replace(os.path.abspath(aux), mark)
        assert os.path.exists(aux) == os.path.exists(golden_path(case + "_" + kind + ".txt")), aux
        if os.path.exists(aux):
            with open(aux) as f, open(golden_path(case + "_" + kind + ".txt")) as g:
                assert f.read() == g.read()
    with open(golden_path(case + "_deck.txt")) as f:
        assert text == f.read()
    with open(mesh) as f, open(golden_path(case + "_mesh.txt")) as g:
        assert f.read() == g.read()


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference_bitwise(case, case_dir, amcfd, oracle_mod):
    deck = amcfd.Deck(case_dir(case))
    orc = oracle_mod.Oracle(deck.params)
    rows = list(csv.DictReader(open(golden_path(case + "_scalars.csv"))))
    dig = json.load(open(golden_path(case + "_digests.json")))
    assert list(orc.shape) == dig["shape"]
    outtrack, monitor = 0.0, []
    for step, row in enumerate(rows, start=1):
        orc.step()
        s = orc.scalars()
        outtrack += s["delt"]
        if outtrack >= s["outtime"]:   # the free surface runs at output times (main.cpp:167-176)
            if deck.params.energyfreesurface:
                monitor.append((s["timet"],) + orc.free_surface())
            outtrack = 0.0
        for key in INT_KEYS:
            if key in ("istat", "iend", "jstat", "jend", "kstat") and s["tpeak"] <= deck.params.tsolid:
                continue
            assert int(row[key]) == s[key], (case, step, key)
        for key in REAL_KEYS:
            assert _same(float(row[key]), s[key]), (case, step, key, row[key], s[key])
        if "resorc" in row:
            assert _same(float(row["resorc"]), s["resorc"]), (case, step, "resorc", row["resorc"], s["resorc"])
        for fld, rec in dig["steps"].get(str(step), {}).items():
            a = orc.field(fld)
            h = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
            assert h == rec["sha256"], (case, step, fld, float(np.linalg.norm(a)), rec["l2"])
    assert orc.finished()
    if deck.params.energyfreesurface:
        # the monitor lines the reference printed, FreeSurfaceManager.cpp:197-206
        want = [l for l in open(golden_path(case + "_FreeSurfaceMonitor.txt")).read().splitlines()
                if l and l[0] == " " and not l.lstrip().startswith("time")]
        got = ["%12.4e%14.4e%14.4e%14.4e%14.4e%14.4e%9d" % m for m in monitor]
        assert got == want


def read_grid_steps(case):
    """{step: (x, y)} from <case>_grid_steps.txt (17 significant digits, which
    round-trip a double exactly)."""
    out = {}
    with open(golden_path(case + "_grid_steps.txt")) as f:
        for line in f:
            w = line.split()
            iy = w.index("y")
            out[int(w[0])] = (np.array([float(v) for v in w[2:iy]]), np.array([float(v) for v in w[iy + 1:]]))
    return out


@pytest.mark.parametrize("case", MOVING)
def test_oracle_moving_grid_matches_reference(case, case_dir, amcfd, oracle_mod):
    """The regenerated x/y node coordinates, step by step, to the bit."""
    deck = amcfd.Deck(case_dir(case))
    orc = oracle_mod.Oracle(deck.params)
    grids = read_grid_steps(case)
    ni, nj = orc.shape[2], orc.shape[1]
    moved = 0
    for step in sorted(grids):
        orc.step()
        x = orc.field("x", shape=(ni,))
        y = orc.field("y", shape=(nj,))
        gx, gy = grids[step]
        assert np.array_equal(x, gx), (case, step, np.abs(x - gx).max())
        assert np.array_equal(y, gy), (case, step, np.abs(y - gy).max())
        moved += 1
    assert moved == len(grids) > 0


def test_moving_mesh_grid_stays_monotone():
    """The shifted zones keep every node strictly ordered, and the moving
    zone keeps its length, so the node under the beam zone moves by exactly
    the beam's x displacement (closed form of the zone update)."""
    for case in MOVING:
        grids = read_grid_steps(case)
        rows = list(csv.DictReader(open(golden_path(case + "_scalars.csv"))))
        for step, (x, y) in grids.items():
            assert np.all(np.diff(x) > 0) and np.all(np.diff(y) > 0), (case, step)
        # steps 1->2 (a +x move): the nodes of the middle zone shift by the
        # beam's displacement, those of the outer zones by less
        x1, x2 = grids[1][0], grids[2][0]
        bx1, bx2 = float(rows[0]["beamposx"]), float(rows[1]["beamposx"])
        shift = x2 - x1
        assert bx2 > bx1
        assert np.isclose(shift.max(), bx2 - bx1, rtol=1e-9, atol=0), (case, shift.max(), bx2 - bx1)
        assert shift.min() >= 0.0 and shift[0] == 0.0 and shift[-1] == 0.0


def test_kept_fields_match_digests():
    dig = json.load(open(golden_path("tiny_digests.json")))
    data = np.load(golden_path("tiny_fields.npz"))
    for key in data.files:
        step, fld = key[1:].split("_", 1)
        a = data[key]
        h = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
        assert h == dig["steps"][step][fld]["sha256"]
