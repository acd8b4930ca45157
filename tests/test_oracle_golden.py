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

CASES = ["tiny", "tiny_cond", "tiny_evap", "tiny_vap", "tiny_vol"]
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
    for step, row in enumerate(rows, start=1):
        orc.step()
        s = orc.scalars()
        for key in INT_KEYS:
            if key in ("istat", "iend", "jstat", "jend", "kstat") and s["tpeak"] <= deck.params.tsolid:
                continue
            assert int(row[key]) == s[key], (case, step, key)
        for key in REAL_KEYS:
            assert _same(float(row[key]), s[key]), (case, step, key, row[key], s[key])
        for fld, rec in dig["steps"].get(str(step), {}).items():
            a = orc.field(fld)
            h = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
            assert h == rec["sha256"], (case, step, fld, float(np.linalg.norm(a)), rec["l2"])
    assert orc.finished()


def test_kept_fields_match_digests():
    dig = json.load(open(golden_path("tiny_digests.json")))
    data = np.load(golden_path("tiny_fields.npz"))
    for key in data.files:
        step, fld = key[1:].split("_", 1)
        a = data[key]
        h = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
        assert h == dig["steps"][step][fld]["sha256"]
