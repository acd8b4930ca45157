"""Output files in the reference's formats (VTS/PVD, residual log, final
geometry), against the files the reference's own output layer wrote for the
same deck (tests/golden/tiny_out_outputs.json, made by
tests/golden/make_output_fixtures.py with oracle/_ref/ref_outputs)."""
import hashlib
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "am-cfd_amd", "amcfd")
FIXTURE = os.path.join(ROOT, "tests", "golden", "tiny_out_outputs.json")


def test_deck_output_strings(case_dir, amcfd):
    deck = amcfd.Deck(case_dir("tiny_out"))
    lib = amcfd.load_library()
    get = lambda key, i=0: lib.amcfd_deck_string(deck._h, key.encode(), i)
    assert get("keyword_id") == b"tiny_out"
    assert get("directory") == b"tiny_out"
    assert get("final_geometry") == b"tiny_out_geometry.txt"
    names = []
    while get("scalar_out", len(names)) is not None:
        names.append(get("scalar_out", len(names)).decode())
    assert names == ["temp", "uvel", "vvel", "wvel", "fracl", "enthalpy"]
    assert get("vector_out") is None and get("no_such_key") is None


@pytest.mark.gpu
def test_cli_writes_the_reference_files(case_dir, tmp_path):
    """Every file of a 20-step run (6 VTS files, the PVD collection, the
    residual log and the final geometry) is byte-identical to the
    reference's."""
    assert os.path.exists(CLI), "build am-cfd_amd first"
    src = case_dir("tiny_out")
    run = tmp_path / "run"
    run.mkdir()
    for f in os.listdir(os.path.dirname(src)):
        if f.startswith("tiny_out"):
            os.link(os.path.join(os.path.dirname(src), f), run / f)
    print("running the CLI on tiny_out", flush=True)
    out = subprocess.run([CLI, "tiny_out.k"], cwd=run, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    want = json.load(open(FIXTURE))["files"]
    got = {}
    for base, _, files in os.walk(run):
        for f in files:
            rel = os.path.relpath(os.path.join(base, f), run)
            if rel.startswith("tiny_out") and rel.endswith((".k",)):
                continue
            got[rel] = open(os.path.join(base, f), "rb").read()
    assert sorted(got) == sorted(want), (sorted(got), sorted(want))
    for rel, rec in want.items():
        data = got[rel]
        if "text" in rec:
            assert data.decode("latin-1") == rec["text"], rel
        assert len(data) == rec["size"], rel
        assert hashlib.sha256(data).hexdigest() == rec["sha256"], rel
