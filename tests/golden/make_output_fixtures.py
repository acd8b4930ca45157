"""Write the output-format fixtures from the compiled reference.

Runs oracle/_ref/ref_outputs (the reference's solver classes and its own
output layer, PrintManager + vtsBinWriter, driven in the reference
executable's order with OMP_NUM_THREADS=1) on a deck of decks.py and stores
in tests/golden/<case>_outputs.json:
  - for every file it wrote: path relative to the run directory, size,
    SHA-256;
  - the full text of the small text files (PVD collection, residual log,
    final geometry, free-surface monitor, extracted sections);
  - the XML header of every VTS file.
Only data is stored.  Run from the repo root:
    python tests/golden/make_output_fixtures.py [case]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import decks  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "ref_outputs")
TEXT_SUFFIXES = (".pvd", ".txt")


def collect(run_dir, skip=()):
    """Every file under run_dir except the deck inputs: name -> record."""
    out = {}
    for base, _, files in os.walk(run_dir):
        for f in sorted(files):
            path = os.path.join(base, f)
            rel = os.path.relpath(path, run_dir)
            if rel in skip:
                continue
            data = open(path, "rb").read()
            rec = {"size": len(data), "sha256": hashlib.sha256(data).hexdigest()}
            if rel.endswith(TEXT_SUFFIXES) and len(data) <= 65536:
                rec["text"] = data.decode("latin-1")
            if rel.endswith(".vts"):   # the XML part, up to the raw appended data
                rec["header"] = data[:data.index(b"<AppendedData encoding=\"raw\">_") + 30].decode("latin-1")
            out[rel] = rec
    return out


def main(case="tiny_out"):
    run = tempfile.mkdtemp(prefix="amcfd_refout_")
    deck = decks.write_case(case, run)
    inputs = {os.path.basename(deck), os.path.basename(deck).replace(".k", "_mesh.k")}
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([EXE, os.path.basename(deck)], cwd=run, env=env, check=True, capture_output=True)
    files = collect(run, skip=inputs)
    dst = os.path.join(HERE, case + "_outputs.json")
    with open(dst, "w") as f:
        json.dump({"case": case, "harness": "oracle/_ref/ref_outputs (OMP_NUM_THREADS=1)", "files": files},
                  f, indent=1, sort_keys=True)
    print(dst, len(files), "files")


if __name__ == "__main__":
    main(*sys.argv[1:])
