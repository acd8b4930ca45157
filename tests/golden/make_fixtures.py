"""Write the golden fixtures under tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_harness (the reference solver classes compiled from
their own sources by oracle/Makefile, driven in the reference executable's
call order, OMP_NUM_THREADS=1) on the decks of decks.py and stores, per case:

  <case>_scalars.csv     every per-step scalar the harness prints
  <case>_digests.json    per step and field: SHA-256 of the raw float64
                         bytes and the L2 norm
  <case>_fields.npz      full fields at a few steps (tiny cases only)
  <case>_deck.txt        the deck text (mesh path replaced by @MESH@)
  <case>_mesh.txt        the mesh text
  <case>_tool.txt, _dt.txt, _heat.txt
                         the toolpath, time-step and heat-source tables, when
                         the deck names them
  <case>_grid_steps.txt  per step, the regenerated x and y node coordinates
                         (moving-mesh decks)
  <case>_FreeSurfaceMonitor.txt  the free-surface monitor (free-surface decks)

Only data is stored; no reference source text is copied.  Run from the repo
root:  python tests/golden/make_fixtures.py [case ...]
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import decks  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
FIELDS = ["enthalpy", "temp", "fracl", "uvel", "vvel", "wvel", "pressure", "pp", "rho", "diff",
          "vis", "hnot", "fraclnot", "unot", "vnot", "wnot", "sourceinput", "dux", "dvy", "dwz"]
# table files a deck may name: suffix, placeholder in <case>_deck.txt, fixture kind
AUX_FILES = [("_tool.crs", "@TOOL@", "tool"), ("_dt.k", "@DT@", "dt"), ("_heat.k", "@HEAT@", "heat")]
# written by the harness only when the deck enables species / G/R / free surface / powder bed
OPTIONAL_FIELDS = ["avgconcentration", "avgconcentrationnot", "concentration", "grad", "rate", "zr", "tempr",
                   "csfrac", "solfrac", "mpPhase", "mpLiqPhase"]
# case -> (fields_every, max_steps, steps whose full fields are kept)
PLAN = {
    "tiny": (1, -1, [2, 40]),
    "tiny_cond": (1, -1, []),
    "tiny_evap": (1, -1, []),
    "tiny_vap": (1, -1, []),
    "tiny_vol": (1, -1, []),
    "tiny_tool": (1, -1, []),
    "tiny_powder": (1, -1, []),
    "tiny_move": (1, -1, []),
    "tiny_move_powder": (1, -1, []),
    "tiny_species": (1, -1, []),
    "tiny_move_species": (1, -1, []),
    "tiny_fs": (1, -1, []),
    "tiny_dtvar": (1, -1, []),
    "tiny_heattable": (1, -1, []),
    "tiny_steady": (1, -1, []),
    "plate120": (50, -1, []),
    "multilayer_small": (5, -1, []),
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def run_case(case, workdir=None, reuse=False):
    """Run the harness for `case` (or, with reuse=True, ingest an earlier
    harness run of the same deck found in `workdir`)."""
    every, max_steps, keep = PLAN[case]
    tmp = workdir or tempfile.mkdtemp(prefix="amcfd_ref_")
    if reuse:
        deck = os.path.join(tmp, case + ".k")
        with open(os.path.join(tmp, "run1.log")) as f:
            info = json.loads(f.read().strip().splitlines()[-1])
    else:
        deck = decks.write_case(case, tmp)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        out = subprocess.run([HARNESS, os.path.basename(deck), ".", str(every), str(max_steps)],
                             cwd=tmp, env=env, check=True, capture_output=True, text=True)
        info = json.loads(out.stdout.strip().splitlines()[-1])
    with open(os.path.join(tmp, "grid.txt")) as f:
        ni, nj, nk = (int(v) for v in f.readline().split())
    shape = (nk, nj, ni)
    steps = sorted({int(n[1:6]) for n in os.listdir(tmp) if n.startswith("f") and n.endswith(".bin")})
    dig = {"shape": list(shape), "steps": {}}
    kept = {}
    for s in steps:
        dig["steps"][str(s)] = {}
        for fld in FIELDS + [f for f in OPTIONAL_FIELDS
                             if os.path.exists(os.path.join(tmp, "f%05d_%s.bin" % (s, f)))]:
            a = np.fromfile(os.path.join(tmp, "f%05d_%s.bin" % (s, fld)), dtype="<f8").reshape(shape)
            dig["steps"][str(s)][fld] = {"sha256": digest(a), "l2": float(np.linalg.norm(a))}
            if s in keep:
                kept["s%d_%s" % (s, fld)] = a
    shutil.copy(os.path.join(tmp, "scalars.csv"), os.path.join(HERE, case + "_scalars.csv"))
    with open(os.path.join(HERE, case + "_digests.json"), "w") as f:
        json.dump(dig, f, indent=0, sort_keys=True)
    if kept:
        np.savez_compressed(os.path.join(HERE, case + "_fields.npz"), **kept)
    mesh = os.path.abspath(os.path.join(tmp, case + "_mesh.k"))
    with open(deck) as f:
        text = f.read().replace(mesh, "@MESH@")
    for suffix, mark, kind in AUX_FILES:
        aux = os.path.abspath(os.path.join(tmp, case + suffix))
        text = text.replace(aux, mark)
        if os.path.exists(aux):
            shutil.copy(aux, os.path.join(HERE, case + "_" + kind + ".txt"))
    monitor = os.path.join(tmp, case + "_FreeSurfaceMonitor.txt")
    if os.path.exists(monitor):
        shutil.copy(monitor, os.path.join(HERE, case + "_FreeSurfaceMonitor.txt"))
    if os.path.exists(os.path.join(tmp, "grid_steps.txt")):
        shutil.copy(os.path.join(tmp, "grid_steps.txt"), os.path.join(HERE, case + "_grid_steps.txt"))
    with open(os.path.join(HERE, case + "_deck.txt"), "w") as f:
        f.write(text)
    shutil.copy(mesh, os.path.join(HERE, case + "_mesh.txt"))
    info["case"] = case
    info["harness"] = "oracle/_ref/ref_harness (OMP_NUM_THREADS=1)"
    with open(os.path.join(HERE, case + "_run.json"), "w") as f:
        json.dump(info, f, indent=1, sort_keys=True)
    if workdir is None:
        shutil.rmtree(tmp)
    return info


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    args = sys.argv[1:]
    if args[:1] == ["--reuse"]:
        # --reuse <case> <dir>: ingest a finished harness run (run1.log, scalars.csv, f*.bin)
        print(run_case(args[1], os.path.abspath(args[2]), reuse=True))
    else:
        for c in args or ["tiny", "tiny_cond", "tiny_evap"]:
            print(run_case(c))
