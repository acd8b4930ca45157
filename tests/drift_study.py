"""Drift of a second solver against the FP64 mode, step by step (GPU).

    python tests/drift_study.py <case> <steps> <every> [ulp]

Runs the plain FP64 mode (bit for bit the C restatement, hence the
reference's single-thread path) and, side by side, either the scan mode or,
with "ulp", a second FP64 solver whose initial enthalpy is moved up by one
ulp.  Prints one JSON line every <every> steps: relative L2 differences of
temp/enthalpy/uvel/pp, relative differences of tpeak and the melt-pool
dimensions, and the difference of the integer melt-pool box; a final line
holds the worst values and the seconds spent in each solver.  The "ulp" run
measures the iteration's own sensitivity to a last-bit perturbation, the
floor any mode that does not keep the reference's rounding lands on.""" 
import sys, os, tempfile, json, time
sys.path.insert(0, os.getcwd()); sys.path.insert(0, os.path.join(os.getcwd(), "tests", "golden"))
import numpy as np, importlib, decks
amcfd = importlib.import_module("am-cfd_amd")
case, nsteps, every = sys.argv[1], int(sys.argv[2]), int(sys.argv[3])
def rel(a, b):
    nb = np.linalg.norm(b); return float(np.linalg.norm(a - b) / nb) if nb else float(np.linalg.norm(a))
d = tempfile.mkdtemp(); path = decks.write_case(case, d)
deck = amcfd.Deck(path)
ref = amcfd.Solver(deck, mode=amcfd.MODE_FP64)
if len(sys.argv) > 4 and sys.argv[4] == "ulp":
    fast = amcfd.Solver(deck, mode=amcfd.MODE_FP64)
    h = fast.field("enthalpy"); fast.set_field("enthalpy", np.nextafter(h, np.inf))
else:
    fast = amcfd.Solver(deck, mode=amcfd.MODE_SCAN)
tr = tf = 0.0
worst = {}
for step in range(1, nsteps + 1):
    t0 = time.perf_counter(); fin = ref.step(); t1 = time.perf_counter(); fast.step(); t2 = time.perf_counter()
    tr += t1 - t0; tf += t2 - t1
    sr, sf = ref.scalars(), fast.scalars()
    rec = {"step": step, "niter": [sr["niter"], sf["niter"]]}
    for k in ("tpeak", "alen", "depth", "width"):
        rec[k] = (sf[k] - sr[k]) / abs(sr[k]) if sr[k] else sf[k] - sr[k]
        worst[k] = max(worst.get(k, 0.0), abs(rec[k]))
    if step % every == 0 or step == nsteps or fin:
        for f in ("temp", "enthalpy", "uvel", "pp"):
            rec[f] = rel(fast.field(f), ref.field(f))
            worst[f] = max(worst.get(f, 0.0), rec[f])
        rec["box"] = [sf[k] - sr[k] for k in ("istat", "iend", "jstat", "jend", "kstat")]
        rec["dims_ref"] = [sr["alen"], sr["depth"], sr["width"]]
        print(json.dumps(rec), flush=True)
    if fin:
        break
print(json.dumps({"case": case, "steps": step, "worst": worst, "s_fp64": tr, "s_scan": tf}), flush=True)
