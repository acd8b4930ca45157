"""Throughput of the per-time-step SIMPLE solve on one MI355X.

Workload: single laser track on a bare plate, 4.0 x 1.2 x 0.6 mm on a
400 x 160 x 80 grid (5.12 M interior cells), P = 300 W, r = 60 um, 0.8 m/s,
dt = 2e-6 s, SIMPLE to residual 1e-6 with maxit 100, plain FP64 in the
reference's program order.  A "step" is one time step: laser update,
iterativeLoopSolve, push-back.  Inputs are built on the device at creation;
the timed region is K steps after W untimed warm-up steps (the melt pool
forms during warm-up).

Prints one JSON line (metric, value = interior cell-steps per second, ...)
with a roofline block for the dominant kernel (the energy x-line sweep,
timed with HIP events on the solver stream) and a CPU baseline measured on
a bounded sample of the same workload with the C restatement.
"""
import argparse
import importlib
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

METRIC = "cell-steps/s on 1 GPU, mixed+warm-start vs plain FP64, ≤1% melt-pool dims"
HBM_PEAK_GBS = 8000.0
CASE = "track400"


def sweep_bytes_per_launch(dom, elem=8):
    """Algorithmic HBM bytes of one whole-domain x-line sweep: 4 passes x
    (8 coefficient reads + 1 read and 1 write of the field) per cell."""
    return 4 * (8 + 1 + 1) * elem * dom


def cpu_baseline(deck_params, dom, gpu_iters_per_step, budget_s=20.0):
    """Time the C restatement (1 thread) on the first SIMPLE iterations of the
    same workload; scale to cell-steps/s with the GPU's iterations per step."""
    import oracle
    orc = oracle.Oracle(deck_params)
    t0 = time.perf_counter()
    iters = 0
    steps = 0
    while time.perf_counter() - t0 < budget_s and steps < 3:
        orc.step()
        steps += 1
        iters = orc.scalars()["itertot"]
    el = time.perf_counter() - t0
    per_iter = el / max(iters, 1)
    value = dom / (per_iter * max(gpu_iters_per_step, 1.0))
    orc.close()
    return {"value": value, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": "C restatement (oracle/), 1 thread, first %d step(s) = %d SIMPLE iterations "
                      "of the 400x160x80 track in %.1f s (%.3f s/iteration), scaled by the GPU's "
                      "%.1f iterations/step" % (steps, iters, el, per_iter, gpu_iters_per_step)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--case", default=CASE)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-probe-mb", type=float, default=0.0,
                    help="only run one device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solve is a single instance whose line sweeps are "
              "Gauss-Seidel ordered across the whole grid; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mb > 0:
        import torch
        n = int(args.copy_probe_mb * 1e6 / 8)
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8}))
        return

    import decks
    amcfd = importlib.import_module("am-cfd_amd")
    tmp = tempfile.mkdtemp(prefix="amcfd_bench_")
    deck_path = decks.write_case(args.case, tmp)
    deck = amcfd.Deck(deck_path)
    p = deck.params
    dom = (p.ni - 2) * (p.nj - 2) * (p.nk - 2)
    solver = amcfd.Solver(deck)
    for w in range(args.warmup):
        solver.step()
        print("warmup step %d: iters %d tpeak %.1f" % (w + 1, solver.scalars()["niter"],
                                                        solver.scalars()["tpeak"]), file=sys.stderr, flush=True)
    solver.set_profiling(True)
    s0 = solver.scalars()
    n0, ms0 = solver.kernel_stats("sweep_h")
    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.step()
        print("timed step %d done" % (k + 1), file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    s1 = solver.scalars()
    n1, ms1 = solver.kernel_stats("sweep_h")
    elapsed = t1 - t0
    iters = s1["itertot"] - s0["itertot"]
    value = dom * args.steps / elapsed
    launches = n1 - n0
    avg_ms = (ms1 - ms0) / max(launches, 1)
    algo = sweep_bytes_per_launch(dom)
    achieved = algo / (avg_ms * 1e-3) / 1e9 if launches else None
    other = {}
    for name in ("assemble_h", "h2t", "sweep_u", "sweep_v", "sweep_w", "sweep_p"):
        n, ms = solver.kernel_stats(name)
        if n:
            other[name] = {"launches": n, "avg_ms": ms / n}
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single track, bare plate, 400x160x80 cells (5.12M), P=300 W, r=60 um, "
                               "0.8 m/s, dt=2e-6 s, maxit 100, tol 1e-6",
                   "mode": "plain FP64, reference program order", "interior_cells": dom,
                   "iters_per_step": iters / args.steps, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_line_sweeps (enthalpy, whole domain)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "launches": launches, "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": algo},
        "kernels": other,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(p, dom, iters / args.steps, args.cpu_budget)
    else:
        result["cpu_baseline"] = None
    solver.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
