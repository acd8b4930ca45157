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
import subprocess
import sys
import tempfile
import time

# the batched leg runs one stream per instance; HIP maps streams onto
# GPU_MAX_HW_QUEUES hardware queues (4 by default), raised here to 32 before
# anything initialises HIP; with 64 instances, 32 workers on 32 queues move
# about 15% more cell-iterations per second than 16 on 16.  The batched
# leg's one-at-a-time comparison runs first, in a child process with the
# default 4: while any process holds 32 queues, single-stream runs of the
# small instances take about 1.7x as long (2.0 against 3.3 M cell-steps/s).
SOLO_CHILD = "--batch-solo-child" in sys.argv
if not SOLO_CHILD and int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 32:
    os.environ["GPU_MAX_HW_QUEUES"] = "32"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

METRIC = "cell-steps/s on 1 GPU, mixed+warm-start vs plain FP64, ≤1% melt-pool dims"
# HBM bytes per launch of the enthalpy sweep on the 400x160x80 track, from two
# rocprofv3 --pmc runs (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00, factors from a
# plain 512 MB device copy): profiles/k_line_sweeps_track400_hbm_traffic.md
SWEEP_TRAFFIC_BYTES = {"track400": 2.129e9 + 0.218e9}
HBM_PEAK_GBS = 8000.0
CASE = "track400"


def sweep_bytes_per_launch(dom, elem=8):
    """Algorithmic HBM bytes of one whole-domain x-line sweep: 4 passes x
    (8 coefficient reads + 1 read and 1 write of the field) per cell."""
    return 4 * (8 + 1 + 1) * elem * dom


def cpu_baseline_reference(case, dom, gpu_iters_per_step, threads, steps):
    """Time the reference's own solver classes (oracle/_ref/ref_harness, built
    from the reference sources with -O3 -fopenmp) on the first `steps` time
    steps of the same deck with `threads` OpenMP threads; scale the measured
    seconds per SIMPLE iteration to cell-steps/s with the GPU's iterations per
    step.  Returns None when the harness is not built."""
    import decks
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    tmp = tempfile.mkdtemp(prefix="amcfd_refcpu_")
    deck = decks.write_case(case, tmp)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([exe, os.path.basename(deck), ".", "0", str(steps)], cwd=tmp, env=env,
                         capture_output=True, text=True, timeout=900, check=True)
    info = json.loads(out.stdout.strip().splitlines()[-1])
    per_iter = info["compute_s"] / max(info["itertot"], 1)
    value = dom / (per_iter * max(gpu_iters_per_step, 1.0))
    return {"value": value, "unit": "cell-steps/s", "cores": threads, "kind": "reference",
            "sample": "reference solver classes (CFDSolverManager etc., g++ -O3 -fopenmp), OMP_NUM_THREADS=%d, "
                      "first %d step(s) = %d SIMPLE iterations of the 400x160x80 track in %.1f s "
                      "(%.4f s/iteration), scaled by the GPU's %.1f iterations/step"
                      % (threads, info["steps"], info["itertot"], info["compute_s"], per_iter,
                         gpu_iters_per_step)}


def cpu_baseline(deck_params, dom, gpu_iters_per_step, budget_s=20.0):
    """Time the C restatement (1 thread) on the first SIMPLE iterations of the
    same workload; scale to cell-steps/s with the GPU's iterations per step.
    Used when the reference harness is not built."""
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


def temp_digest(solver):
    import hashlib
    return hashlib.sha256(solver.field("temp").tobytes()).hexdigest()


def batch_solo(amcfd, decks, tmp, n, warmup, steps):
    """Instances 0..n-1 one at a time: seconds over the timed steps and, per
    instance, its total iteration count and a digest of its temperature field."""
    ts, runs = 0.0, []
    for s in range(n):
        b = amcfd.Solver(amcfd.Deck(decks.write_sweep_instance(s, tmp)))
        b.run(max_steps=warmup)
        t0 = time.perf_counter()
        b.run(max_steps=steps)
        ts += time.perf_counter() - t0
        runs.append([b.scalars()["itertot"], temp_digest(b)])
        b.close()
    return {"seconds": ts, "runs": runs}


def batch_leg(amcfd, decks, tmp, n, warmup, steps, workers):
    import numpy as np
    """Instances 0..n-1 of the seeded parameter sweep (uneven grids, 12.5 um
    cells, P/v/r drawn per instance) in the plain FP64 mode: one at a time in
    a child process with the default hardware-queue count, then advanced side
    by side with amcfd_run_batch over the same steps.  Both runs compute the
    same bits; the ratio is the gain from sharing the device."""
    dks = [amcfd.Deck(decks.write_sweep_instance(s, tmp)) for s in range(n)]
    cells = [(d.params.ni - 2) * (d.params.nj - 2) * (d.params.nk - 2) for d in dks]
    work = sum(cells) * steps
    print("batch leg: %d instances, %d cells" % (n, sum(cells)), file=sys.stderr, flush=True)
    env = dict(os.environ, GPU_MAX_HW_QUEUES="4")
    cmd = [sys.executable, os.path.abspath(__file__), "--batch-solo-child", "--batch", str(n),
           "--batch-warmup", str(warmup), "--batch-steps", str(steps)]
    out = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, check=True, text=True, timeout=900).stdout
    solo = json.loads(out.strip().splitlines()[-1])
    ts = solo["seconds"]
    print("batch leg: one at a time %.3f s" % ts, file=sys.stderr, flush=True)
    side = [amcfd.Solver(d) for d in dks]
    amcfd.run_batch(side, warmup, workers)
    i0 = sum(s.scalars()["itertot"] for s in side)
    t0 = time.perf_counter()
    amcfd.run_batch(side, steps, workers)
    tb = time.perf_counter() - t0
    iters = sum(s.scalars()["itertot"] for s in side) - i0
    print("batch leg: side by side %.3f s" % tb, file=sys.stderr, flush=True)
    done = [[s.scalars()["itertot"], temp_digest(s)] for s in side]
    ref = [(s.field("temp"), s.scalars()) for s in side]
    for s in side:
        s.close()
    same = solo["runs"] == done
    # the same instances and steps in the scan mode, side by side, measured
    # against the FP64 results
    scan = [amcfd.Solver(d, mode=amcfd.MODE_SCAN) for d in dks]
    amcfd.run_batch(scan, warmup, workers)
    t0 = time.perf_counter()
    amcfd.run_batch(scan, steps, workers)
    tsc = time.perf_counter() - t0
    print("batch leg: scan mode side by side %.3f s" % tsc, file=sys.stderr, flush=True)
    def drift(solvers):
        l2, dims = 0.0, 0.0
        for s, (t64, s64) in zip(solvers, ref):
            t = s.field("temp")
            l2 = max(l2, float(np.linalg.norm(t - t64) / np.linalg.norm(t64)))
            for key in ("alen", "depth", "width"):
                if s64[key] > 0:
                    dims = max(dims, abs(s.scalars()[key] - s64[key]) / s64[key])
            s.close()
        return l2, dims
    l2, dims = drift(scan)
    # the floor: FP64 runs whose enthalpy is moved up by one ulp after the
    # warm-up steps, when the melt pool has formed (a one-ulp change of the
    # uniform initial field is mostly rounded away before it can grow)
    ulp = [amcfd.Solver(d) for d in dks]
    amcfd.run_batch(ulp, warmup, workers)
    for s in ulp:
        s.set_field("enthalpy", np.nextafter(s.field("enthalpy"), np.inf))
    amcfd.run_batch(ulp, steps, workers)
    l2u, dimsu = drift(ulp)
    return {"value": work / tb, "unit": "cell-steps/s", "instances": n, "workers": workers,
            "interior_cells": sum(cells),
            "warmup_steps": warmup, "timed_steps": steps, "iters_per_instance_step": iters / (n * steps),
            "one_at_a_time_value": work / ts, "speedup_vs_one_at_a_time": ts / tb,
            "identical_to_one_at_a_time": bool(same),
            "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), "one_at_a_time_hw_queues": 4,
            "scan_mode": {"value": work / tsc, "speedup_vs_fp64_one_at_a_time": ts / tsc,
                          "max_temp_rel_l2_vs_fp64": l2, "max_melt_pool_dim_rel_diff_vs_fp64": dims,
                          "fp64_one_ulp_floor": {"max_temp_rel_l2_vs_fp64": l2u,
                                                 "max_melt_pool_dim_rel_diff_vs_fp64": dimsu},
                          "note": "the same instances and steps with the x-line solves by parallel scans, side by "
                                  "side; temperature and melt-pool length/depth/width against the FP64 side-by-side "
                                  "results after the timed steps (the scan runs differ from step 1 on), next to the "
                                  "same measures for FP64 runs whose enthalpy was moved up by one ulp after the "
                                  "warm-up steps"},
            "note": "instances 0..n-1 of the seeded parameter sweep (numpy PCG64 seed 20261203+s), plain FP64 "
                    "mode, each on its own stream, taken largest first by the worker threads (amcfd_run_batch); "
                    "one at a time beforehand in a child process, same steps, compared by iteration count and a "
                    "digest of the temperature field"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--case", default=CASE)
    ap.add_argument("--mode", choices=["fp64", "scan"], default="fp64",
                    help="fp64: reference program order, bit for bit; scan: parallel line solves")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fast", action="store_true", help="skip the scan-mode leg")
    ap.add_argument("--batch", type=int, default=64,
                    help="instances of the seeded parameter sweep in the batched leg (0 skips it)")
    ap.add_argument("--batch-workers", type=int, default=32)
    ap.add_argument("--batch-warmup", type=int, default=5)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="OpenMP threads of the reference CPU baseline (the host cores used)")
    ap.add_argument("--cpu-steps", type=int, default=2, help="time steps in the CPU baseline sample")
    ap.add_argument("--copy-probe-mb", type=float, default=0.0,
                    help="only run one device copy of this size (counter calibration)")
    ap.add_argument("--batch-solo-child", action="store_true",
                    help="internal: run the batched leg's one-at-a-time comparison and print it as JSON")
    args = ap.parse_args()
    if args.batch_solo_child:
        import decks
        amcfd = importlib.import_module("am-cfd_amd")
        print(json.dumps(batch_solo(amcfd, decks, tempfile.mkdtemp(prefix="amcfd_solo_"), args.batch,
                                    args.batch_warmup, args.batch_steps)))
        return
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
    solver = amcfd.Solver(deck, mode=amcfd.MODE_SCAN if args.mode == "scan" else amcfd.MODE_FP64)
    for w in range(args.warmup):
        solver.step()
        print("warmup step %d: iters %d tpeak %.1f" % (w + 1, solver.scalars()["niter"],
                                                        solver.scalars()["tpeak"]), file=sys.stderr, flush=True)
    # the timed steps carry HIP events around the enthalpy sweep alone (the
    # roofline's kernel); the other families are timed in one more step
    # after the timed region
    solver.set_profiling(2)
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
    solver.set_profiling(1)
    solver.step()
    other = {}
    for name in ("assemble_h", "h2t", "sweep_u", "sweep_v", "sweep_w", "sweep_p"):
        n, ms = solver.kernel_stats(name)
        if n:
            other[name] = {"launches": n, "avg_ms": ms / n, "timed": "one step after the timed region"}
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
                   "mode": ("plain FP64, reference program order" if args.mode == "fp64" else
                            "FP64, x-line solves by parallel scans, reference sweep order"), "interior_cells": dom,
                   "iters_per_step": iters / args.steps, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_line_sweeps (enthalpy, whole domain)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": SWEEP_TRAFFIC_BYTES.get(args.case),
                     "launches": launches, "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": algo},
        "kernels": other,
    }
    solver.close()
    if args.mode == "fp64" and not args.no_fast:
        # the scan mode on the same workload, same warm-up and step count
        fast = amcfd.Solver(deck, mode=amcfd.MODE_SCAN)
        for w in range(args.warmup):
            fast.step()
        fast.set_profiling(2)
        f0 = fast.scalars()["itertot"]
        n0f, ms0f = fast.kernel_stats("sweep_h")
        t0 = time.perf_counter()
        for k in range(args.steps):
            fast.step()
        el = time.perf_counter() - t0
        n1f, ms1f = fast.kernel_stats("sweep_h")
        avgf = (ms1f - ms0f) / max(n1f - n0f, 1)
        result["fast_modes"] = {"scan": {
            "value": dom * args.steps / el, "ms_per_step": 1e3 * el / args.steps,
            "iters_per_step": (fast.scalars()["itertot"] - f0) / args.steps,
            "speedup_vs_fp64": (dom * args.steps / el) / value,
            "sweep_avg_launch_ms": avgf,
            "sweep_frac": algo / (avgf * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "note": "FP64, x-line solves by parallel scans in the reference's sweep order; drift vs the "
                    "FP64 mode equals a 1-ulp perturbation's (tests/drift_study.py)"}}
        fast.close()
    if args.mode == "fp64" and args.batch > 0:
        result.setdefault("fast_modes", {})["batch"] = batch_leg(amcfd, decks, tmp, args.batch, args.batch_warmup,
                                                                 args.batch_steps, args.batch_workers)
    result["cpu_baseline"] = None
    if not args.no_cpu:
        print("cpu baseline: reference harness, %d threads" % args.cpu_threads, file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline_reference(args.case, dom, iters / args.steps, args.cpu_threads,
                                                        args.cpu_steps)
        if result["cpu_baseline"] is None:
            result["cpu_baseline"] = cpu_baseline(p, dom, iters / args.steps, args.cpu_budget)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
