"""Throughput of the per-time-step SIMPLE solve on one MI355X.

Headline workload: single laser track on a bare plate, 4.0 x 1.2 x 0.6 mm on
a 400 x 160 x 80 grid (5.12 M interior cells), P = 300 W, r = 60 um,
0.8 m/s, dt = 2e-6 s, maxit 100, tolerance 1e-6.  A "step" is one time step
(laser update, iterativeLoopSolve, push-back) with the inputs resident on the
device; the timed region is K steps after W untimed warm-up steps (the melt
pool forms during warm-up).

The JSON line's value is the headline form: the mixed-precision mode
(FP32 enthalpy coefficients, delta-form line passes in parallel order, the
SIMPLE loop from a captured graph) with the learned warm start, trained on
steps recorded from the 64-instance parameter sweep.  Next to it:
  - the plain FP64 mode (the reference's program order, bit for bit) and the
    mixed mode without the warm start, on the same workload and steps;
  - drift: from one FP64 state after the warm-up, the mixed run and an FP64
    run moved by one ulp, over the same steps, against the plain FP64 run;
  - the 64 uneven sweep instances one at a time and side by side;
  - the warm start's ablation on steps that converge;
  - the roofline of the dominant kernel and the reference's CPU path.
"""
import argparse
import importlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time

# The side-by-side legs run one stream per instance; HIP maps streams onto
# GPU_MAX_HW_QUEUES hardware queues (4 by default), raised to 32 for them.
# They run in a child process of their own: once a process has spread work
# over 32 queues, its later single-stream runs are slower (measured: the
# 400x160x80 track ~1.8x), so the headline legs run where that never happened.
BATCH_CHILD = "--batch-child" in sys.argv
if BATCH_CHILD and int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 32:
    os.environ["GPU_MAX_HW_QUEUES"] = "32"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

METRIC = "cell-steps/s on 1 GPU, mixed+warm-start vs plain FP64, ≤1% melt-pool dims"
HBM_PEAK_GBS = 8000.0
CASE = "track400"
PROFILES = os.path.join(ROOT, "profiles")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def fp64_sweep_bytes(dom):
    """Algorithmic HBM bytes of one FP64 whole-domain sweep (k_line_sweeps):
    4 Gauss-Seidel passes x (8 coefficient reads + 1 read and 1 write of the
    field) x 8 B per cell."""
    return 4 * (8 + 1 + 1) * 8 * dom


def mixed_half_pass_bytes(dom):
    """Algorithmic HBM bytes of one red or black x-line half pass of the mixed
    mode (k_xline_mx): half the cells x (7 FP32 coefficients + FP64 source,
    read and write of the FP64 enthalpy) = 52 B per cell."""
    return (dom // 2) * (7 * 4 + 8 + 8 + 8)


def mixed_assembly_bytes(dom):
    """Algorithmic HBM bytes of the mixed mode's enthalpy assembly
    (k_energy_assemble_mx) on a deck without a volumetric source: reads u, v,
    w, rho, diff, hnot, fracl, fraclnot, h (9 FP64 fields), writes 7 FP32
    coefficients and FP64 su: 9 x 8 + 7 x 4 + 8 = 108 B per cell."""
    return dom * (9 * 8 + 7 * 4 + 8)


def profile_traffic(name):
    """HBM bytes per launch that a committed rocprofv3 PMC summary states (a
    line 'traffic_bytes_per_launch: <kernel> <bytes>'), or None."""
    for fn in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if not fn.endswith(".md"):
            continue
        for line in open(os.path.join(PROFILES, fn)):
            m = re.match(r"\s*traffic_bytes_per_launch:\s*(\S+)\s+([0-9.eE+]+)", line)
            if m and m.group(1) == name:
                return float(m.group(2)), fn
    return None


def cpu_baseline_reference(case, dom, gpu_iters_per_step, threads, steps):
    """The reference's own solver classes (oracle/_ref, built from the
    reference's sources with -O3 -fopenmp) on the first `steps` time steps of
    the same deck with `threads` OpenMP threads; seconds per SIMPLE iteration
    scaled to cell-steps/s with the GPU's iterations per step.  x-lines of
    more than 398 interior cells need the build with widened x capacity
    (oracle/Makefile, _ref/ref_harness_wide).  None if not built."""
    import decks
    tmp = tempfile.mkdtemp(prefix="amcfd_refcpu_")
    deck = decks.write_case(case, tmp)
    wide = decks.CASES[case]["mesh"][0][1] and sum(decks.CASES[case]["mesh"][0][1]) > 398
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_wide" if wide else "ref_harness")
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    proc = subprocess.run([exe, os.path.basename(deck), ".", "0", str(steps)], cwd=tmp, env=env,
                          capture_output=True, text=True, timeout=900, check=True)
    info = json.loads(proc.stdout.strip().splitlines()[-1])
    per_iter = info["compute_s"] / max(info["itertot"], 1)
    value = dom / (per_iter * max(gpu_iters_per_step, 1.0))
    build = ("reference sources with _nx_ raised to 512 (oracle/Makefile ref_harness_wide; the unmodified "
             "capacity of 400 x-nodes cannot hold this grid)" if wide else "reference sources, unmodified")
    return {"value": value, "unit": "cell-steps/s", "cores": threads, "kind": "reference",
            "sample": "reference solver classes (CFDSolverManager etc., g++ -O3 -fopenmp, %s), "
                      "OMP_NUM_THREADS=%d, first %d step(s) = %d SIMPLE iterations of the %s deck in %.1f s "
                      "(%.4f s/iteration), scaled by the GPU's %.1f iterations/step"
                      % (build, threads, info["steps"], info["itertot"], case, info["compute_s"], per_iter,
                         gpu_iters_per_step)}


def cpu_baseline_port(deck_params, dom, gpu_iters_per_step, budget_s=20.0):
    """The C restatement (oracle/, 1 thread) on the first iterations: used
    where the reference harness is not built."""
    import oracle
    orc = oracle.Oracle(deck_params)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s and steps < 3:
        orc.step()
        steps += 1
    iters = orc.scalars()["itertot"]
    el = time.perf_counter() - t0
    per_iter = el / max(iters, 1)
    orc.close()
    return {"value": dom / (per_iter * max(gpu_iters_per_step, 1.0)), "unit": "cell-steps/s", "cores": 1,
            "kind": "port", "sample": "C restatement (oracle/), 1 thread, first %d step(s) = %d iterations in %.1f s"
                                      % (steps, iters, el)}


def temp_digest(solver):
    import hashlib
    return hashlib.sha256(solver.field("temp").tobytes()).hexdigest()


def timed_steps(solver, k, label):
    s0 = solver.scalars()["itertot"]
    t0 = time.perf_counter()
    for i in range(k):
        solver.step()
        log("%s: timed step %d" % (label, i + 1))
    el = time.perf_counter() - t0
    return el, (solver.scalars()["itertot"] - s0) / k


def drift(np, a, ref):
    """Relative L2 per field and melt-pool dimension differences of solver a
    against solver ref."""
    out = {}
    for f in ("temp", "enthalpy", "uvel", "vvel", "wvel", "pressure"):
        x, y = a.field(f), ref.field(f)
        n = float(np.linalg.norm(y))
        out["rel_l2_" + f] = float(np.linalg.norm(x - y) / n) if n else float(np.linalg.norm(x))
    sa, sr = a.scalars(), ref.scalars()
    for k in ("alen", "depth", "width"):
        out["rel_diff_" + k] = abs(sa[k] - sr[k]) / sr[k] if sr[k] else abs(sa[k])
    out["rel_diff_tpeak"] = abs(sa["tpeak"] - sr["tpeak"]) / sr["tpeak"]
    return out


def batch_solo(amcfd, decks, tmp, n, warmup, steps):
    """Instances 0..n-1 one at a time in the FP64 mode: seconds over the
    timed steps, and per instance its iteration count and temperature digest."""
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


def batch_grouped(amcfd, decks, tmp, n, warmup, steps):
    """Instances 0..n-1 in the mixed mode, grouped (amcfd_run_grouped: every
    kernel of an iteration one launch over all instances, captured): seconds
    over the timed steps, iterations, and per instance its iteration count
    and temperature digest."""
    ss = [amcfd.Solver(amcfd.Deck(decks.write_sweep_instance(s, tmp)), mode=amcfd.MODE_MIXED) for s in range(n)]
    amcfd.run_grouped(ss, warmup)
    i0 = sum(s.scalars()["itertot"] for s in ss)
    t0 = time.perf_counter()
    amcfd.run_grouped(ss, steps)
    tg = time.perf_counter() - t0
    iters = sum(s.scalars()["itertot"] for s in ss) - i0
    runs = [[s.scalars()["itertot"], temp_digest(s)] for s in ss]
    for s in ss:
        s.close()
    return tg, iters, runs


def batch_side_by_side(amcfd, decks, tmp, n, warmup, steps, workers, record_steps, train_iters, ws_out):
    """The seeded parameter sweep side by side (amcfd_run_batch, one stream
    per instance) in the FP64 mode and in the mixed mode (captured loops);
    the mixed instances then record the warm start's training samples over
    record_steps more steps, and the network is trained on them and saved to
    ws_out.  Runs in its own process (32 hardware queues)."""
    dks = [amcfd.Deck(decks.write_sweep_instance(s, tmp)) for s in range(n)]
    cells = [(d.params.ni - 2) * (d.params.nj - 2) * (d.params.nk - 2) for d in dks]
    net = amcfd.WarmStartNet(seed=7) if record_steps > 0 else None

    def side_by_side(mode, record):
        ss = [amcfd.Solver(d, mode=mode) for d in dks]
        amcfd.run_batch(ss, warmup, workers)
        i0 = sum(s.scalars()["itertot"] for s in ss)
        t0 = time.perf_counter()
        amcfd.run_batch(ss, steps, workers)
        tb = time.perf_counter() - t0
        iters = sum(s.scalars()["itertot"] for s in ss) - i0
        runs = [[s.scalars()["itertot"], temp_digest(s)] for s in ss]
        if record:
            # training samples from the steps after the timed ones (the
            # melt pool has formed; recording stays out of the timed region)
            for s in ss:
                net.attach(s, amcfd.WS_RECORD)
            amcfd.run_batch(ss, record, workers)
        for s in ss:
            s.close()
        return tb, iters, runs

    tb, iters, runs = side_by_side(amcfd.MODE_FP64, 0)
    log("batch: FP64 side by side %.3f s" % tb)
    tm, itm, mruns = side_by_side(amcfd.MODE_MIXED, record_steps)
    log("batch: mixed side by side %.3f s" % tm)
    out = {"cells": cells, "fp64": [tb, iters, runs], "mixed": [tm, itm, mruns], "train": None}
    if net is not None and net.samples() > 0:
        t0 = time.perf_counter()
        n1 = (3 * train_iters) // 5
        l0, _ = net.train(n1, batch=8, lr=1e-3)
        _, l1 = net.train(train_iters - n1, batch=8, lr=2e-4)
        out["train"] = {"samples": net.samples(), "iterations": train_iters, "batch": 8,
                        "schedule": "Adam, lr 1e-3 for %d iterations then 2e-4" % n1, "loss_first": l0,
                        "loss_last": l1, "seconds": time.perf_counter() - t0,
                        "from": "%d steps of each of the %d sweep instances in the mixed mode, after their "
                                "%d warm-up and %d timed steps" % (record_steps, n, warmup, steps)}
        log("warm start trained on %d samples: loss %.3e -> %.3e" % (net.samples(), l0, l1))
        net.save(ws_out)
    if net is not None:
        net.close()
    return out


def batch_leg(amcfd, decks, tmp, args, ws_out):
    """The 64 sweep instances (uneven grids, 12.5 um cells, P/v/r drawn per
    instance): FP64 one at a time here, then the side-by-side legs and the
    warm start's training in a child process."""
    n, warmup, steps = args.batch, args.batch_warmup, args.batch_steps
    log("batch leg: %d instances one at a time" % n)
    solo = batch_solo(amcfd, decks, tmp, n, warmup, steps)
    ts = solo["seconds"]
    log("batch leg: FP64 one at a time %.3f s" % ts)
    tg, itg, gruns = batch_grouped(amcfd, decks, tmp, n, warmup, steps)
    log("batch leg: mixed grouped %.3f s" % tg)
    cmd = [sys.executable, os.path.abspath(__file__), "--batch-child", "--batch", str(n), "--batch-warmup",
           str(warmup), "--batch-steps", str(steps), "--batch-workers", str(args.batch_workers),
           "--record-steps", str(args.record_steps), "--train-iters", str(args.train_iters), "--ws-out", ws_out]
    child = subprocess.run(cmd, stdout=subprocess.PIPE, check=True, text=True, timeout=1200).stdout
    c = json.loads(child.strip().splitlines()[-1])
    work = sum(c["cells"]) * steps
    tb, iters, runs = c["fp64"]
    tm, itm, mruns = c["mixed"]
    res = {"instances": n, "interior_cells": sum(c["cells"]), "warmup_steps": warmup, "timed_steps": steps,
           "workers": args.batch_workers,
           "fp64_one_at_a_time": {"value": work / ts, "seconds": ts, "hw_queues": 4},
           "fp64_side_by_side": {"value": work / tb, "seconds": tb, "iters_per_instance_step": iters / (n * steps),
                                 "identical_to_one_at_a_time": runs == solo["runs"],
                                 "speedup_vs_fp64_one_at_a_time": ts / tb, "hw_queues": 32},
           "mixed_side_by_side": {"value": work / tm, "seconds": tm, "iters_per_instance_step": itm / (n * steps),
                                  "speedup_vs_fp64_one_at_a_time": ts / tm, "speedup_vs_fp64_side_by_side": tb / tm,
                                  "hw_queues": 32},
           "mixed_grouped": {"value": work / tg, "seconds": tg, "iters_per_instance_step": itg / (n * steps),
                             "identical_to_side_by_side": gruns == mruns,
                             "speedup_vs_fp64_one_at_a_time": ts / tg, "speedup_vs_fp64_side_by_side": tb / tg,
                             "speedup_vs_mixed_side_by_side": tm / tg, "hw_queues": 4},
           "note": "instances 0..n-1 of the seeded sweep (numpy PCG64 seed 20261203+s); grouped = "
                   "amcfd_run_grouped, the instances in lock step with every kernel of an iteration one launch "
                   "over all of them (one stream, captured graph, in this process); side by side = "
                   "amcfd_run_batch, each instance on its own stream, the mixed instances running their SIMPLE "
                   "loops from captured graphs (a child process with 32 hardware queues); FP64 side by side "
                   "compared with one at a time by iteration count and temperature digest"}
    return res, c["train"]


def ablation_leg(amcfd, decks, tmp):
    """The warm start on steps that converge: tiny_cond at powers that melt
    nothing, to 150 iterations, so every step ends on the reference's own
    tests; trained on 12 recorded runs, applied to 2 held-out ones."""
    import numpy as np

    def deck(power, speed):
        return decks.write_case("tiny_cond", os.path.join(tmp, "abl_%g_%g" % (power, speed)), power=power,
                                speed=speed, maxit=150, finaltime=0.2e-3 - 1e-9, lasertime=0.08e-3)
    net = amcfd.WarmStartNet(seed=1)
    for power in (20.0, 30.0, 40.0, 50.0):
        for speed in (0.7, 1.0, 1.3):
            s = amcfd.Solver(deck(power, speed), mode=amcfd.MODE_MIXED)
            net.attach(s, amcfd.WS_RECORD)
            s.run()
            s.close()
    l0, l1 = net.train(1500, batch=8, lr=1e-3)
    _, l2 = net.train(1000, batch=8, lr=2e-4)
    out = {"train_samples": net.samples(), "loss_first": l0, "loss_last": l2, "held_out": []}
    for power, speed in ((35.0, 0.9), (45.0, 1.15)):
        row = {"power_w": power, "speed_m_s": speed}
        for label, use in (("plain", False), ("warm_start", True)):
            s = amcfd.Solver(deck(power, speed), mode=amcfd.MODE_MIXED)
            if use:
                net.attach(s, amcfd.WS_APPLY)
            its, on = [], []
            t0 = time.perf_counter()
            fin = False
            while not fin:
                fin = s.step()
                sc = s.scalars()
                its.append(sc["niter"])
                on.append(sc["powerindicator"])
            el = time.perf_counter() - t0
            row[label] = {"iterations_laser_on": int(sum(i for i, p in zip(its, on) if p == 1)),
                          "iterations_cooling": int(sum(i for i, p in zip(its, on) if p == 0)),
                          "max_iterations_per_step": int(max(its)), "seconds": el}
            if use:
                row[label]["steps_from_guess"], row[label]["steps_guarded"] = s.warm_start_counts()
            row[label + "_temp"] = s.field("temp")
            s.close()
        a, b = row.pop("warm_start_temp"), row.pop("plain_temp")
        row["temp_rel_l2_warm_vs_plain"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        out["held_out"].append(row)
    net.close()
    out["note"] = ("tiny_cond (conduction only) at 20-50 W: nothing melts, every step converges (heat balance "
                   "within 1% with the laser on, resorh < 1e-6 after it) within maxit 150; the molten decks "
                   "run to maxit, where a better start cannot shorten a step")
    return out


def profile_leg(args):
    """Steps of the workload for a profiler: a calibration copy, then the
    solver's own kernels (host-driven loop, so each launch is a dispatch)."""
    import torch
    import decks
    amcfd = importlib.import_module("am-cfd_amd")
    tmp = tempfile.mkdtemp(prefix="amcfd_prof_")
    deck = amcfd.Deck(decks.write_case(args.case, tmp))
    mode = amcfd.MODE_MIXED if args.profile_leg == "mixed" else amcfd.MODE_FP64
    s = amcfd.Solver(deck, mode=mode)
    if mode == amcfd.MODE_MIXED:
        s.set_graph(False)
    for w in range(args.warmup):
        s.step()
        log("profile leg: warm-up step %d" % (w + 1))
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")   # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.step()
        log("profile leg: step %d" % (k + 1))
    el = time.perf_counter() - t0
    sc = s.scalars()
    print(json.dumps({"profile_leg": args.profile_leg, "case": args.case, "steps": args.steps,
                      "ms_per_step": 1e3 * el / args.steps, "itertot": sc["itertot"],
                      "calibration": "3 x torch copy_ of 2^28 float32 (1 GiB read + 1 GiB write each)"}))
    s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--case", default=CASE)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch", type=int, default=64,
                    help="instances of the seeded parameter sweep in the batch leg (0 skips it)")
    ap.add_argument("--batch-workers", type=int, default=16)
    ap.add_argument("--batch-warmup", type=int, default=5)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--record-steps", type=int, default=3,
                    help="steps of each mixed sweep instance recorded for the warm start after the timed ones")
    ap.add_argument("--train-iters", type=int, default=2000)
    ap.add_argument("--no-ablation", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="OpenMP threads of the reference CPU baseline (the host cores used)")
    ap.add_argument("--cpu-steps", type=int, default=6, help="time steps in the CPU baseline sample")
    ap.add_argument("--profile-leg", choices=("mixed", "fp64"), default=None,
                    help="only run --warmup and --steps steps of the workload in this mode with the host-driven "
                         "loop (every kernel its own dispatch), after a device copy of 1 GiB (read 1 GiB, write "
                         "1 GiB) that calibrates memory counters; for rocprofv3")
    ap.add_argument("--batch-child", action="store_true",
                    help="internal: the batch leg's side-by-side runs and the warm start's training, as JSON")
    ap.add_argument("--ws-out", default=None, help="internal: where the batch child saves the trained network")
    args = ap.parse_args()
    if args.batch_child:
        import decks
        amcfd = importlib.import_module("am-cfd_amd")
        print(json.dumps(batch_side_by_side(amcfd, decks, tempfile.mkdtemp(prefix="amcfd_side_"), args.batch,
                                            args.batch_warmup, args.batch_steps, args.batch_workers,
                                            args.record_steps, args.train_iters, args.ws_out)))
        return
    if args.profile_leg:
        profile_leg(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the solve of one instance runs on one device (its line sweeps span "
              "the whole grid); --gpus must be 1. Several instances share a device with amcfd_run_grouped.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import decks
    amcfd = importlib.import_module("am-cfd_amd")
    tmp = tempfile.mkdtemp(prefix="amcfd_bench_")
    result = {"metric": METRIC}

    # -- the sweep; the warm start trained on its mixed-mode instances ----------
    net = amcfd.WarmStartNet(seed=7)
    batch, train = None, None
    if args.batch > 0:
        ws_path = os.path.join(tmp, "warm_start.bin")
        batch, train = batch_leg(amcfd, decks, tmp, args, ws_path)
        if train:
            net.load(ws_path)

    # -- the 400x160x80 track: FP64, mixed, mixed + warm start --------------------
    deck = amcfd.Deck(decks.write_case(args.case, tmp))
    p = deck.params
    dom = (p.ni - 2) * (p.nj - 2) * (p.nk - 2)
    fp = amcfd.Solver(deck, mode=amcfd.MODE_FP64)
    for w in range(args.warmup):
        fp.step()
        log("fp64 warm-up step %d: %d iterations, tpeak %.1f" % (w + 1, fp.scalars()["niter"], fp.scalars()["tpeak"]))
    # the drift runs start from this state
    mx_d = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    mx_d.copy_state_from(fp)
    ulp = amcfd.Solver(deck, mode=amcfd.MODE_FP64)
    ulp.copy_state_from(fp)
    ulp.set_field("enthalpy", np.nextafter(ulp.field("enthalpy"), np.inf))
    ws_d = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    ws_d.copy_state_from(fp)
    # the FP64 mode's own iteration error on these steps: the same start with
    # twice the iterations per step
    it2_dir = os.path.join(tmp, "maxit2")
    os.makedirs(it2_dir, exist_ok=True)
    it2 = amcfd.Solver(amcfd.Deck(decks.write_case(args.case, it2_dir, maxit=2 * int(p.maxit))), mode=amcfd.MODE_FP64)
    it2.copy_state_from(fp)
    wsa_d = None
    if train:
        net.attach(ws_d, amcfd.WS_APPLY)
        # without the guard, for comparison: every step from the guess
        wsa_d = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
        wsa_d.copy_state_from(fp)
        net.attach(wsa_d, amcfd.WS_APPLY_ALWAYS)
    # FP64 timed steps (events around the enthalpy sweep alone)
    fp.set_profiling(2)
    n0, ms0 = fp.kernel_stats("sweep_h")
    el64, it64 = timed_steps(fp, args.steps, "fp64")
    n1, ms1 = fp.kernel_stats("sweep_h")
    v64 = dom * args.steps / el64
    sweep64_ms = (ms1 - ms0) / max(n1 - n0, 1)
    runs = [(mx_d, "mixed (drift)"), (ulp, "fp64 + 1 ulp"), (ws_d, "mixed + warm start (drift)"),
            (it2, "fp64 at twice maxit")]
    if wsa_d is not None:
        runs.append((wsa_d, "mixed + unguarded warm start (drift)"))
    for s, label in runs:
        for k in range(args.steps):
            s.step()
        log("%s: %d steps" % (label, args.steps))
    drift_block = {"from": "one FP64 state after %d warm-up steps, %d steps on" % (args.warmup, args.steps),
                   "mixed": drift(np, mx_d, fp), "mixed_warm_start": drift(np, ws_d, fp),
                   "fp64_one_ulp": drift(np, ulp, fp), "fp64_twice_maxit": drift(np, it2, fp),
                   "mixed_vs_fp64_twice_maxit": drift(np, mx_d, it2),
                   "note": "each run starts from the same FP64 state (amcfd_copy_state) and runs the same steps; "
                           "fp64_one_ulp moves that state's enthalpy up by one ulp; a step here runs to maxit "
                           "without reaching the tolerance, so a last-bit difference grows into this floor; "
                           "fp64_twice_maxit runs the FP64 mode with twice the iterations per step: its distance "
                           "from the plain FP64 run is the iteration error the maxit-bound steps carry; "
                           "mixed_vs_fp64_twice_maxit is the mixed run against that one"}
    drift_block["mixed_warm_start"]["steps_from_guess"], drift_block["mixed_warm_start"]["steps_guarded"] = \
        ws_d.warm_start_counts()
    if wsa_d is not None:
        drift_block["mixed_warm_start_unguarded"] = drift(np, wsa_d, fp)
        drift_block["mixed_warm_start_unguarded"]["note"] = (
            "every step started from the guess (mode 3); these steps end at maxit, so the start moves the "
            "answer: the reason mode 2 keeps such runs on the plain start")
        wsa_d.close()
    for s in (mx_d, ulp, ws_d, it2):
        s.close()
    fp.close()
    # mixed, and mixed + warm start, from the start of the run (captured loops)
    legs = {}
    for label, use_ws in (("mixed", False), ("mixed_warm_start", True)):
        s = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
        if use_ws and train:
            net.attach(s, amcfd.WS_APPLY)
        for w in range(args.warmup):
            s.step()
        el, it = timed_steps(s, args.steps, label)
        legs[label] = {"value": dom * args.steps / el, "ms_per_step": 1e3 * el / args.steps, "iters_per_step": it,
                       "speedup_vs_fp64": (dom * args.steps / el) / v64}
        if use_ws:
            a, g = s.warm_start_counts()
            legs[label].update({"trained": bool(train), "steps_from_guess": a, "steps_guarded": g,
                                "note": "mode 2: a step starts from the network's guess when the step before "
                                        "converged; every step of this workload ends at maxit (FP64 too), so the "
                                        "guard keeps the plain start and the iteration count is maxit either way; "
                                        "the warm start's gain shows on converging steps (warm_start_ablation)"})
        if label == "mixed":
            # kernel times from one more step, host-driven so events bracket them
            s.set_graph(False)
            s.set_profiling(1)
            s.step()
            for name in ("assemble_h", "sweep_h", "h2t"):
                n, ms = s.kernel_stats(name)
                legs[label][name + "_avg_ms"] = ms / max(n, 1)
        s.close()
    head = legs["mixed_warm_start"]

    # -- roofline of the dominant kernels of the headline path --------------------
    half_ms = legs["mixed"]["sweep_h_avg_ms"] / 8.0   # 4 red-black passes = 8 launches per event pair
    asm_ms = legs["mixed"]["assemble_h_avg_ms"]
    xb, ab = mixed_half_pass_bytes(dom), mixed_assembly_bytes(dom)
    scheme = {0: "0", 1: "1", 2: "2"}[p.discretization]
    cand = [("k_xline_mx<float,8>", "k_xline_mx<float, 8> (enthalpy x-lines, one red or black half pass)", xb,
             half_ms, 8),
            ("k_energy_assemble_mx<%s>" % scheme, "k_energy_assemble_mx (enthalpy coefficients, FP32)", ab, asm_ms, 1)]
    cand.sort(key=lambda c: -c[3] * c[4])   # largest device time per iteration first
    key, name, algo, ms, per_iter = cand[0]
    achieved = algo / (ms * 1e-3) / 1e9
    traffic = profile_traffic(key)
    roof = {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic[0] if traffic else None,
            "traffic_source": ("profiles/" + traffic[1]) if traffic else None,
            "avg_launch_ms": ms, "launches_per_iteration": per_iter, "algorithmic_bytes_per_launch": algo,
            "timing": "HIP events on the solver stream, host-driven iteration after the timed region"}
    other = []
    for _, nm, al, t, k in cand[1:]:
        other.append({"kernel": nm, "avg_launch_ms": t, "achieved_gbs": al / (t * 1e-3) / 1e9,
                      "frac": al / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, "launches_per_iteration": k,
                      "algorithmic_bytes_per_launch": al})
    a64 = fp64_sweep_bytes(dom) / (sweep64_ms * 1e-3) / 1e9
    other.append({"kernel": "k_line_sweeps<256,20,true,false> (FP64 mode, Gauss-Seidel order)",
                  "avg_launch_ms": sweep64_ms, "achieved_gbs": a64, "frac": a64 / HBM_PEAK_GBS,
                  "bound": "latency of the Gauss-Seidel dependency chain",
                  "algorithmic_bytes_per_launch": fp64_sweep_bytes(dom)})

    result.update({
        "value": head["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic",
        "config": {"workload": "single track, bare plate, 400x160x80 cells (5.12M), P=300 W, r=60 um, 0.8 m/s, "
                               "dt=2e-6 s, maxit 100, tol 1e-6",
                   "mode": "mixed precision (FP32 enthalpy coefficients, FP32 line solves on FP64 residuals, "
                           "red-black x-line passes), SIMPLE loop from captured graphs, learned warm start",
                   "interior_cells": dom, "iters_per_step": head["iters_per_step"],
                   "parallelism": "single GPU"},
        "speedup_vs_fp64": head["speedup_vs_fp64"],
        "fp64": {"value": v64, "ms_per_step": 1e3 * el64 / args.steps, "iters_per_step": it64,
                 "mode": "plain FP64, reference program order (bit for bit)"},
        "mixed": legs["mixed"], "mixed_warm_start": head,
        "drift": drift_block,
        "warm_start_training": train,
        "roofline": roof, "kernels": other,
    })
    if batch:
        result["batch"] = batch
    net.close()
    if not args.no_ablation:
        log("warm-start ablation")
        result["warm_start_ablation"] = ablation_leg(amcfd, decks, tmp)
    result["cpu_baseline"] = None
    if not args.no_cpu:
        log("cpu baseline: reference harness, %d threads" % args.cpu_threads)
        result["cpu_baseline"] = cpu_baseline_reference(args.case, dom, it64, args.cpu_threads, args.cpu_steps)
        if result["cpu_baseline"] is None:
            result["cpu_baseline"] = cpu_baseline_port(p, dom, it64)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
