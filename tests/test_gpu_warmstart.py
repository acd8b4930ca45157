"""Learned warm start on the GPU: the BF16-MFMA network's loss and every
parameter gradient against the FP64 restatement (oracle/nn_oracle.py), and
training on trajectories the solver records."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import nn_oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def layer_slices():
    out, o = [], 0
    for ci, co, _, _ in nn_oracle.LAYERS:
        nw = co * 27 * ci
        out.append((slice(o, o + nw), slice(o + nw, o + nw + co)))
        o += nw + co
    return out


def gradient_errors(amcfd, bias, head_scale, seed, shape):
    net = amcfd.WarmStartNet(seed=seed)
    rng = np.random.default_rng(seed)
    p = net.params()
    sl = layer_slices()
    for l, (w, b) in enumerate(sl):
        p[b] = bias + rng.normal(size=p[b].size) * 0.05
    p[sl[8][0]] *= head_scale
    net.set_params(p)
    pb = net.params().astype(np.float64)
    for w, b in sl:   # the device computes with BF16 weights and FP32 biases
        pb[w] = nn_oracle.bf16(pb[w])
    N, D, H, W = shape
    x = nn_oracle.bf16(rng.normal(size=(N, D, H, W, 16)).astype(np.float32))
    tgt = rng.normal(size=(N, D, H, W, 2)).astype(np.float32)
    mask = (rng.uniform(size=(N, D, H, W)) > 0.2).astype(np.float32)
    print("network gradient check, bias %.2f" % bias, flush=True)
    lg, gg = net.loss_grad(x, tgt, mask)
    net.close()
    t64, m64 = tgt.astype(np.float64), mask.astype(np.float64)
    lo, go = nn_oracle.loss_grad(pb, x, t64, m64)                 # plain FP64
    le, ge = nn_oracle.loss_grad(pb, x, t64, m64, emulate=True)   # FP64 with the device's BF16 roundings
    print("loss gpu %.6e oracle %.6e emulated %.6e" % (lg, lo, le))
    errs = []
    for l, (w, b) in enumerate(sl):
        for name, s in (("w", w), ("b", b)):
            a, r, q = gg[s].astype(np.float64), go[s], ge[s]
            e = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            eq = np.linalg.norm(a - q) / max(np.linalg.norm(q), 1e-30)
            cos = float(a @ q / max(np.linalg.norm(a) * np.linalg.norm(q), 1e-30))
            print("layer %d %s rel-L2 vs FP64 %.3e vs emulated %.3e cos %.6f |g| %.3e" % (
                l, name, e, eq, cos, np.linalg.norm(r)))
            errs.append((l, name, e, eq, cos))
    return (lg, lo, le), errs


def test_network_gradient_matches_the_fp64_oracle(amcfd):
    """BF16 inputs, weights and activations with FP32 accumulation against
    FP64 on the same (BF16-rounded) weights and inputs.  With every ReLU
    active (positive biases: no kinks) the device agrees with FP64 that
    rounds to BF16 where the device does to 3% per layer, with a cosine of
    0.9995 or better (the deepest sums, over 8192 positions in layer 0, carry
    the last few ReLUs that rounding still switches): that pins the kernels.
    With He-initialised weights and small biases, a pre-activation
    within rounding of zero can switch its ReLU, which in the 1/8-resolution
    layers (a handful of positions per sum) moves a gradient by several
    percent; there the bound is 10% and a cosine of 0.99."""
    (lg, lo, le), errs = gradient_errors(amcfd, 1.0, 10.0, 11, (2, 16, 16, 16))   # biases of 1: few kinks
    assert abs(lg - le) <= 1e-3 * abs(le) and abs(lg - lo) <= 2e-2 * abs(lo)
    for l, name, e, eq, cos in errs:
        assert eq <= 3e-2 and cos >= 0.9995, (l, name, eq, cos)
    (lg, lo, le), errs = gradient_errors(amcfd, 0.0, 10.0, 3, (2, 8, 16, 16))
    assert abs(lg - le) <= 1e-3 * abs(le) and abs(lg - lo) <= 2e-2 * abs(lo)
    for l, name, e, eq, cos in errs:
        assert e <= 1e-1 and eq <= 1e-1 and cos >= 0.99, (l, name, e, eq, cos)


def cond_deck(case_dir, power, speed):
    return case_dir("tiny_cond", power=power, speed=speed)


def test_training_on_recorded_steps_lowers_the_loss(case_dir, amcfd):
    net = amcfd.WarmStartNet(seed=1)
    for power, speed in ((150.0, 0.8), (200.0, 1.0), (250.0, 1.2)):
        s = amcfd.Solver(cond_deck(case_dir, power, speed), mode=amcfd.MODE_MIXED)
        net.attach(s, amcfd.WS_RECORD)
        s.run()
        s.close()
    n = net.samples()
    print("recorded samples", n, flush=True)
    assert n == 60
    first, last = net.train(200, batch=8, lr=1e-3)
    print("loss %.4e -> %.4e" % (first, last))
    assert last < 0.5 * first
    net.close()


def no_melt_deck(case_dir, power, speed):
    """tiny_cond at a power that melts nothing (peak T below the solidus), to
    150 iterations: every step ends on the reference's own tests (heat balance
    within 1% with the laser on, resorh < 1e-6 after it), so the number of
    iterations measures how far each step starts from its solution."""
    return case_dir("tiny_cond", power=power, speed=speed, maxit=150, finaltime=0.2e-3 - 1e-9,
                    lasertime=0.08e-3)


def run_iters(amcfd, path, net=None):
    s = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    if net is not None:
        net.attach(s, amcfd.WS_APPLY)
    its, on = [], []
    fin = False
    while not fin:
        fin = s.step()
        sc = s.scalars()
        its.append(sc["niter"])
        on.append(sc["powerindicator"])
    T = s.field("temp")
    s.close()
    return np.array(its), np.array(on), T


def test_warm_start_shortens_converging_steps(case_dir, amcfd):
    """Ablation: trained on twelve recorded runs (four powers x three speeds),
    applied to two held-out ones.  The warm-started steps converge on the same
    tests in fewer iterations (laser on and cooling), to the same answer
    within the tolerance the tests allow."""
    net = amcfd.WarmStartNet(seed=1)
    for power in (20.0, 30.0, 40.0, 50.0):
        for speed in (0.7, 1.0, 1.3):
            s = amcfd.Solver(no_melt_deck(case_dir, power, speed), mode=amcfd.MODE_MIXED)
            net.attach(s, amcfd.WS_RECORD)
            s.run()
            s.close()
    print("recorded", net.samples(), flush=True)
    a, b = net.train(1500, batch=8, lr=1e-3)
    c, d = net.train(1000, batch=8, lr=2e-4)
    print("loss %.3e -> %.3e -> %.3e" % (a, b, d))
    for power, speed in ((35.0, 0.9), (45.0, 1.15)):
        it0, on, T0 = run_iters(amcfd, no_melt_deck(case_dir, power, speed))
        it1, on1, T1 = run_iters(amcfd, no_melt_deck(case_dir, power, speed), net)
        assert np.array_equal(on, on1)
        assert it0.max() < 150 and it1.max() < 150   # every step converged, both ways
        e = float(np.linalg.norm(T1 - T0) / np.linalg.norm(T0))
        print("P=%.0f v=%.2f laser on: %d -> %d iterations, cooling: %d -> %d, T rel-L2 %.1e" % (
            power, speed, it0[on == 1].sum(), it1[on == 1].sum(), it0[on == 0].sum(), it1[on == 0].sum(), e))
        assert it1[on == 1].sum() < 0.7 * it0[on == 1].sum()
        assert it1[on == 0].sum() < 0.9 * it0[on == 0].sum()
        assert e < 3e-3
    net.close()


def test_guard_keeps_maxit_bound_runs_on_the_plain_start(case_dir, amcfd):
    """Mode 2 starts a step from the guess only after a converged step: on
    the molten single track cut to 10 iterations a step, whose steps end at
    maxit once the pool has formed, no step starts from it and the run is
    bit for bit the plain one; mode 3 starts every step from it.  On the
    no-melt deck, whose steps converge, mode 2 uses the guess on every step
    but the first."""
    path = case_dir("tiny", maxit=10)
    net = amcfd.WarmStartNet(seed=5)
    plain = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    guarded = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    always = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    for s in (plain, guarded, always):   # the first step converges, before the pool forms
        s.run(max_steps=3)
    net.attach(guarded, amcfd.WS_APPLY)
    net.attach(always, amcfd.WS_APPLY_ALWAYS)
    prev = [plain.scalars()["niter"]]
    for step in range(3, 9):
        print("guard, molten track, step %d" % (step + 1), flush=True)
        for s in (plain, guarded, always):
            s.step()
        prev.append(plain.scalars()["niter"])
    print("iterations per step", prev)
    assert all(n == 10 for n in prev), prev   # every step from the third on ends at maxit
    assert guarded.warm_start_counts() == (0, 6)
    assert always.warm_start_counts() == (6, 0)
    for f in ("enthalpy", "uvel", "vvel", "wvel", "pp"):
        assert np.array_equal(plain.field(f), guarded.field(f)), f
    assert not np.array_equal(plain.field("enthalpy"), always.field("enthalpy"))
    net.set_factors(tfactor=0.0)   # untrained: apply a zero change, so that every step still converges
    s = amcfd.Solver(no_melt_deck(case_dir, 30.0, 1.0), mode=amcfd.MODE_MIXED)
    net.attach(s, amcfd.WS_APPLY)
    n = s.run()
    assert s.warm_start_counts() == (n - 1, 1)
    net.close()
