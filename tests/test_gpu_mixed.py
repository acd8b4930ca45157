"""Mixed-precision mode (AMCFD_MODE_MIXED) on the GPU: FP32 enthalpy
coefficients, delta-form line passes (FP64 residual, FP32 solve) in
parallel order.  It is not bit-compatible with the reference by design; the
tests hold it to the FP64 mode (itself bitwise equal to the oracle) within
stated bounds, check reproducibility and the line lengths it adds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb else float(np.linalg.norm(a))


def run_pair(amcfd, path, steps, schedule=None):
    ref = amcfd.Solver(path, mode=amcfd.MODE_FP64)
    mx = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    if schedule:
        mx.set_mixed_schedule(schedule)
    rows = []
    for step in range(1, steps + 1):
        print("mixed vs fp64 %s step %d" % (path, step), flush=True)
        ref.step()
        mx.step()
        sr, sm = ref.scalars(), mx.scalars()
        rows.append((step, sr["niter"], sm["niter"], rel_l2(mx.field("temp"), ref.field("temp")), sr, sm))
    return ref, mx, rows


def test_mixed_mode_is_reproducible(case_dir, amcfd):
    path = case_dir("tiny")
    out = []
    for _ in range(2):
        s = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
        s.run(max_steps=5)
        out.append({f: s.field(f) for f in ("enthalpy", "uvel", "pp")})
        s.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f


def test_mixed_mode_converging_steps_match_fp64(case_dir, amcfd):
    """Steps that converge: tiny_cond at powers that melt nothing, to 150
    iterations, so that every step of both modes ends on the reference's own
    tests.  The temperature stays within 1e-4 (relative L2) of the FP64
    mode's over the whole run (measured: 3.4e-5 and 4.8e-5), and the
    iteration counts agree on nearly every step."""
    for power in (30.0, 45.0):
        path = case_dir("tiny_cond", power=power, speed=1.0, maxit=150, finaltime=0.2e-3 - 1e-9,
                        lasertime=0.08e-3)
        ref = amcfd.Solver(path, mode=amcfd.MODE_FP64)
        mx = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
        worst, same, n = 0.0, 0, 0
        fin = False
        while not fin:
            fin = ref.step()
            mx.step()
            worst = max(worst, rel_l2(mx.field("temp"), ref.field("temp")))
            n += 1
            same += ref.scalars()["niter"] == mx.scalars()["niter"]
            assert ref.scalars()["niter"] < 150 and mx.scalars()["niter"] < 150
        print("P=%.0f W: %d steps, max T rel-L2 %.2e, equal iteration counts on %d" % (power, n, worst, same))
        assert worst <= 1e-4
        assert same >= 0.9 * n


def test_mixed_mode_conduction_tracks_fp64(case_dir, amcfd):
    """Conduction only, maxit 40: from step 6 on most steps end at maxit
    without converging, and there the answer depends on the iterative
    method itself.  The yardstick is the FP64 mode's own iteration error, the
    distance between its runs at maxit 40 and 80 (measured 1.9e-2 by step 20);
    the mixed mode's distance from the FP64 run stays within 1.5 times it
    (measured 1.2e-2), and within 3e-4 on the first five steps, which
    converge (measured 2.1e-4: this deck melts, without flow)."""
    p40 = amcfd.Deck(case_dir("tiny_cond"))   # parsed before the next deck overwrites the file
    p80 = amcfd.Deck(case_dir("tiny_cond", maxit=80))
    ref = amcfd.Solver(p40, mode=amcfd.MODE_FP64)
    mx = amcfd.Solver(p40, mode=amcfd.MODE_MIXED)
    longer = amcfd.Solver(p80, mode=amcfd.MODE_FP64)
    e_mx, e_it = [], []
    for step in range(1, 21):
        print("conduction step %d" % step, flush=True)
        for s in (ref, mx, longer):
            s.step()
        e_mx.append(rel_l2(mx.field("temp"), ref.field("temp")))
        e_it.append(rel_l2(longer.field("temp"), ref.field("temp")))
        print("step %d niter %d/%d/%d  mixed %.2e  FP64 maxit 80 vs 40 %.2e" % (
            step, ref.scalars()["niter"], mx.scalars()["niter"], longer.scalars()["niter"], e_mx[-1], e_it[-1]))
    assert max(e_mx[:5]) < 3e-4
    assert max(e_mx) <= 1.5 * max(e_it)
    assert np.isfinite(mx.field("enthalpy")).all()


def test_mixed_mode_molten_runs(case_dir, amcfd):
    """Molten single track (flow on, maxit-bound): finite, and the pool
    tracks the FP64 run's to within the iteration's own sensitivity."""
    ref, mx, rows = run_pair(amcfd, case_dir("tiny"), 12)
    for step, n64, nmx, eT, sr, sm in rows:
        print("step %d niter %d/%d T %.2e tpeak %.1f/%.1f depth %.3e/%.3e" % (
            step, n64, nmx, eT, sr["tpeak"], sm["tpeak"], sr["depth"], sm["depth"]))
    assert np.isfinite(mx.field("uvel")).all() and np.isfinite(mx.field("enthalpy")).all()
    assert max(r[3] for r in rows) < 5e-2


def test_mixed_mode_lines_beyond_the_fp64_cap(amcfd, tmp_path):
    """x-lines of 2000 interior cells: refused by the FP64 sweeps (1024 cap),
    solved by the mixed mode's wave solver; conduction only, compared with
    the oracle's FP64 run."""
    import decks
    import oracle
    oracle.load()
    decks.CASES["tiny_xlong"] = dict(
        mesh=(([0.1e-3, 4.0e-3, 0.1e-3], [10, 1980, 10], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [3, 3], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [2, 4], [1.0, 1.0])),
        deck=dict(finaltime=0.02e-3 - 1e-9, maxit=40, dt=5e-6, lasertime=0.02e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6, flow=False))
    path = decks.write_case("tiny_xlong", str(tmp_path))
    deck = amcfd.Deck(path)
    assert deck.params.ni - 2 == 2000
    with pytest.raises(amcfd.AmcfdError):
        amcfd.Solver(deck, mode=amcfd.MODE_FP64)
    mx = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    orc = oracle.Oracle(deck.params)
    for step in range(4):
        orc.step()
        mx.step()
        e = rel_l2(mx.field("temp"), orc.field("temp"))
        print("tiny_xlong step %d T rel-L2 vs oracle %.2e niter %d/%d" % (
            step + 1, e, mx.scalars()["niter"], orc.scalars()["niter"]))
        assert e < 1e-3


@pytest.mark.parametrize("case,steps", [("tiny", 8), ("tiny_cond", 8), ("tiny_tool", 8), ("tiny_powder", 30)])
def test_captured_loop_equals_host_loop(case_dir, amcfd, case, steps):
    """The mixed mode's SIMPLE loop from a captured graph (device-side box,
    flow decision and convergence test; blocks of 8 iterations, the rest of a
    block turned into no-ops) gives the bits of the host-driven loop, and the
    same iteration counts.  tiny_powder adds a powder layer mid-run, which
    rebuilds the graph."""
    path = case_dir(case)
    a = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    b = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    a.set_graph(False)
    b.set_graph(True, block=8)
    for step in range(steps):
        print("graph vs host %s step %d" % (case, step + 1), flush=True)
        fa, fb = a.step(), b.step()
        sa, sb = a.scalars(), b.scalars()
        assert sa["niter"] == sb["niter"] and sa["itertot"] == sb["itertot"], (step, sa["niter"], sb["niter"])
        assert sa["nk"] == sb["nk"]
        for key in ("tpeak", "ratio", "resorh", "resorm", "alen", "depth", "width"):
            assert sa[key] == sb[key] or (np.isnan(sa[key]) and np.isnan(sb[key])), (step, key)
        if fa:
            break
    for f in ("enthalpy", "temp", "uvel", "vvel", "wvel", "pressure", "pp"):
        assert np.array_equal(a.field(f), b.field(f)), f


@pytest.mark.parametrize("schedule", ["xyzx", "xxyz", "yzxx"])
def test_mixed_schedules_with_y_and_z_passes(case_dir, amcfd, schedule):
    """Pass schedules that alternate the implicit direction: y-line and
    z-line passes (one lane per line, Jacobi between lines) between the
    red-black x passes.  On the converging no-melt deck the result stays
    within 1e-4 of the FP64 mode's, like the default x-only schedule; on the
    molten track it stays finite and reproducible."""
    path = case_dir("tiny_cond", power=30.0, speed=1.0, maxit=150, finaltime=0.2e-3 - 1e-9, lasertime=0.08e-3)
    ref = amcfd.Solver(path, mode=amcfd.MODE_FP64)
    mx = amcfd.Solver(path, mode=amcfd.MODE_MIXED)
    mx.set_mixed_schedule(schedule)
    worst = 0.0
    fin = False
    while not fin:
        fin = ref.step()
        mx.step()
        worst = max(worst, rel_l2(mx.field("temp"), ref.field("temp")))
        assert mx.scalars()["niter"] < 150
    print("%s: max T rel-L2 %.2e" % (schedule, worst))
    assert worst <= 1e-4
    out = []
    for _ in range(2):
        s = amcfd.Solver(case_dir("tiny"), mode=amcfd.MODE_MIXED)
        s.set_mixed_schedule(schedule)
        s.run(max_steps=6)
        out.append(s.field("enthalpy"))
        assert np.isfinite(out[-1]).all()
        s.close()
    assert np.array_equal(out[0], out[1])


def test_solver_takes_its_launch_shapes_from_the_tuning_table(case_dir, amcfd, tmp_path, monkeypatch):
    """A mixed-mode solver starts from the library's tuning-table row for its
    interior cell count (am-cfd_amd/tuning.txt, written by autotune.py), and
    amcfd_set_graph still overrides it."""
    deck = amcfd.Deck(case_dir("tiny"))
    p = deck.params
    cells = (p.ni - 2) * (p.nj - 2) * (p.nk - 2)
    s = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    assert s.graph_settings() == (True,) + amcfd.tuning_choice(cells)
    s.set_graph(True, 16, 128)
    s.set_box_fused(False)
    s.set_line_batch(3)
    s.set_assembly_overlap(False)
    assert s.graph_settings() == (True, 16, 128, False, 3, False)


def test_wide_grids_run_without_yz_passes(amcfd, tmp_path):
    """y- and z-line lengths bound only the schedules that use y/z passes: a
    grid with 400 cells in y runs the default x-only schedule, and a y/z
    schedule on it is refused with AMCFD_E_CAPACITY."""
    import decks
    decks.CASES["tiny_wide_y"] = dict(
        mesh=(([0.3e-3], [12], [1.0]), ([4.0e-3], [400], [1.0]), ([0.15e-3], [6], [1.0])),
        deck=dict(finaltime=0.01e-3 - 1e-9, maxit=20, dt=5e-6, lasertime=0.01e-3, xstart=0.15e-3,
                  ystart=2.0e-3, speed=1.0, power=50.0, radius=50e-6, flow=False))
    s = amcfd.Solver(decks.write_case("tiny_wide_y", str(tmp_path)), mode=amcfd.MODE_MIXED)
    with pytest.raises(amcfd.AmcfdError) as e:
        s.set_mixed_schedule("xyzx")
    assert e.value.code == -7
    s.run()
    assert np.isfinite(s.field("temp")).all()


FLOW_FIELDS = ("enthalpy", "uvel", "vvel", "wvel", "pp", "pressure")


@pytest.mark.parametrize("case,steps", [("tiny", 12), ("tiny_vap", 10), ("tiny_tool", 10), ("tiny_powder", 30)])
def test_fused_box_passes_equal_the_half_pass_launches(case_dir, amcfd, case, steps):
    """The melt-pool box passes in one persistent launch per variable give
    the bits of one launch per red or black half pass, step after step, from
    the captured loop and from the host loop alike."""
    deck = amcfd.Deck(case_dir(case))
    fused = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    fused.set_box_fused(True)
    split = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    split.set_box_fused(False)
    host = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    host.set_box_fused(True)
    host.set_graph(False)
    for step in range(steps):
        print("fused box passes, %s step %d" % (case, step + 1), flush=True)
        for s in (fused, split, host):
            s.step()
        assert fused.scalars()["niter"] == split.scalars()["niter"] == host.scalars()["niter"]
        for f in FLOW_FIELDS:
            assert np.array_equal(fused.field(f), split.field(f)), (step, f)
            assert np.array_equal(fused.field(f), host.field(f)), (step, f)
    assert fused.scalars()["tpeak"] > fused.params.tsolid   # the box passes ran


def test_expired_waits_end_a_mixed_step(case_dir, amcfd):
    """The persistent box passes wait with a bound: with the device error
    flag raised, every wait gives up at once, every kernel ends and the step
    returns AMCFD_E_HIP."""
    s = amcfd.Solver(case_dir("tiny"), mode=amcfd.MODE_MIXED)
    s.run(max_steps=4)
    assert s.scalars()["tpeak"] > s.params.tsolid
    s.test_expire_waits()
    with pytest.raises(amcfd.AmcfdError) as e:
        s.step()
    assert e.value.code == -5


@pytest.mark.parametrize("inst", [0, 3])
def test_assembly_overlap_gives_the_same_bits(amcfd, tmp_path, inst):
    """The captured loop with the next iteration's boundary launch and far
    assembly tiles beside the flow part gives the bits of the loop without
    that overlap and of the host loop, on sweep instances whose steps end by
    convergence inside a block of iterations (the ghost values set ahead are
    put back), every field including the ghost cells."""
    import decks
    deck = amcfd.Deck(decks.write_sweep_instance(inst, str(tmp_path)))
    on = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    on.set_assembly_overlap(True)
    few = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    few.set_assembly_overlap(16)   # many tiles per workgroup
    off = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    off.set_assembly_overlap(False)
    host = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    host.set_graph(False)
    converged = 0
    for step in range(4):
        print("assembly overlap, instance %d step %d" % (inst, step + 1), flush=True)
        for s in (on, few, off, host):
            s.step()
        sa = on.scalars()
        converged += sa["niter"] < on.params.maxit
        for other in (few, off, host):
            sb = other.scalars()
            assert sa["niter"] == sb["niter"], (step, sa["niter"], sb["niter"])
            for key in ("tpeak", "ratio", "resorh", "resorm", "alen", "depth", "width"):
                assert sa[key] == sb[key] or (np.isnan(sa[key]) and np.isnan(sb[key])), (step, key, sa[key], sb[key])
            for f in ("enthalpy", "temp", "fracl", "uvel", "vvel", "wvel", "pressure", "pp"):
                assert np.array_equal(on.field(f), other.field(f)), (step, f)
    assert converged >= 1   # a step ended inside a block of iterations
