"""State copy (amcfd_copy_state), the mixed mode side by side over instances
of the seeded parameter sweep, and the warm start recording from solvers
that step on several host threads at once."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("enthalpy", "temp", "fracl", "uvel", "vvel", "wvel", "pressure", "pp")
SCALARS = ("timet", "niter", "itertot", "tpeak", "alen", "depth", "width", "beamposx", "powerindicator")


def assert_same(a, b):
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f
    sa, sb = a.scalars(), b.scalars()
    for k in SCALARS:
        assert sa[k] == sb[k], k


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_copied_state_continues_bit_for_bit(case_dir, amcfd, mode):
    """A solver given another's state after 4 steps continues exactly as the
    original does: same fields and scalars after 4 more steps each (the
    molten single track, flow on; mixed mode with its captured loop)."""
    m = amcfd.MODE_MIXED if mode == "mixed" else amcfd.MODE_FP64
    deck = amcfd.Deck(case_dir("tiny"))
    a = amcfd.Solver(deck, mode=m)
    a.run(max_steps=4)
    b = amcfd.Solver(deck, mode=m)
    b.copy_state_from(a)
    assert_same(a, b)
    for step in range(4):
        print("copy_state %s step %d" % (mode, step + 5), flush=True)
        a.step()
        b.step()
    assert a.scalars()["step"] == b.scalars()["step"]
    assert_same(a, b)


def test_copied_state_across_modes(case_dir, amcfd):
    """An FP64 state taken over by a mixed-mode solver: the first steps after
    it stay near the FP64 run's own (the drift comparison of the bench)."""
    deck = amcfd.Deck(case_dir("tiny"))
    fp = amcfd.Solver(deck, mode=amcfd.MODE_FP64)
    fp.run(max_steps=4)
    mx = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    mx.copy_state_from(fp)
    assert np.array_equal(mx.field("enthalpy"), fp.field("enthalpy"))
    fp.step()
    mx.step()
    T, T0 = mx.field("temp"), fp.field("temp")
    e = float(np.linalg.norm(T - T0) / np.linalg.norm(T0))
    print("FP64 -> mixed, one step: T rel-L2 %.2e" % e)
    assert e < 1e-2


def test_copy_state_rejects_other_grids(case_dir, amcfd):
    a = amcfd.Solver(case_dir("tiny"))
    b = amcfd.Solver(case_dir("plate120"))
    with pytest.raises(amcfd.AmcfdError):
        b.copy_state_from(a)


def test_mixed_batch_of_sweep_instances_is_the_solo_runs(amcfd, tmp_path):
    """Six instances of the seeded sweep (uneven grids, 0.3-2.5 M cells each)
    in the mixed mode, side by side with their captured loops on six host
    threads, give instance by instance the bits of running each alone."""
    import decks
    dks = [amcfd.Deck(decks.write_sweep_instance(s, str(tmp_path))) for s in range(6)]
    print("mixed batch of 6 sweep instances, 4 steps", flush=True)
    batch = [amcfd.Solver(d, mode=amcfd.MODE_MIXED) for d in dks]
    assert amcfd.run_batch(batch, 4, 6) == [4] * 6
    for s, (d, b) in enumerate(zip(dks, batch)):
        print("solo instance %d" % s, flush=True)
        solo = amcfd.Solver(d, mode=amcfd.MODE_MIXED)
        solo.run(max_steps=4)
        assert solo.scalars()["itertot"] == b.scalars()["itertot"], s
        for f in ("enthalpy", "uvel", "vvel", "wvel", "pp"):
            assert np.array_equal(solo.field(f), b.field(f)), (s, f)
        solo.close()
        b.close()


def test_warm_start_records_from_concurrent_solvers(amcfd, tmp_path):
    """One network recording from four solvers stepping side by side keeps
    one sample per solver step, the solvers' results are those of solvers
    without it, and training on the samples runs; destroying the network
    while solvers are attached detaches them."""
    import decks
    dks = [amcfd.Deck(decks.write_sweep_instance(s, str(tmp_path))) for s in range(4)]
    net = amcfd.WarmStartNet(seed=3)
    rec = [amcfd.Solver(d, mode=amcfd.MODE_MIXED) for d in dks]
    for s in rec:
        net.attach(s, amcfd.WS_RECORD)
    print("recording from 4 solvers side by side, 3 steps", flush=True)
    assert amcfd.run_batch(rec, 3, 4) == [3] * 4
    assert net.samples() == 12
    plain = [amcfd.Solver(d, mode=amcfd.MODE_MIXED) for d in dks]
    amcfd.run_batch(plain, 3, 4)
    for a, b in zip(rec, plain):
        assert np.array_equal(a.field("enthalpy"), b.field("enthalpy"))
    l0, l1 = net.train(20, batch=4, lr=1e-3)
    print("loss %.3e -> %.3e" % (l0, l1))
    assert np.isfinite(l0) and np.isfinite(l1)
    net.close()
    for a, b in zip(rec, plain):
        a.step()   # no network left: a plain step
        b.step()
        assert np.array_equal(a.field("enthalpy"), b.field("enthalpy"))
