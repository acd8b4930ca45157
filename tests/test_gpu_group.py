"""Grouped lock-step runs (amcfd_run_grouped): mixed-mode instances of the
seeded parameter sweep on uneven grids, every kernel of an iteration one
launch over all of them, against each instance run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("enthalpy", "temp", "fracl", "uvel", "vvel", "wvel", "pressure", "pp")
SCALARS = ("timet", "niter", "itertot", "tpeak", "alen", "depth", "width", "ratio", "resorh")


def assert_same(a, b, tag):
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f), equal_nan=True), (tag, f)
    sa, sb = a.scalars(), b.scalars()
    for k in SCALARS:
        assert sa[k] == sb[k] or (np.isnan(sa[k]) and np.isnan(sb[k])), (tag, k, sa[k], sb[k])


def test_grouped_instances_are_the_solo_runs(amcfd, tmp_path, case_dir):
    """Five sweep instances (0.3-2.5 M cells, x-lines of different size
    classes) and the 22x14x14 single track, grouped for 3 steps and then 2
    more in a second call, give instance by instance the bits of running
    each alone for 5 steps."""
    import decks
    paths = [decks.write_sweep_instance(s, str(tmp_path)) for s in range(5)] + [case_dir("tiny")]
    dks = [amcfd.Deck(p) for p in paths]
    group = [amcfd.Solver(d, mode=amcfd.MODE_MIXED) for d in dks]
    print("grouped: %d instances, 3 steps" % len(group), flush=True)
    assert amcfd.run_grouped(group, 3) == [3] * len(group)
    print("grouped: 2 more steps", flush=True)
    assert amcfd.run_grouped(group, 2) == [2] * len(group)
    for s, (d, g) in enumerate(zip(dks, group)):
        print("solo instance %d" % s, flush=True)
        solo = amcfd.Solver(d, mode=amcfd.MODE_MIXED)
        solo.run(max_steps=5)
        assert_same(solo, g, s)
        solo.close()
        g.close()


def test_grouped_instances_finish_at_their_own_ends(amcfd, case_dir):
    """Instances whose runs end at different steps leave the group as they
    finish (the group is rebuilt without them) and the rest carry on."""
    short = amcfd.Solver(amcfd.Deck(case_dir("tiny_cond")), mode=amcfd.MODE_MIXED)
    long_ = amcfd.Solver(amcfd.Deck(case_dir("tiny")), mode=amcfd.MODE_MIXED)
    print("grouped until both finish", flush=True)
    done = amcfd.run_grouped([short, long_], -1)
    ref_s = amcfd.Solver(amcfd.Deck(case_dir("tiny_cond")), mode=amcfd.MODE_MIXED)
    ref_l = amcfd.Solver(amcfd.Deck(case_dir("tiny")), mode=amcfd.MODE_MIXED)
    ks = ref_s.run()
    kl = ref_l.run()
    assert done == [ks, kl], (done, ks, kl)
    assert_same(short, ref_s, "tiny_cond")
    assert_same(long_, ref_l, "tiny")


def test_grouped_refuses_solvers_it_cannot_take(amcfd, case_dir):
    deck = amcfd.Deck(case_dir("tiny"))
    mixed = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    fp64 = amcfd.Solver(deck, mode=amcfd.MODE_FP64)
    with pytest.raises(amcfd.AmcfdError):
        amcfd.run_grouped([mixed, fp64], 1)
    assert mixed.scalars()["step"] == 0   # refused before any step
    with pytest.raises(amcfd.AmcfdError):
        amcfd.run_grouped([mixed, mixed], 1)
    yz = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    yz.set_mixed_schedule("xyzx")
    with pytest.raises(amcfd.AmcfdError):
        amcfd.run_grouped([yz], 1)
