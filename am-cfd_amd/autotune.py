"""Autotuner of the mixed mode's captured-loop launch shapes.

The captured SIMPLE loop of AMCFD_MODE_MIXED has shapes that change its speed
and never its results: the iterations per graph launch (the host reads the
scalars once per launch, and iterations after convergence are no-ops), the
workgroups that stride over the melt-pool box, whether a box variable's
red-black passes run in one persistent launch or one launch per half pass,
the line batch of the enthalpy's x-line passes (consecutive lines per
workgroup), and whether the next iteration's assembly runs beside the flow
part (amcfd_set_assembly_overlap).  Grouped runs (amcfd_run_grouped) launch each instance's kernels
with that instance's shapes, so the table serves them too.
This tool measures every candidate on one representative deck per grid-size
class, each solver alone (beside other solvers, fewer launches count for
more: the fused box passes measured 30% faster side by side and equal
alone), and writes the
table the library reads when it creates a mixed-mode solver (tuning.txt
beside libamcfd.so, or the file named by $AMCFD_TUNING):

    python am-cfd_amd/autotune.py small.k medium.k large.k [--out am-cfd_amd/tuning.txt]

For each deck it runs --warmup steps, then times --steps steps per candidate
from that same state (amcfd_copy_state), --repeat times, and keeps the median.
The graph block, box workgroups and box form are searched together; then the
line batch with the best of them, then the assembly overlap with the best of
those.
Every candidate's fields must equal the first candidate's bit for bit, or the
tool stops.  A candidate replaces the built-in shapes (8, 256, fused, 1, no overlap) only
when it is faster by more than --margin.  Class boundaries lie at the
geometric means of neighbouring decks' interior cell counts.  Needs one GPU.
"""
import argparse
import importlib
import math
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
amcfd = importlib.import_module(os.path.basename(HERE))

GRAPH_BLOCKS = (4, 8, 16)
BOX_BLOCKS = (128, 256, 512)
BOX_FUSED = (1, 0)
LINE_BATCHES = (2, 4, 8)
BUILT_IN = (8, 256, 1, 1, 0)
FIELDS = ("enthalpy", "uvel", "vvel", "wvel", "pp", "pressure")


def cells_of(deck):
    p = deck.params
    return (p.ni - 2) * (p.nj - 2) * (p.nk - 2)


def measure(deck, base, gb, bb, bf, lb, ov, steps, repeat):
    """Median seconds per step of `steps` steps from base's state, and the
    fields at the end of the last repetition."""
    times, fields = [], None
    for _ in range(repeat):
        s = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
        s.copy_state_from(base)
        s.set_graph(True, gb, bb)
        s.set_box_fused(bool(bf))
        s.set_line_batch(lb)
        s.set_assembly_overlap(ov)
        s.step()   # captures the graph; untimed
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step()
        times.append((time.perf_counter() - t0) / steps)
        fields = {f: s.field(f) for f in FIELDS}
        s.close()
    return statistics.median(times), fields


def tune_deck(path, warmup, steps, repeat, margin, log):
    import numpy as np
    deck = amcfd.Deck(path)
    base = amcfd.Solver(deck, mode=amcfd.MODE_MIXED)
    base.run(max_steps=warmup)
    results, ref = {}, [None]

    def candidate(shape):
        t, fields = measure(deck, base, *shape, steps, repeat)
        if ref[0] is None:
            ref[0] = fields
        else:
            for f in FIELDS:
                if not np.array_equal(fields[f], ref[0][f], equal_nan=True):
                    raise SystemExit("%s: graph_block %d, box_blocks %d, box_fused %d, line_batch %d, overlap %d "
                                     "changed %s: launch shapes must not change results" % ((path,) + shape + (f,)))
        results[shape] = t
        log("%s  graph_block %2d  box_blocks %4d  box_fused %d  line_batch %d  overlap %d  %.3f ms/step"
            % ((os.path.basename(path),) + shape + (1e3 * t,)))

    for bf in BOX_FUSED:
        for gb in GRAPH_BLOCKS:
            for bb in BOX_BLOCKS:
                candidate((gb, bb, bf, 1, 0))
    first = min(results, key=results.get)
    for lb in LINE_BATCHES:
        candidate(first[:3] + (lb, 0))
    second = min(results, key=results.get)
    candidate(second[:4] + (1,))
    base.close()
    best = min(results, key=results.get)
    if results[best] > (1.0 - margin) * results[BUILT_IN]:
        best = BUILT_IN
    return cells_of(deck), best, results


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("decks", nargs="+", help="one representative deck per grid-size class")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--margin", type=float, default=0.015,
                    help="least fractional gain over the built-in shapes (8, 256, 1, 1, 0) that replaces them")
    args = ap.parse_args()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    rows = []
    for path in args.decks:
        cells, best, res = tune_deck(path, args.warmup, args.steps, args.repeat, args.margin, log)
        rows.append((cells, best, res[best], res[BUILT_IN], os.path.basename(path)))
        log("%s: %d cells -> graph_block %d, box_blocks %d, box_fused %d, line_batch %d, overlap %d (%.3f ms/step; "
            "built-in %.3f)" % ((path, cells) + best + (1e3 * res[best], 1e3 * res[BUILT_IN])))
    rows.sort()
    lines = ["# Launch shapes of the mixed mode's captured loop, by grid size (interior cells).",
             "# Written by am-cfd_amd/autotune.py; read by libamcfd.so when a mixed-mode solver is created.",
             "# Command: python am-cfd_amd/autotune.py " + " ".join(os.path.basename(d) for d in args.decks)
             + " --warmup %d --steps %d --repeat %d --margin %g" % (args.warmup, args.steps, args.repeat, args.margin),
             "# max_cells  graph_block  box_blocks  box_fused  line_batch  overlap   # representative deck: cells, "
             "ms/step (built-in 8/256/1/1/0)"]
    for n, (cells, best, t, t0, name) in enumerate(rows):
        upper = int(math.sqrt(cells * rows[n + 1][0])) if n + 1 < len(rows) else 2 ** 62
        lines.append("%d %d %d %d %d %d   # %s: %d cells, %.3f ms/step (%.3f)"
                     % ((upper,) + best + (name, cells, 1e3 * t, 1e3 * t0)))
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
