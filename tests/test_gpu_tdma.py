"""amcfd_tdma_lines: the reference's x-line TDMA over a batch of lines in
device memory, bit for bit against the recurrence of
CFDSolverManager.cpp:2500-2522 (and, with the guard, :2421-2450)."""
import numpy as np
import pytest

from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def reference_lines(aw, ap, ae, d, phi, n, nlines, guard):
    """The reference's recurrence, line by line, in Python floats (IEEE
    double, the same operations in the same order)."""
    phi = phi.copy()
    for l in range(nlines):
        b = l * n
        pr = [0.0] * n
        qr = [0.0] * n
        qr[0] = float(phi[b])
        for i in range(1, n - 1):
            c = b + i
            den = float(ap[c]) - float(aw[c]) * pr[i - 1]
            if guard and abs(den) <= 1e-12:
                den += 1e-6
            pr[i] = float(ae[c]) / den
            qr[i] = (float(d[c]) + float(aw[c]) * qr[i - 1]) / den
        for i in range(n - 2, 0, -1):
            phi[b + i] = pr[i] * phi[b + i + 1] + qr[i]
    return phi


def make_batch(rng, n, nlines, tiny_pivots=False):
    aw = rng.uniform(0.0, 1.0, n * nlines)
    ae = rng.uniform(0.0, 1.0, n * nlines)
    ap = aw + ae + rng.uniform(0.01, 0.5, n * nlines)
    if tiny_pivots:   # rows whose pivot the guard moves (solid-pinned cells have ap = 1e20 instead)
        k = rng.choice(n * nlines, size=n * nlines // 17, replace=False)
        aw[k], ae[k], ap[k] = 0.0, 0.0, 1e-13
    d = rng.normal(size=n * nlines) * 1e3
    phi = rng.normal(size=n * nlines) * 1e3
    return aw, ap, ae, d, phi


@pytest.mark.parametrize("n,nlines,guard", [(402, 160, 0), (37, 1000, 1), (3, 5, 0)])
def test_tdma_lines_matches_the_reference_recurrence(amcfd, n, nlines, guard):
    rng = np.random.default_rng(n * 7919 + nlines)
    aw, ap, ae, d, phi = make_batch(rng, n, nlines, tiny_pivots=bool(guard))
    want = reference_lines(aw, ap, ae, d, phi, n, min(nlines, 200), guard)
    bufs = [DeviceArray(a) for a in (aw, ap, ae, d, phi)]
    print("tdma_lines n=%d nlines=%d guard=%d" % (n, nlines, guard), flush=True)
    amcfd.tdma_lines(*[b.addr for b in bufs], n, nlines, guard)
    got = bufs[4].get()
    m = min(nlines, 200) * n
    assert np.array_equal(got[:m], want[:m])
    for b in bufs:
        b.free()


def test_tdma_lines_rejects_bad_shapes(amcfd):
    with pytest.raises(amcfd.AmcfdError) as e:
        amcfd.tdma_lines(0, 0, 0, 0, 0, 10, 4)
    assert e.value.code == -1
    buf = DeviceArray(np.zeros(8))
    with pytest.raises(amcfd.AmcfdError):
        amcfd.tdma_lines(buf.addr, buf.addr, buf.addr, buf.addr, buf.addr, 2, 4)
    buf.free()
