"""amcfd_tdma_lines_noalloc against amcfd_tdma_lines through the C ABI:
identical bits on two calls in a row, the second with a smaller batch (the
fixed scratch keeps the first call's contents), and the size limit."""
import numpy as np
import pytest

from hipmem import DeviceArray
from test_gpu_tdma import make_batch

pytestmark = pytest.mark.gpu


def solve(amcfd, fn, arrays, n, nlines, guard):
    bufs = [DeviceArray(a) for a in arrays]
    fn(*[b.addr for b in bufs], n, nlines, guard)
    out = bufs[4].get()
    for b in bufs:
        b.free()
    return out


def test_noalloc_matches_the_allocating_entry_point(amcfd):
    rng = np.random.default_rng(7)
    for n, nlines, guard in ((402, 2000, 0), (41, 300, 1)):   # second call smaller
        print("noalloc vs alloc n=%d nlines=%d guard=%d" % (n, nlines, guard), flush=True)
        arrays = make_batch(rng, n, nlines, tiny_pivots=bool(guard))
        a = solve(amcfd, amcfd.tdma_lines, arrays, n, nlines, guard)
        b = solve(amcfd, amcfd.tdma_lines_noalloc, arrays, n, nlines, guard)
        assert np.array_equal(a, b)
        assert not np.array_equal(a, arrays[4])


def test_noalloc_rejects_batches_beyond_its_scratch(amcfd):
    buf = DeviceArray(np.zeros(8))
    with pytest.raises(amcfd.AmcfdError) as e:
        amcfd.tdma_lines_noalloc(buf.addr, buf.addr, buf.addr, buf.addr, buf.addr, 1025, 1024)
    assert e.value.code == -7
    buf.free()
