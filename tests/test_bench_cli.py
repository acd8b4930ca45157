"""bench.py's command-line contract on the host (no GPU needed): more than one
GPU is refused with a message, and the child-process mode used for the
one-at-a-time comparison is recognised before anything starts HIP."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_queue_count_left_alone_for_the_solo_child():
    """The bench raises GPU_MAX_HW_QUEUES to 32 for its side-by-side leg, but
    not in the child that runs the instances one at a time."""
    code = ("import os, sys; sys.argv = ['bench.py'] + sys.argv[1:]; "
            "os.environ['GPU_MAX_HW_QUEUES'] = '4'; "
            "sys.path.insert(0, %r); import bench; print(os.environ['GPU_MAX_HW_QUEUES'])" % ROOT)
    parent = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, cwd=ROOT)
    child = subprocess.run([sys.executable, "-c", code, "--batch-solo-child"], capture_output=True, text=True,
                           timeout=120, cwd=ROOT)
    assert parent.stdout.strip() == "32", parent.stderr
    assert child.stdout.strip() == "4", child.stderr
