"""bench.py's command-line contract on the host (no GPU needed): more than one
GPU is refused with a message, and the child-process mode that runs the
side-by-side leg is recognised before anything starts HIP."""
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


def test_queue_count_raised_only_in_the_side_by_side_child():
    """The side-by-side leg runs in a child process with GPU_MAX_HW_QUEUES
    raised to 32; the parent, which runs the single-stream legs, keeps the
    default 4."""
    code = ("import os, sys; sys.argv = ['bench.py'] + sys.argv[1:]; "
            "os.environ['GPU_MAX_HW_QUEUES'] = '4'; "
            "sys.path.insert(0, %r); import bench; print(os.environ['GPU_MAX_HW_QUEUES'])" % ROOT)
    parent = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, cwd=ROOT)
    child = subprocess.run([sys.executable, "-c", code, "--batch-child"], capture_output=True, text=True,
                           timeout=120, cwd=ROOT)
    assert parent.stdout.strip() == "4", parent.stderr
    assert child.stdout.strip() == "32", child.stderr
