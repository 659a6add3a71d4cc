"""bench.py's command line on a machine without a GPU: more than one GPU is
refused with a message (one GPU is what the library measures), and without a
visible GPU the default run stops with a nonzero status instead of timing
anything."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_no_gpu_no_timing():
    import torch

    if torch.cuda.is_available():
        return
    r = _run("--steps", "1", "--warmup", "0", "--no-cpu")
    assert r.returncode != 0
    assert r.stdout.strip() == ""
