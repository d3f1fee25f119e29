"""bench.py's command line without a GPU: more than one GPU is refused with
a message (this build measures one GPU), and the per-workload defaults of
--steps / --warmup are the documented ones."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_help_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for wl in ("frame", "batched", "diff", "diff_batched", "brickmap", "procgen", "engine",
               "copy"):
        assert wl in r.stdout
    assert "default: 200 for single frames" in " ".join(r.stdout.split())
