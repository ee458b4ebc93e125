"""bench.py's command-line contract, checked without a GPU: it measures one GPU
and says so for any other count, and its workloads are named by what they are."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_explanation():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""   # no bench line for an unsupported count


def test_workload_names():
    r = _run("--help")
    assert r.returncode == 0
    for name in ("batched32x4096", "eeg32x30504", "wide128x16384", "uneven512", "served"):
        assert name in r.stdout
    assert _run("--workload", "c3").returncode != 0   # numbered names are not accepted
