"""bench.py's command line without a GPU: one GPU only (a larger --gpus exits
with the reason), and a box without a visible GPU gets a clear error rather
than a CPU measurement."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=300, env=env)


@pytest.mark.parametrize("n", ["2", "8"])
def test_more_than_one_gpu_refused(n):
    r = _bench("--gpus", n)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_no_gpu_is_an_error():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", CUDA_VISIBLE_DEVICES="-1")
    r = _bench("--workload", "sphere32_gjk", "--steps", "2", "--warmup", "0", env=env)
    assert r.returncode == 1
    assert "no GPU visible" in r.stderr and r.stdout == ""
