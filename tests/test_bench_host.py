"""bench.py's host-side logic (no GPU): the headline rule and the one-GPU
guard."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_guided_headline_needs_the_bound_and_a_speedup():
    assert bench.guided_is_headline(8.05e-7, 1.14e6, 6.8e5)
    assert not bench.guided_is_headline(1.2e-6, 1.14e6, 6.8e5)   # misses the error bound
    assert not bench.guided_is_headline(5.9e-7, 1.2e6, 1.7e6)    # slower than the plain step
    assert not bench.guided_is_headline(None, 1.14e6, 6.8e5)     # no reference run
    assert bench.guided_is_headline(1e-6, 2.0, 1.0)              # the bound is inclusive


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
