"""Host-side parts of bench.py: the one-GPU rule and the CPU baseline legs on tiny samples."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_baseline_legs_on_a_tiny_sample():
    import bench

    flat = bench.hmmwv_params(18)
    rate, rate1, secs = bench.cpu_baseline(18, flat, 16, 200, 2)
    assert rate > 0 and rate1 > 0 and secs > 0
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref18.so")):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ref = bench.cpu_reference(8, 200, 2, n_single=2)
    assert ref is not None and ref[0] > 0 and ref[1] > 0
