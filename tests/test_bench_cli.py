"""bench.py's command-line contract on the host: it measures one GPU and says
so for any other count, and its helpers read the committed fixtures."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_roofline_inputs_come_from_committed_files():
    sys.path.insert(0, ROOT)
    import bench
    bpp, rays = bench.fixture_bytes_per_path()
    assert 1000 < bpp < 20000 and 1 < rays < 17  # at most max_depth extension rays plus shadow rays
    tpp = bench.measured_traffic_per_path()
    assert tpp is not None and 0 < tpp < bpp
