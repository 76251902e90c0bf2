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


def test_image_error_measures_rmse_flips_and_ray_counts():
    import numpy as np
    sys.path.insert(0, ROOT)
    import bench
    rng = np.random.default_rng(0)
    ref = rng.uniform(0.1, 1.0, (20, 30, 3))
    same = bench.image_error(ref.astype(np.float32), ref, 1)
    assert same["rel_rmse"] < 1e-7 and same["flip_fraction"] == 0 and same["pixels"] == 600
    moved = ref.copy()
    moved.reshape(-1, 3)[::10] *= 1.01  # 60 of 600 pixels moved by 1%
    e = bench.image_error(moved, ref, 1, (100, 40), {"extend_rays": 101, "shadow_rays": 40})
    assert abs(e["flip_fraction"] - 0.1) < 1e-9 and 1e-3 < e["rel_rmse"] < 1e-2
    assert e["rays"]["rel_diff"] == round(1 / 141, 7)
    strided = bench.image_error(moved, ref, 10)  # exactly the moved pixels
    assert strided["pixels"] == 60 and strided["flip_fraction"] == 1.0


def test_library_override_needs_explicit_permission():
    env = dict(os.environ, ZIGTRACE_AMD_LIB="/nonexistent/libother.so")
    env.pop("ZIGTRACE_AMD_ALLOW_OTHER_LIB", None)
    r = subprocess.run([sys.executable, "-c", "import zigtrace_amd"], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "ZIGTRACE_AMD_ALLOW_OTHER_LIB" in r.stderr
