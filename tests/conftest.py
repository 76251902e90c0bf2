import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; calls the HIP library")


def pytest_report_header(config):
    """Names the native library the tests load (path and hash)."""
    try:
        from zigtrace_amd import _native
        info = _native.loaded_library()
        return f"libzigtrace_amd: {info['path']} sha256:{info['sha256_16']}" + ("" if info["default"] else " (override)")
    except Exception as e:  # the CPU-only tests still run; the GPU ones fail loudly on import
        return f"libzigtrace_amd: not loaded ({e})"


SCENES = os.path.join(ROOT, "examples")


def scene_text(name: str) -> str:
    with open(os.path.join(SCENES, f"{name}.txt")) as f:
        return f.read()


@pytest.fixture(scope="session")
def gpu_ctx():
    import zigtrace_amd as zt
    ctx = zt.Context(0)
    yield ctx
    ctx.close()
