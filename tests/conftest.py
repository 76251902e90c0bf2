import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; calls the HIP library")


SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def scene_text(name: str) -> str:
    with open(os.path.join(SCENES, f"{name}.txt")) as f:
        return f.read()


@pytest.fixture(scope="session")
def gpu_ctx():
    import zigtrace_amd as zt
    ctx = zt.Context(0)
    yield ctx
    ctx.close()
