import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "master-rad_amd"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device) to run")


def gpu_available() -> bool:
    try:
        import mrad
        t = mrad.GpuRayTracer(1, 1, 1, True)
        ok = t.is_available()
        t.close()
        return ok
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("gpu-marked test run without a usable HIP device (libmrad reports it unavailable)")
