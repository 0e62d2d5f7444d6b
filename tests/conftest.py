import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the native gfx950 path")


def _have_gpu() -> bool:
    try:
        from quantumsimevo_amd import capi
        return capi.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the GPU tests; fails loudly if the library or device is missing."""
    from quantumsimevo_amd import capi
    if capi.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    c = capi.Context(0)
    yield c
    c.close()
