import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle_ctypes
    oracle_ctypes.lib()
    return oracle_ctypes


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device; GPU tests fail loudly without it."""
    from dvsrender4deepspace_amd import capi
    capi.load_library()
    if capi.device_count() < 1:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return capi
