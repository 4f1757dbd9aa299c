import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libxdem.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def lib():
    """libxdem.so with a visible GPU.  A GPU test fails (never skips) when the
    library or the device is missing: there is no CPU fallback."""
    from metakonga_amd import abi
    L = abi.load()
    n = L.xdem_device_count()
    assert n > 0, "no GPU visible to libxdem.so"
    return L
