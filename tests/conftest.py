import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libe4hip.so)")
    config.addinivalue_line("markers", "slow: full-size GPU checks")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The device library, initialised once per session.  Fails loudly when the
    HIP library is missing: there is no CPU fallback."""
    import eilmer4_amd as e4
    L = e4.load()
    assert L.e4hip_init(0) == 0, e4.lib.last_error()
    return e4
