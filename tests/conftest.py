import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libzorro_hip.so")
    config.addinivalue_line("markers", "slow: oracle runs that take tens of seconds on the CPU")


@pytest.fixture(scope="session")
def hip():
    """The HIP library handle; GPU tests fail loudly when it is missing."""
    from zorro_amd import _lib
    return _lib.load()
