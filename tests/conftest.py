import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def ora():
    from oracle import uni10_oracle
    return uni10_oracle


@pytest.fixture(scope="session")
def tn():
    """The HIP library, initialised on device 0.  Fails (never skips) when the
    library or the device is missing: GPU tests must run the native path."""
    from trim_amd import tn as _tn
    _tn.load()
    _tn.init(0)
    return _tn
