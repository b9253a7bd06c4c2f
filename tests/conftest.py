import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU box)")


@pytest.fixture(scope="session")
def gpu():
    """The native library with at least one visible GPU. A GPU test fails (it
    does not skip) when the library or the device is missing."""
    from scx_amd import _lib

    n = _lib.device_count()
    assert n >= 1, "no ROCm GPU visible to libscxpca"
    return _lib
