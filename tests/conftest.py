import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libtamc_gpu.so")


@pytest.fixture(scope="session")
def gpu():
    """Initialise the device once; a GPU test fails loudly if the native library or device is missing."""
    from tamc_amd import _lib
    _lib.check(_lib.lib().tamc_init())
    return _lib.lib()
