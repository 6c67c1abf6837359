import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and librtkhip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import rtkref
    rtkref.lib()  # builds oracle/_build/librtkref.so if needed
    return rtkref


@pytest.fixture(scope="session")
def hip():
    """The HIP package; GPU tests require a device and the built library."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import rtk_amd
    rtk_amd.lib()
    return rtk_amd
