import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu_lib():
    from openradioss_amd import lib as L
    lib = L.load()
    if L.device_count() < 1:
        pytest.fail("gpu test run without a HIP device")
    return lib
