import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import common  # noqa: E402,F401  (sets sys.path for the package and the oracle)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run on the GPU host")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def rbd():
    """The product binding; GPU tests require the HIP library and a device."""
    import rbdgrad
    rbdgrad.lib()
    if rbdgrad.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests must run on the GPU host")
    return rbdgrad
