import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import oracle as _o
    return _o()


@pytest.fixture(scope="session")
def reference():
    from oracle_lib import reference as _r
    r = _r()
    if r is None:
        pytest.skip("compiled reference (oracle/_ref/libamgref.so) not built")
    return r


@pytest.fixture(scope="session")
def net16():
    from blood_network_amd import make_network
    return make_network(16, 9)


@pytest.fixture(scope="session")
def net30():
    from blood_network_amd import make_network
    return make_network(30, 10)


def pytest_collection_modifyitems(config, items):
    # GPU tests need a device: skip them cleanly where HIP sees none
    try:
        from blood_network_amd import device_available
        have = device_available()
    except ImportError:
        have = False
    if have:
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
