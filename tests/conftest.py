import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle32():
    from oracle.pyoracle import Oracle
    return Oracle("f32")


@pytest.fixture(scope="session")
def oracle64():
    from oracle.pyoracle import Oracle
    return Oracle("f64")


@pytest.fixture(scope="session")
def device():
    from specklesource_amd import track
    dev = track.Device(0)
    yield dev
    dev.close()
