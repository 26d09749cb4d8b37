import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libsfx.so")


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    return o


@pytest.fixture(scope="session")
def sfx():
    """The product library; GPU tests fail loudly if it is missing."""
    from field_amd import sfx as s
    assert s.device_count() >= 1, "no HIP device visible"
    return s
