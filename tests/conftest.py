import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def dp():
    import dplasma_amd
    dplasma_amd.libs()
    return dplasma_amd


@pytest.fixture(scope="session")
def ctx(dp):
    c = dp.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def dev(dp):
    d = dp.Device(0)
    yield d
    d.close()
