import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C ABI")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("2017multiphase_amd")


@pytest.fixture(scope="session")
def scenes():
    return importlib.import_module("2017multiphase_amd.scenes")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle
