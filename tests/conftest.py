import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def lee():
    return importlib.import_module("3d_mpi_test_amd")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle
