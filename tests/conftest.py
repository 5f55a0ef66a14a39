import importlib
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
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libf3d_hip.so on the device)")


@pytest.fixture(scope="session")
def f3d():
    """The flame3d-mhd_amd package (directory name has a hyphen, so import by string)."""
    return importlib.import_module("flame3d-mhd_amd")


@pytest.fixture(scope="session")
def cases():
    return importlib.import_module("flame3d-mhd_amd.cases")


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
