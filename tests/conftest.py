import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def bem():
    return importlib.import_module("3d_bem_amd")


@pytest.fixture(scope="session")
def meshes():
    return importlib.import_module("3d_bem_amd.meshes")


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle
