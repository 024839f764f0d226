import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP path)")


@pytest.fixture(scope="session")
def orc():
    """the CPU oracle (checker)"""
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_lib():
    """the product library; on a GPU test it must load and see a device"""
    import nla3d_amd
    nla3d_amd.lib()
    return nla3d_amd
