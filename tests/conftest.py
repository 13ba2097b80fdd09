import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) and runs the HIP path")


@pytest.fixture(scope="session")
def mr():
    import mrcpp_amd
    mrcpp_amd.lib()
    return mrcpp_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.Oracle.lib()
    return oracle
