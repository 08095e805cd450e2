import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hsmu-spgemm_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP path")


@pytest.fixture(scope="session")
def ctx():
    import hsmu
    c = hsmu.Context(0)
    yield c
    c.close()
