import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpmx kernels)")


@pytest.fixture(scope="session")
def ctx():
    import parmoon_amd as pa
    c = pa.Context(0)
    yield c
    c.close()
