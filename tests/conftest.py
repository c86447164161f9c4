import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpgpu.so)")


@pytest.fixture(scope="session")
def ctx():
    from proteusgpu_amd import api
    c = api.Context(0)
    yield c
    c.close()
