import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import build as ob
    from oracle import oracle as o
    if not os.path.exists(o._LIB_PATH):
        ob.build()
    return o
