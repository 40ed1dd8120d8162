import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP device)")
    config.addinivalue_line("markers", "slow: takes more than ~20 s")


@pytest.fixture(scope="session")
def root_dir():
    return ROOT
