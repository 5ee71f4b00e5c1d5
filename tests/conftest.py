import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the gfx950 kernels")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import ads_oracle
    ads_oracle.build()
    return ads_oracle
