import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "nqcdynamics.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

KEY = 0x4E5143444D493335  # "NQCDMI35"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnqcd_mi355x.so)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def key():
    return KEY
