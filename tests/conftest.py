import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "ccpp-physics_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def kd_lw():
    from ccpprad import synthetic as syn
    return syn.make_kdist(True)


@pytest.fixture(scope="session")
def kd_sw():
    from ccpprad import synthetic as syn
    return syn.make_kdist(False)


@pytest.fixture(scope="session")
def lut_lw():
    from ccpprad import synthetic as syn
    return syn.make_cloud_lut(True)


@pytest.fixture(scope="session")
def lut_sw():
    from ccpprad import synthetic as syn
    return syn.make_cloud_lut(False)
