import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libflipgpu.so")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.build(ref=True)
    return oracle


@pytest.fixture(scope="session")
def fg():
    import flipsolver2d_amd as fg

    fg.lib()
    return fg


@pytest.fixture(scope="session")
def ctx(fg):
    c = fg.Context(0)
    yield c
    c.close()
