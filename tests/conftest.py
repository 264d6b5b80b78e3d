import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsvo.so on the device)")


@pytest.fixture(scope="session")
def ctx():
    import sparsevoxeloctree_amd as S

    c = S.Context(0)
    yield c
    c.close()


@pytest.fixture
def no_device_errors(ctx):
    ctx.device_errors()
    yield
    assert ctx.device_errors() == 0, "kernels raised device error bits"
