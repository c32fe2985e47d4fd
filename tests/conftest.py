import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "fem-rs_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def fr():
    import femref
    return femref


@pytest.fixture(scope="session")
def fg():
    import femgpu
    return femgpu


@pytest.fixture(scope="session")
def ctx(fg):
    return fg.default_context()
