import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (REPO, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the library's kernels")


@pytest.fixture(scope="session")
def ctx():
    """One device context for all GPU tests of the session."""
    from serenity_amd import xc
    c = xc.Context(0)
    yield c
    c.close()
