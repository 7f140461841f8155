import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bfs-gpu-sycl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI)")


@pytest.fixture(scope="session")
def engine():
    """One engine context for all GPU tests (fails loudly if the build is missing)."""
    import bfsx
    eng = bfsx.Engine(0)
    yield eng
    eng.close()
