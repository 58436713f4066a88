import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "triplyperiodic-modified_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def gpu_available():
    # the HIP library itself is the device check: fail loudly if it is
    # missing, skip only when no GPU is present at all
    if not os.path.exists("/dev/kfd"):
        pytest.skip("no GPU visible (/dev/kfd missing)")
    import tpx
    tpx.load()
    return True
