import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def qsr():
    """The product binding; on a GPU run the library and a device must exist."""
    import lcdrsimulator_amd as q
    q.load()
    if q.device_count() < 1:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return q


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O
