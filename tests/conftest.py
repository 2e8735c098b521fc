import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def O():
    """The CPU FP64 oracle (built on first use)."""
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def H():
    import helmholtzbem_amd
    return helmholtzbem_amd
