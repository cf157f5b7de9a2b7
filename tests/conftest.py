import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libamls_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: minutes-long CPU case")


@pytest.fixture(scope="session")
def blas1():
    """Single-threaded BLAS for the oracle: numpy and scipy ship two OpenBLAS
    builds whose spinning worker threads slow each other down."""
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        yield


@pytest.fixture(scope="session")
def gpu_ctx():
    from amls_amd import HipContext
    ctx = HipContext(0)
    yield ctx
    ctx.close()
