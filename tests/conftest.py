import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbnn_flip.so on the device)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    if not os.path.exists(oracle_lib.ORACLE_SO):
        oracle_lib.build()
    return oracle_lib


@pytest.fixture(scope="session")
def bnn():
    """The HIP library; the test fails (not skips) when it is not built."""
    import banana_amd

    banana_amd.load()
    return banana_amd
