import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsvm.so on cuda:0)")


@pytest.fixture(scope="session")
def engine():
    """libsvm initialised on device 0 for the whole GPU session."""
    from shadowvm_amd import pipeline as P
    P.init(device=0, tuple_per_bulk=3_000_000)
    yield P
    P.shutdown()
