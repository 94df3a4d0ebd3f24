import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
# a failing drop-in call (void, like the reference's) would otherwise end the whole
# test process with exit(-1); with this it records the error for qcuGetLastError()
# and the test that made it fails on its own
os.environ.setdefault("QCU_EXIT_ON_ERROR", "0")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqcu.so on cuda device 0)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
