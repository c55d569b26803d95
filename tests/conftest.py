import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libqll_amd.so")


@pytest.fixture(scope="session")
def gpu():
    import qll_amd
    qll_amd.init()
    return qll_amd
