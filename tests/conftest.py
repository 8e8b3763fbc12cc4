import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgfhip on cuda:0)")


@pytest.fixture(scope="session")
def oracle():
    import oracle.gforacle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def gf():
    import __graft_entry__

    return __graft_entry__.load_package()
