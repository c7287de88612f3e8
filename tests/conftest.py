import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd")
for p in (HERE, os.path.join(PKG, "python")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="module")
def tmpg():
    """the ctypes binding with libtmpg.so loaded (raises when it is missing)"""
    import tmpg as t
    t.load_library()
    return t
