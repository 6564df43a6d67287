import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ref():
    """The compiled reference (oracle/_ref); skips when it was not built."""
    import oracle
    r = oracle.ref()
    if r is None:
        pytest.skip("oracle/_ref/libphc_ref.so not built")
    return r


@pytest.fixture(scope="session")
def port():
    import oracle
    return oracle.port()


@pytest.fixture(scope="session")
def host():
    """The product's Fortran host library (C-callable layer)."""
    from polyhomocont_amd import host as h
    return h.library()


@pytest.fixture(scope="session")
def ctx():
    from polyhomocont_amd import device
    c = device.Context(0)
    yield c
    c.close()
