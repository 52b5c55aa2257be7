import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    from oracle_bind import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def device():
    import libgail_amd as g
    d = g.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="session")
def ref():
    """The reference's own sources compiled into oracle/_ref (skipped where not built)."""
    from oracle_bind import reference
    r = reference()
    if r is None:
        pytest.skip("compiled reference not built (oracle/_ref)")
    return r
