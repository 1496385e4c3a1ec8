import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def orc():
    o = ge.load_oracle()
    if not os.path.exists(o.LIB_PATH):
        o.build()
    return o


@pytest.fixture(scope="session")
def gpj():
    return ge.load_package()


@pytest.fixture(scope="session")
def gpu(gpj):
    """The package, after checking that the HIP library loads and a device exists.
    Fails loudly (no skip) so a GPU run never passes on a missing extension."""
    gpj.lib()
    assert gpj.device_available(), "no HIP device visible to a gpu-marked test"
    return gpj
