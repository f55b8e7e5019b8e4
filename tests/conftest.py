import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def gpu():
    from parallelpopgen_amd import capi
    lib = capi.load()
    n = capi.C.c_int32()
    rc = lib.gf_device_count(capi.C.byref(n))
    if rc != 0 or n.value < 1:
        raise RuntimeError("no HIP device visible: %s" % lib.gf_last_error().decode())
    return capi
