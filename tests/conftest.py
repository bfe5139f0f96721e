import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: longer CPU-side check")


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.pyoracle import Reference, have_reference
    if not have_reference():
        pytest.skip("compiled HONEI reference (oracle/_ref) not built here")
    return Reference()


@pytest.fixture(scope="session")
def hip():
    """The device layer; on a GPU host the library must load (no fallback)."""
    from honei_amd import device, lib
    lib()
    cnt = __import__("ctypes").c_int()
    rc = lib().hsolve_device_count(__import__("ctypes").byref(cnt))
    assert rc == 0 and cnt.value >= 1, "no HIP device visible"
    return device


@pytest.fixture
def rng():
    return np.random.default_rng(20241202)
