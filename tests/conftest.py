"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on a
CPU-only host.  The oracle (oracle/) is the checker: only tests, smoke() and
bench.py's CPU baseline load it."""
import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from tay_amd import api  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
MODELS_LIB = os.path.join(ROOT, "oracle", "_build", "libmodels.so")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: large CPU oracle runs")


@pytest.fixture(scope="session")
def oracle():
    lib = api.TayLib(ORACLE_LIB)
    lib.tay_threads_start(min(os.cpu_count() or 1, 16), 0)
    return lib


@pytest.fixture(scope="session")
def models():
    return ctypes.CDLL(MODELS_LIB)


@pytest.fixture(scope="session")
def tay():
    """The GPU library.  Loading fails loudly when it is not built."""
    return api.TayLib(api.LIBRARY)


@pytest.fixture(scope="session")
def gpu(tay):
    if not tay.tay_gpu_available():
        pytest.fail("no GPU available to the GPU library: " + tay.tay_gpu_last_error().decode())
    return tay
