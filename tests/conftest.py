import os
import subprocess
import sys

import pytest

try:  # before anything loads the HIP library, so that it shares torch's HIP runtime
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)
# a persistent batch launch whose lane-group barrier waits this long gives up with an error
# (the library's default is 60 s; the test scenes need milliseconds)
os.environ.setdefault("MT_BATCH_STALL_MS", "20000")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "build", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "oracle.cpp")
    if not os.path.exists(lib) or os.path.getmtime(lib) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session", autouse=True)
def oracle_built():
    _ensure_oracle()
    yield


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is visible; fails loudly if the HIP library is missing."""
    import ctypes
    from mitrace_amd import _lib
    L = _lib.load()  # raises ImportError when the library is not built: no fallback
    n = ctypes.c_int(0)
    if L.mt_device_count(ctypes.byref(n)) != 0 or n.value == 0:
        pytest.skip("no HIP device visible")
    return L
