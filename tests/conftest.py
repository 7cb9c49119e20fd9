"""Shared fixtures: the package binding (hmec), the oracle and GPU gating."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "h-mecs-2d_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _load_hmec():
    spec = importlib.util.spec_from_file_location("hmec", os.path.join(PKG, "hmec.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hmec"] = mod
    spec.loader.exec_module(mod)
    return mod


hm = _load_hmec()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def hmec():
    return hm


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


def gpu_available():
    try:
        import ctypes
        lib = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        if lib.hipGetDeviceCount(ctypes.byref(n)) != 0:
            return False
        return n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no HIP device")
    return True
