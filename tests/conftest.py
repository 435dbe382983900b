import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libwgm.so on cuda:0)")


@pytest.fixture(scope="session")
def rt():
    import gpu_workgroup_managment_amd as wgm

    r = wgm.Runtime(0)
    yield r
    r.close()


@pytest.fixture(scope="session")
def uk():
    """libwgm_userk.so: caller-written kernels plus plain device-buffer helpers
    (hipMalloc / copies) used by the tests of caller kernels."""
    path = os.path.join(ROOT, "tests", "kernels", "libwgm_userk.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for f in ("userk_rounds", "userk_mega"):
        getattr(lib, f).restype = ctypes.c_void_p
    lib.userk_alloc.argtypes = [ctypes.c_longlong, ctypes.POINTER(ctypes.c_void_p)]
    lib.userk_free.argtypes = [ctypes.c_void_p]
    lib.userk_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong]
    lib.userk_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong]
    return lib


class Dev:
    """A device copy of a numpy array (through the userk helpers)."""

    def __init__(self, uk, host):
        self.uk, self.host = uk, np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert uk.userk_alloc(self.host.nbytes, ctypes.byref(self.ptr)) == 0
        assert uk.userk_h2d(self.ptr, self.host.ctypes.data, self.host.nbytes) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.uk.userk_d2h(out.ctypes.data, self.ptr, out.nbytes) == 0
        return out

    def free(self):
        self.uk.userk_free(self.ptr)


@pytest.fixture(scope="session")
def dev(uk):
    """dev(numpy array) -> Dev."""
    return lambda host: Dev(uk, host)
