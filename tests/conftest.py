import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def relerr(a, b):
    """Per-point relative difference; 0 where the values are identical (0 == 0)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(a == b, 0.0, d / np.abs(b))
    return r


def flux_err(a, b):
    """Error measure for V (last axis z): the emergent points V(1), V(nz) per point
    relative; interior points relative to the system's max |V| (V changes sign
    through the midplane of a symmetric column, so it has no per-point scale there)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max(axis=-1, keepdims=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(a == b, 0.0, np.abs(a - b) / scale)
    r[..., 0] = relerr(a[..., 0], b[..., 0])
    r[..., -1] = relerr(a[..., -1], b[..., -1])
    return r


class DeviceArray:
    """A device buffer from the HIP runtime the library itself runs on
    (libamdhip64.so.7 of /opt/rocm, already loaded by the library), so the
    test needs no other GPU runtime: hipMalloc / hipMemcpy / hipFree by
    ctypes.  The test allocates its inputs and outputs; the entry point under
    test allocates nothing."""
    _hip = None

    def __init__(self, nbytes):
        import ctypes as C
        import ocotillo_amd as oco
        oco.load_library()
        if DeviceArray._hip is None:
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            DeviceArray._hip = h
        self.n = int(nbytes)
        p = C.c_void_p()
        assert DeviceArray._hip.hipMalloc(C.byref(p), self.n) == 0
        self.ptr = p.value

    def upload(self, a):
        import ctypes as C
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert DeviceArray._hip.hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.n, 1) == 0
        return self

    def download(self, shape):
        import ctypes as C
        out = np.empty(shape, dtype=np.float64)
        assert out.nbytes == self.n
        assert DeviceArray._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.n, 2) == 0
        return out

    def free(self):
        import ctypes as C
        if self.ptr:
            DeviceArray._hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None
