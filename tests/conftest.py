import os
import sys

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
