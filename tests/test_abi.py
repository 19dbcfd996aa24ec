"""The C ABI library loads and exports every entry point include/ocotillo.h
declares (no GPU needed: nothing here launches work)."""
import ctypes
import os
import re

from conftest import REPO


def declared():
    text = open(os.path.join(REPO, "include", "ocotillo.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oco_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    import ocotillo_amd
    assert declared() == sorted(ocotillo_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    import ocotillo_amd
    L = ctypes.CDLL(ocotillo_amd.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name


def test_strerror_and_version_need_no_device():
    import ocotillo_amd
    L = ocotillo_amd.load_library()
    assert b"tridiag stage 2" in L.oco_strerror(2)
    assert b"tridiag stage 1" in L.oco_strerror(1)
    assert L.oco_version().startswith(b"ocotillo")


def test_no_cpu_fallback_in_binding():
    """The product binding must not import the oracle or compute on the host."""
    src = open(os.path.join(REPO, "ocotillo_amd", "__init__.py")).read()
    assert "oracle" not in src.replace("never computes", "")
