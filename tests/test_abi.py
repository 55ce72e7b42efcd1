"""The C-ABI library loads and exports every entry point declared in
include/hymd_amd.h (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hymd_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hymd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "hymd_update_field" in names and "hymd_cbf" in names and len(names) >= 25


def test_library_exports_header():
    from hymd_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libhymd_amd.so is not built (python -m hymd_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert _lib.lib().hymd_abi_version() == 1


def test_python_binding_covers_header():
    from hymd_amd import _lib
    assert sorted(_lib.SIGNATURES) == _declared()


def test_error_path_without_device():
    """A bad argument is rejected on the host with a message, before any HIP call."""
    from hymd_amd import _lib
    L = _lib.lib()
    box = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    st = L.hymd_respa_kick_drift(None, None, None, -1, None, 0.0, None, 0.0, 0.1, box, None)
    assert st == -1
    assert b"bad argument" in L.hymd_last_error()
    h = ctypes.c_void_p()
    mesh = (ctypes.c_int32 * 3)(7, 8, 8)
    st = L.hymd_pm_create(ctypes.byref(h), mesh, box, 2, None)
    assert st == -1 and b"even" in L.hymd_last_error()
