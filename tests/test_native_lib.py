"""The C-ABI library loads and exports every function include/vjx.h declares."""

import ctypes
import os
import re

from vajax_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "vjx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vjx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "vjx_lu_refactor" in names and "vjx_transient_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.EXPORTED)


def test_no_device_is_reported_cleanly():
    # never fails; 0 without a GPU
    assert _lib.lib.vjx_device_count() >= 0
    assert _lib.lib.vjx_version() >= 100


def test_pattern_validation_errors():
    import numpy as np

    h = ctypes.c_void_p()
    indptr = np.array([0, 2, 3], np.int32)
    indices = np.array([1, 0, 1], np.int32)  # row 0 not sorted
    rc = _lib.lib.vjx_pattern_create(2, 3, _lib.ptr(indptr), _lib.ptr(indices), ctypes.byref(h))
    assert rc == -1
    assert "sorted" in _lib.last_error()
    indices = np.array([0, 1, 5], np.int32)
    rc = _lib.lib.vjx_pattern_create(2, 3, _lib.ptr(indptr), _lib.ptr(indices), ctypes.byref(h))
    assert rc == -1 and "range" in _lib.last_error()
