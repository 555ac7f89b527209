"""The C-ABI library loads and exports every entry point of include/*.h
(no GPU work: runs on the CPU-only host too)."""
import ctypes
import glob
import os
import re

import pytest

from nexd_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(nexd\w*)\s*\(", text))
    return sorted(names)


def test_headers_declare_entry_points():
    names = _declared()
    assert "nexd_create" in names and "nexd_run" in names
    assert set(capi.EXPORTS) <= set(names)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libnexd_gpu.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_without_context_is_a_string():
    lib = capi.load()
    assert isinstance(lib.nexd_last_error(None), bytes)


def test_group_create_rejects_empty_group_without_a_gpu_call():
    lib = capi._load_batch()
    h = ctypes.c_void_p()
    arr = (ctypes.c_void_p * 1)()
    rc = lib.nexd_group_create(arr, 0, ctypes.byref(h))
    assert rc == -1
    lib.nexd_group_last_error.restype = ctypes.c_char_p
    lib.nexd_group_last_error.argtypes = [ctypes.c_void_p]
    assert b"empty" in lib.nexd_group_last_error(None)
