"""CPU-side checks of the native library: it loads and exports every symbol of include/or_gpu.h,
and its argument checks fail loudly without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from openradioss_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "or_gpu.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(or_\w+)\(", text, re.M)))


def test_header_and_loader_agree():
    assert header_symbols() == sorted(L.EXPORTS)


def test_library_exports_every_symbol():
    lib = L.load()
    for name in header_symbols():
        assert hasattr(lib, name), name
    assert lib.or_version() >= 100


def test_null_arguments_are_rejected():
    lib = L.load()
    assert lib.or_model_create(None, 0, 10, 1, 0) == L.OR_ERR_ARG
    assert lib.or_device_count(None) == L.OR_ERR_ARG
    assert b"null" in lib.or_last_error()


def test_bad_sizes_are_rejected_before_any_device_call():
    lib = L.load()
    h = C.c_void_p()
    assert lib.or_model_create(C.byref(h), 0, 0, 1, 0) == L.OR_ERR_ARG
    assert lib.or_model_create(C.byref(h), 0, 10, 1, 7) == L.OR_ERR_ARG
    assert lib.or_model_create(C.byref(h), 0, 10, 1 << 29, 0) == L.OR_ERR_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(ImportError):
        L.load()
