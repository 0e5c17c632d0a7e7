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
    assert lib.or_model_create(None, 0, 10, 1, 0, 0, 0) == L.OR_ERR_ARG
    assert lib.or_device_count(None) == L.OR_ERR_ARG
    assert b"null" in lib.or_last_error()


def test_time_history_calls_reject_null_arguments():
    lib = L.load()
    assert lib.or_model_th_setup(None, 0, None, 0, None, 0, None) == L.OR_ERR_ARG
    assert lib.or_model_run_th(None, None, None, None, None, None) == L.OR_ERR_ARG
    assert lib.or_model_ehour(None, None) == L.OR_ERR_ARG
    assert lib.or_model_set_ehour(None, 0.0) == L.OR_ERR_ARG
    assert b"null" in lib.or_last_error()


def test_bad_sizes_are_rejected_before_any_device_call():
    lib = L.load()
    h = C.c_void_p()
    assert lib.or_model_create(C.byref(h), 0, 0, 1, 0, 0, 0) == L.OR_ERR_ARG        # no nodes
    assert lib.or_model_create(C.byref(h), 0, 10, 1, 0, 0, 7) == L.OR_ERR_ARG       # bad precision
    assert lib.or_model_create(C.byref(h), 0, 10, 1 << 29, 0, 0, 0) == L.OR_ERR_ARG  # rows beyond int32
    assert lib.or_model_create(C.byref(h), 0, 10, 0, 5, 1, 0) == L.OR_ERR_ARG       # NPT=1 not on this path
    assert lib.or_model_create(C.byref(h), 0, 10, 0, 5, 11, 0) == L.OR_ERR_ARG      # NPT > 10


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(ImportError):
        L.load()


def test_group_dropins_reject_bad_input_before_any_device_call():
    import numpy as np
    from openradioss_amd import engine as E
    from openradioss_amd import model as M
    lib = L.load()
    m = M.square_tube(width=4.0, length=2.0)
    qp = E.q4_params(m)
    dt = C.c_double(1.0)
    z = np.zeros(8)
    # nel out of 1..128
    assert lib.or_cforc3_group(0, C.byref(qp), 10, 0, 0, *([z.ctypes.data_as(C.c_void_p)] * 16), 8,
                               z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), 1.0,
                               C.byref(dt)) == L.OR_ERR_ARG
    qp.npt = 1
    ixc = np.array([[0, 1, 2, 3, 4, 1, 1]], np.int32)
    iadc = np.array([[1, 2, 3, 4]], np.int32)
    arr = [np.zeros(64) for _ in range(16)]
    arr[3] = ixc
    args = [a.ctypes.data_as(C.c_void_p) for a in arr]
    assert lib.or_cforc3_group(0, C.byref(qp), 4, 1, 0, *args, 4, np.zeros(32).ctypes.data_as(C.c_void_p),
                               iadc.ctypes.data_as(C.c_void_p), 1.0, C.byref(dt)) == L.OR_ERR_ARG
    assert b"NPT" in lib.or_last_error()


def test_footprint_estimate_needs_no_device():
    """Admission by footprint is computed on the host: no device call, exact row-slot count."""
    from openradioss_amd import model as M
    from openradioss_amd import serve as S
    m = M.square_tube(width=12.0, length=24.0)
    f64, f32 = S.footprint(m, "fp64"), S.footprint(m, "split")
    assert 0 < f32 < f64
    # element state and skyline rows halve in the split form, nodes stay FP64
    nodes = m.nn * (8 * 13 + 4 + 8 * 5)
    assert f64 - f32 > 0.3 * (f64 - nodes)
    b = M.hex_model_to_model(M.cantilever_bar(4, 4, 16))
    assert S.footprint(b, "fp64") > 0
    L = __import__("openradioss_amd.lib", fromlist=["load"])
    import ctypes as C
    out = C.c_int64()
    assert L.load().or_model_footprint_estimate(10, 1, 0, 0, 0, 0, None, None, C.byref(out)) == L.OR_ERR_ARG
