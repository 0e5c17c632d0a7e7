"""Loader of the HIP C-ABI library (include/or_gpu.h) and its ctypes signatures.

The product path is the native library built from openradioss_amd/csrc for
gfx950.  If it is missing the import of ``load()`` fails loudly: there is no
CPU fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libor_gpu.so")

OR_OK = 0
OR_ERR_ARG = 1
OR_ERR_HIP = 2
OR_ERR_NODEV = 3
OR_ERR_STATE = 4
OR_ERR_MSTOP = 5
OR_PREC_FP64 = 0
OR_MSTOP_JACOBIAN = 1
OR_MSTOP_TIME_STEP = 2
OR_PREC_SPLIT = 1

# Every symbol declared in include/or_gpu.h.
EXPORTS = (
    "or_version", "or_last_error", "or_device_count", "or_device_memory", "or_model_create", "or_model_destroy",
    "or_model_set_hex_params", "or_model_set_q4_params", "or_model_upload_nodes", "or_model_upload_displacement", "or_model_upload_hex",
    "or_model_upload_q4", "or_model_upload_skyline", "or_model_set_time", "or_model_get_time", "or_model_run",
    "or_model_sync", "or_model_run_output", "or_model_download_nodes", "or_model_download_hex", "or_model_download_q4",
    "or_model_energy", "or_model_kernel_times", "or_model_set_timing", "or_model_footprint", "or_model_footprint_estimate", "or_sforc3_group",
    "or_cforc3_group", "or_nodal_update", "or_batch_create", "or_batch_destroy", "or_batch_add", "or_batch_remove", "or_batch_run",
    "or_batch_sync", "or_batch_active", "or_model_set_variant", "or_model_autotune", "or_batch_autotune",
    "or_math_eval", "or_sforc3_group_static", "or_model_assembly_stats", "or_model_run_ecrit", "or_model_wfext",
    "or_model_th_setup", "or_model_run_th", "or_model_ehour", "or_model_set_ehour",
)


class OrHexParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "rho0", "bulk", "shear", "ca", "cb", "cn", "epmx", "sigmx", "cc", "epdr", "fisokin", "asrate",
        "m_exp")] + [(n, C.c_int32) for n in ("icc", "vp", "israte", "iform")] + [
        (n, C.c_double) for n in ("qa", "qb", "qh", "vns1", "vns2", "dtfac", "dtmx")]


class OrQ4Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "rho", "young", "nu", "shear", "ssp", "ca", "cb", "cn", "epmx", "ymax", "cc", "epdr", "m_exp",
        "asrate")] + [(n, C.c_int32) for n in ("icc", "vp", "israte", "iform", "npt", "ipla", "ithk", "ishear")] + [
        (n, C.c_double) for n in ("hm", "hf", "hr", "fsh", "dtfac", "dtmin")]


class OrTimeState(C.Structure):
    _fields_ = [("tt", C.c_double), ("dt1", C.c_double), ("dt2", C.c_double), ("dt12", C.c_double),
                ("dt2old", C.c_double), ("ncycle", C.c_int64), ("mstop", C.c_int32)]


class OrEcrit(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("enint", "encin", "enrot", "wfext")] + [
        ("itypts", C.c_int32), ("pad", C.c_int32), ("nelts", C.c_int64)]


class OrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"or_gpu error {code}: {msg}")
        self.code = code


_lib = None

_P = C.c_void_p
_SIGS = {
    "or_version": ([], C.c_int),
    "or_last_error": ([], C.c_char_p),
    "or_device_count": ([_P], C.c_int),
    "or_model_create": ([_P, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_int], C.c_int),
    "or_model_destroy": ([_P], C.c_int),
    "or_model_set_hex_params": ([_P, _P], C.c_int),
    "or_model_set_q4_params": ([_P, _P], C.c_int),
    "or_model_upload_nodes": ([_P] * 8, C.c_int),
    "or_model_upload_displacement": ([_P, _P], C.c_int),
    "or_model_upload_hex": ([_P] * 10, C.c_int),
    "or_model_upload_q4": ([_P] * 14, C.c_int),
    "or_model_upload_skyline": ([_P] * 4, C.c_int),
    "or_model_set_time": ([_P, _P], C.c_int),
    "or_model_get_time": ([_P, _P], C.c_int),
    "or_model_run": ([_P, C.c_int64], C.c_int),
    "or_model_sync": ([_P], C.c_int),
    "or_model_download_nodes": ([_P] * 5, C.c_int),
    "or_model_download_hex": ([_P] * 8, C.c_int),
    "or_model_download_q4": ([_P] * 12, C.c_int),
    "or_model_energy": ([_P, _P], C.c_int),
    "or_model_kernel_times": ([_P, C.c_int, _P, _P, _P, _P], C.c_int),
    "or_model_set_timing": ([_P, C.c_int], C.c_int),
    "or_model_footprint": ([_P, _P], C.c_int),
    "or_model_assembly_stats": ([_P, _P], C.c_int),
    "or_model_footprint_estimate": ([C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_int, C.c_int, _P, _P, _P],
                                    C.c_int),
    "or_sforc3_group": ([C.c_int, _P, C.c_int64, C.c_int32, C.c_int64] + [_P] * 11 + [C.c_int64, _P, _P,
                                                                                      C.c_double, _P, _P], C.c_int),
    "or_sforc3_group_static": ([C.c_int, _P, C.c_int64, C.c_int32, C.c_int64] + [_P] * 11 + [C.c_int64, _P, _P,
                                                                                             C.c_double, _P, _P],
                               C.c_int),
    "or_cforc3_group": ([C.c_int, _P, C.c_int64, C.c_int32, C.c_int64] + [_P] * 16 + [C.c_int64, _P, _P,
                                                                                      C.c_double, _P], C.c_int),
    "or_math_eval": ([C.c_int, C.c_int, C.c_int64, _P, _P, _P], C.c_int),
    "or_model_run_output": ([_P, _P], C.c_int),
    "or_model_run_ecrit": ([_P, _P], C.c_int),
    "or_model_wfext": ([_P, _P], C.c_int),
    "or_model_th_setup": ([_P, C.c_int64, _P, C.c_int64, _P, C.c_int64, _P], C.c_int),
    "or_model_run_th": ([_P] * 6, C.c_int),
    "or_model_ehour": ([_P, _P], C.c_int),
    "or_model_set_ehour": ([_P, C.c_double], C.c_int),
    "or_nodal_update": ([C.c_int, C.c_int64, C.c_int64] + [_P] * 12 + [C.c_double, C.c_double, _P], C.c_int),
    "or_device_memory": ([C.c_int, _P, _P], C.c_int),
    "or_batch_create": ([_P, C.c_int], C.c_int),
    "or_batch_destroy": ([_P], C.c_int),
    "or_batch_add": ([_P, _P, C.c_double], C.c_int),
    "or_batch_remove": ([_P, _P], C.c_int),
    "or_batch_run": ([_P, C.c_int64], C.c_int),
    "or_batch_sync": ([_P], C.c_int),
    "or_batch_active": ([_P, _P], C.c_int),
    "or_model_set_variant": ([_P, C.c_int, C.c_int], C.c_int),
    "or_model_autotune": ([_P, C.c_int32, _P], C.c_int),
    "or_batch_autotune": ([_P, C.c_int32, _P], C.c_int),
}


def load():
    """Load libor_gpu.so (built by __graft_entry__.build()); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"native library {LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int):
    if rc != OR_OK:
        raise OrError(rc, load().or_last_error().decode())


def device_memory(device: int = 0) -> tuple[int, int]:
    """(free, total) bytes of device memory."""
    f, t = C.c_int64(), C.c_int64()
    check(load().or_device_memory(device, C.byref(f), C.byref(t)))
    return f.value, t.value


def device_count() -> int:
    n = C.c_int(0)
    rc = load().or_device_count(C.byref(n))
    return n.value if rc == OR_OK else 0
