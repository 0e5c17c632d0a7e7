"""Host side of the GPU path: a resident model driven through the C ABI.

``GpuHexModel`` uploads a :class:`openradioss_amd.model.HexModel` once, then
advances it cycle by cycle on the GPU with all state resident in HBM, in the
call order of RESOL for this path (resol.F:2721 DT1=DT2 ... 8605 TT+=DT2).
``sforc3_group`` is the group-level drop-in of SFORC3 with the reference's
host array layouts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L
from .model import HexModel


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def hex_params(model: HexModel) -> L.OrHexParams:
    p = model.params()
    hp = L.OrHexParams()
    for name, _ in L.OrHexParams._fields_:
        setattr(hp, name, p[name])
    return hp


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class GpuHexModel:
    """Resident hex model on one GPU (precision 'fp64' or 'split')."""

    def __init__(self, model: HexModel, precision: str = "fp64", device: int = 0, state: dict | None = None):
        lib = L.load()
        self.model = model
        self.prec = {"fp64": L.OR_PREC_FP64, "split": L.OR_PREC_SPLIT}[precision]
        h = C.c_void_p()
        L.check(lib.or_model_create(C.byref(h), device, model.nn, model.ne, self.prec))
        self._h = h
        self._lib = lib
        try:
            self._params = hex_params(model)
            L.check(lib.or_model_set_hex_params(h, C.byref(self._params)))
            x = _f64(model.x)
            v = _f64(model.v)
            imp_code = None
            vimp = None
            if model.imp_node.shape[0]:
                imp_code = np.zeros(model.nn, np.int32)
                vimp = np.zeros((model.nn, 3))
                bit = np.array([4, 2, 1], np.int32)
                for n, dof, val in zip(model.imp_node, model.imp_dof, model.imp_val):
                    imp_code[n] |= bit[dof]
                    vimp[n, dof] = val
            L.check(lib.or_model_upload_nodes(h, _p(x), _p(v), _p(_f64(model.ms)),
                                              _p(np.ascontiguousarray(model.bcs, np.int32)), _p(imp_code),
                                              _p(vimp)))
            st = model.initial_state() if state is None else state
            L.check(lib.or_model_upload_hex(h, _p(np.ascontiguousarray(model.conn, np.int32)), _p(_f64(st["sig"])),
                                            _p(_f64(st["eint"])), _p(_f64(st["rho"])), _p(_f64(model.vol0)),
                                            _p(_f64(st["pla"])), _p(_f64(st["epsd"])), _p(_f64(st["offg"])),
                                            _p(_f64(st["qvis"]))))
            L.check(lib.or_model_upload_skyline(h, _p(np.ascontiguousarray(model.adsky, np.int64)),
                                                _p(np.ascontiguousarray(model.iads, np.int64))))
        except Exception:
            self.close()
            raise

    # -- lifecycle -------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.or_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping --------------------------------------------------------
    def run(self, ncycles: int, sync: bool = True):
        L.check(self._lib.or_model_run(self._h, int(ncycles)))
        if sync:
            self.sync()

    def sync(self):
        L.check(self._lib.or_model_sync(self._h))

    def time_state(self) -> L.OrTimeState:
        ts = L.OrTimeState()
        L.check(self._lib.or_model_get_time(self._h, C.byref(ts)))
        return ts

    def set_time_state(self, ts: L.OrTimeState):
        L.check(self._lib.or_model_set_time(self._h, C.byref(ts)))

    # -- results ---------------------------------------------------------
    def nodes(self):
        nn = self.model.nn
        x = np.empty((nn, 3))
        v = np.empty((nn, 3))
        d = np.empty((nn, 3))
        L.check(self._lib.or_model_download_nodes(self._h, _p(x), _p(v), _p(d), None))
        return x, v, d

    def hex_state(self) -> dict:
        ne = self.model.ne
        out = dict(sig=np.empty((6, ne)), eint=np.empty(ne), rho=np.empty(ne), pla=np.empty(ne),
                   epsd=np.empty(ne), offg=np.empty(ne), qvis=np.empty(ne))
        L.check(self._lib.or_model_download_hex(self._h, _p(out["sig"]), _p(out["eint"]), _p(out["rho"]),
                                                _p(out["pla"]), _p(out["epsd"]), _p(out["offg"]),
                                                _p(out["qvis"])))
        return out

    def energy(self) -> np.ndarray:
        out = np.zeros(2)
        L.check(self._lib.or_model_energy(self._h, _p(out)))
        return out

    def set_timing(self, enable: bool):
        L.check(self._lib.or_model_set_timing(self._h, int(bool(enable))))

    def kernel_times(self, reset: bool = True):
        """(element ms, node ms, cycles) averaged per launch since the last reset."""
        e = C.c_double()
        n = C.c_double()
        c = C.c_int64()
        L.check(self._lib.or_model_kernel_times(self._h, int(reset), C.byref(e), C.byref(n), C.byref(c)))
        return e.value, n.value, c.value

    def footprint(self) -> int:
        b = C.c_int64()
        L.check(self._lib.or_model_footprint(self._h, C.byref(b)))
        return b.value


def sforc3_group(params: L.OrHexParams, x, v, ixs, gbuf: dict, fsky, iads, nft: int, nel: int, dt1: float,
                 dt2t: float, device: int = 0):
    """SFORC3 on one group (reference layouts; see or_sforc3_group in include/or_gpu.h).

    x, v: X(3,NUMNOD) as (numnod,3); ixs: IXS(11,*) as (ncol,11) int32 (1-based ids);
    gbuf: dict of GBUF arrays of the group (sig as (6,nel) = SIG(NEL,6) in Fortran order);
    fsky: FSKY(8,LSKY) as (lsky,8); iads: IADS(8,*) as (ncol,8) 1-based.
    Updates gbuf and fsky in place; returns (dt2t, mstop).
    """
    lib = L.load()
    d = C.c_double(dt2t)
    ms = C.c_int32(0)
    arrs = [gbuf[k] for k in ("sig", "eint", "rho", "vol", "pla", "epsd", "off", "qvis")]
    for a in arrs + [fsky]:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    L.check(lib.or_sforc3_group(device, C.byref(params), x.shape[0], nel, nft, _p(_f64(x)), _p(_f64(v)),
                                _p(np.ascontiguousarray(ixs, np.int32)), *[_p(a) for a in arrs], fsky.shape[0],
                                _p(fsky), _p(np.ascontiguousarray(iads, np.int32)), float(dt1), C.byref(d),
                                C.byref(ms)))
    return d.value, ms.value
