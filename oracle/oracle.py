"""ctypes front end of the C oracle (oracle/or_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_SRC = os.path.join(_HERE, "or_oracle.c")


def build(force: bool = False) -> str:
    """Compile the oracle with contraction off (matching the reference release flags)."""
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(_SRC):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                               "-fopenmp", "-Wall", _SRC, "-o", _LIB, "-lm"])
    return _LIB


class HexParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "rho0", "bulk", "shear", "ca", "cb", "cn", "epmx", "sigmx", "cc", "epdr", "fisokin",
        "asrate", "m_exp")] + [(n, C.c_int) for n in ("icc", "vp", "israte", "iform")] + [
        (n, C.c_double) for n in ("qa", "qb", "qh", "vns1", "vns2", "dtfac", "dtmx")]


class TimeState(C.Structure):
    _fields_ = [("tt", C.c_double), ("dt1", C.c_double), ("dt2", C.c_double), ("dt12", C.c_double),
                ("dt2old", C.c_double), ("ncycle", C.c_int64), ("mstop", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(_LIB)
        _lib.or_const.restype = C.c_double
        _lib.or_const.argtypes = [C.c_int]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def hex_params(model) -> HexParams:
    p = model.params()
    hp = HexParams()
    for name, _ in HexParams._fields_:
        setattr(hp, name, p[name])
    return hp


def initial_time_state() -> TimeState:
    """TT=DT1=DT2=0, DT2OLD=DTINI/ONEP1 with DTINI=EP06 (starter contrl.F:1125-1130)."""
    ts = TimeState()
    ts.tt = 0.0
    ts.dt1 = 0.0
    ts.dt2 = 0.0
    ts.dt12 = 0.0
    ts.dt2old = 1.0e6 / lib().or_const(1)
    ts.ncycle = 0
    ts.mstop = 0
    return ts


class HexOracle:
    """Runs the hex path on the CPU with the reference's array layouts."""

    def __init__(self, model, nthreads: int = 1, track_hourglass: bool = False):
        self.m = model
        st = model.initial_state()
        self.x = np.ascontiguousarray(model.x, dtype=np.float64).copy()
        self.v = np.ascontiguousarray(model.v, dtype=np.float64).copy()
        self.a = np.zeros_like(self.x)
        self.d = np.zeros_like(self.x)
        self.stifn = np.zeros(model.nn)
        self.sig = np.ascontiguousarray(st["sig"])
        self.eint = st["eint"]
        self.rho = st["rho"]
        self.pla = st["pla"]
        self.epsd = st["epsd"]
        self.offg = st["offg"]
        self.qvis = st["qvis"]
        self.ehou = np.zeros(model.ne) if track_hourglass else None
        self.fsky = np.zeros((model.lsky, 8))
        self.conn = np.ascontiguousarray(model.conn, dtype=np.int32)
        self.iads = np.ascontiguousarray(model.iads, dtype=np.int64)
        self.adsky = np.ascontiguousarray(model.adsky, dtype=np.int64)
        self.vol0 = np.ascontiguousarray(model.vol0)
        self.ms = np.ascontiguousarray(model.ms)
        self.bcs = np.ascontiguousarray(model.bcs, dtype=np.int32)
        self.imp_node = np.ascontiguousarray(model.imp_node, dtype=np.int32)
        self.imp_dof = np.ascontiguousarray(model.imp_dof, dtype=np.int32)
        self.imp_val = np.ascontiguousarray(model.imp_val, dtype=np.float64)
        self.params = hex_params(model)
        self.ts = initial_time_state()
        self.nthreads = nthreads

    def cycle(self, n: int = 1):
        L = lib()
        for _ in range(n):
            L.or_cycle_hex(C.c_int64(self.m.nn), C.c_int64(self.m.ne), C.c_int64(self.m.lsky),
                           _p(self.conn), _p(self.x), _p(self.v), _p(self.a), _p(self.d), _p(self.ms),
                           _p(self.stifn), _p(self.sig), _p(self.eint), _p(self.rho), _p(self.vol0),
                           _p(self.pla), _p(self.epsd), _p(self.offg), _p(self.qvis), _p(self.ehou),
                           _p(self.iads), _p(self.adsky), _p(self.fsky), _p(self.bcs),
                           C.c_int64(self.imp_node.shape[0]), _p(self.imp_node), _p(self.imp_dof),
                           _p(self.imp_val), C.byref(self.params), C.byref(self.ts), C.c_int(self.nthreads))

    def energy(self):
        out = np.zeros(2)
        lib().or_energy(C.c_int64(self.m.nn), C.c_int64(self.m.ne), _p(self.eint), _p(self.vol0), _p(self.ms),
                        _p(self.v), _p(out))
        return out
