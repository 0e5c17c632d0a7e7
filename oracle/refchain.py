"""Cycle driver around the compiled reference routines (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY.  Runs the hex cycle with the reference's own
SFORC3 leaves (per group of 128 elements, as FORINT does), its ACCELE,
VELOCITY and DEPLA; the skyline sum (ASSPAR4 loop asspar4.F:164-180), the BCS
zeroing (bcs1.F) and the DT2 pipeline (resol.F:6125-6497) are plain
additions / min operations restated here.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libref.so")
_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB, mode=os.RTLD_LAZY)
        _lib.ref_hex_forces.restype = None
        _lib.ref_accele.restype = None
        _lib.ref_veldep.restype = None
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def param_arrays(p: dict):
    prm = np.array([p["rho0"], p["bulk"], p["shear"], p["ca"], p["cb"], p["cn"], p["epmx"], p["sigmx"], p["cc"],
                    p["epdr"], p["fisokin"], p["asrate"], p["m_exp"], p["qa"], p["qb"], p["qh"], p["vns1"],
                    p["vns2"], p["dtfac"], p["dtmx"]], dtype=np.float64)
    iprm = np.array([p["icc"], p["vp"], p["israte"], p["iform"]], dtype=np.int32)
    return prm, iprm


class RefHex:
    """Reference-routine run of a HexModel (same arrays as oracle.HexOracle)."""

    GROUP = 128

    def __init__(self, model):
        self.m = model
        st = model.initial_state()
        self.x = model.x.copy()
        self.v = model.v.copy()
        self.d = np.zeros_like(self.x)
        self.sig = np.ascontiguousarray(st["sig"])
        self.eint, self.rho, self.pla = st["eint"], st["rho"], st["pla"]
        self.epsd, self.offg, self.qvis = st["epsd"], st["offg"], st["qvis"]
        self.vol0 = model.vol0.copy()
        self.fsky = np.zeros((model.lsky, 8))
        self.conn1 = np.ascontiguousarray(model.conn + 1, dtype=np.int32)
        self.iads1 = np.ascontiguousarray(model.iads + 1, dtype=np.int32)
        self.prm, self.iprm = param_arrays(model.params())
        self.tt = 0.0
        self.dt1 = 0.0
        self.dt2 = 0.0
        self.dt12 = 0.0
        self.dt2old = 1.0e6 / (1.0 + 1.0 / 10.0)
        self.ncycle = 0
        self.mstop = 0
        self.ehou = 0.0

    def forces(self, dt1: float) -> float:
        """SFORC3 over all groups; returns DT2T."""
        L = lib()
        dt2t = C.c_double(1.0e6)
        ne = self.m.ne
        for e0 in range(0, ne, self.GROUP):
            e1 = min(ne, e0 + self.GROUP)
            nel = e1 - e0
            sig = np.ascontiguousarray(self.sig[:, e0:e1])
            grp = [np.ascontiguousarray(a[e0:e1]) for a in
                   (self.eint, self.rho, self.vol0, self.pla, self.epsd, self.offg, self.qvis)]
            conn = np.ascontiguousarray(self.conn1[e0:e1])
            iads = np.ascontiguousarray(self.iads1[e0:e1])
            ms = C.c_int(0)
            eh = C.c_double(0.0)
            L.ref_hex_forces(C.c_int(nel), C.c_int(self.m.nn), _p(self.x), _p(self.v), _p(conn), _p(sig),
                             *[_p(a) for a in grp], C.c_int(self.m.lsky), _p(self.fsky), _p(iads), _p(self.prm),
                             _p(self.iprm), C.c_double(dt1), C.byref(dt2t), C.byref(ms), C.byref(eh))
            self.sig[:, e0:e1] = sig
            for dst, src in zip((self.eint, self.rho, self.vol0, self.pla, self.epsd, self.offg, self.qvis), grp):
                dst[e0:e1] = src
            self.mstop |= ms.value
            self.ehou += eh.value
        return dt2t.value

    def cycle(self, n: int = 1):
        m = self.m
        for _ in range(n):
            self.dt1 = self.dt2
            self.dt2 = 1.0e6
            dt2t = self.forces(self.dt1)
            # ASSPAR4: A(1,N) = A(1,N) + FSKY(1,K) for K in node order, from A = 0
            a = np.zeros((m.nn, 3))
            counts = np.diff(m.adsky)
            for j in range(int(counts.max())):
                has = counts > j
                a[has] = a[has] + self.fsky[m.adsky[:-1][has] + j, :3]
            if dt2t < self.dt2:
                self.dt2 = dt2t
            self.dt2 = min(self.dt2, float(np.float32(1.1)) * self.dt2old, m.dtmx)
            self.dt2old = self.dt2
            self.dt12 = 0.5 * (self.dt1 + self.dt2)
            L = lib()
            ms = np.ascontiguousarray(m.ms)
            L.ref_accele(C.c_int(m.nn), _p(a), _p(self.v), _p(ms))
            # BCS1V global frame
            for bit, c in ((4, 0), (2, 1), (1, 2)):
                sel = (m.bcs & bit) != 0
                a[sel, c] = 0.0
                self.v[sel, c] = 0.0
            for n_, dof, val in zip(m.imp_node, m.imp_dof, m.imp_val):
                a[n_, dof] = (val - self.v[n_, dof]) / self.dt12
            L.ref_veldep(C.c_int(m.nn), _p(a), _p(self.v), _p(self.x), _p(self.d), C.c_double(self.dt2),
                         C.c_double(self.dt12))
            self.tt = self.tt + self.dt2
            self.ncycle += 1
