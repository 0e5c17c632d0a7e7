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
        for f in ("ref_q4_forces", "ref_accele6", "ref_veldep6", "ref_shell_consts"):
            getattr(_lib, f).restype = None
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
            if self.dt2 <= 0.0:
                self.mstop |= 2  # resol.F:8575, 9599
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


def q4_param_arrays(p: dict):
    prm = np.zeros(24)
    vals = [p["rho"], p["young"], p["nu"], p["shear"], p["a11"], p["a12"], p["ssp"], p["ca"], p["cb"], p["cn"],
            p["epmx"], p["ymax"], p["cc"], p["epdr"], p["asrate"], p["m_exp"], p["hm"], p["hf"], p["hr"], p["fsh"],
            p["dtfac"], p["dtmin"]]
    prm[:len(vals)] = vals
    iprm = np.array([p["icc"], p["vp"], p["israte"], p["iform"], p["npt"], p["ipla"], p["ithk"], p["ishear"]],
                    dtype=np.int32)
    return prm, iprm


def shell_consts():
    out = np.zeros(4)
    lib().ref_shell_consts(_p(out))
    return out


class RefModel:
    """Reference-routine run of a general Model (hex groups through SFORC3 leaves,
    Q4 groups through the CFORC3 chain, nodal update through ACCELE / VELOCITY / DEPLA
    with rotational DOF).  Same arrays as oracle.ModelOracle."""

    GROUP = 128

    def __init__(self, model):
        m = self.m = model
        nn = m.nn
        self.x, self.v, self.vr = m.x.copy(), m.v.copy(), m.vr.copy()
        self.d = np.zeros((nn, 3))
        self.fsky = np.zeros((m.lsky, 8))
        if m.hexp is not None:
            st = m.hexp.initial_state()
            self.hsig = np.ascontiguousarray(st["sig"])
            self.heint, self.hrho, self.hpla = st["eint"], st["rho"], st["pla"]
            self.hepsd, self.hoffg, self.hqvis = st["epsd"], st["offg"], st["qvis"]
            self.hvol0 = m.hexp.vol0.copy()
            self.hconn1 = np.ascontiguousarray(m.hexp.conn + 1, dtype=np.int32)
            self.hiads1 = np.ascontiguousarray(m.hexp.iads + 1, dtype=np.int32)
            self.hprm, self.hiprm = param_arrays(m.hexp.params(m.dtmx))
        if m.q4 is not None:
            self.q = {k: np.ascontiguousarray(v) for k, v in m.q4.initial_state().items()}
            self.qthke = m.q4.thke.astype(np.float64).copy()
            self.qconn1 = np.ascontiguousarray(m.q4.conn + 1, dtype=np.int32)
            self.qiadc1 = np.ascontiguousarray(m.q4.iadc + 1, dtype=np.int32)
            self.qprm, self.qiprm = q4_param_arrays(m.q4.params())
        self.tt = self.dt1 = self.dt2 = self.dt12 = 0.0
        self.dt2old = 1.0e6 / (1.0 + 1.0 / 10.0)
        self.ncycle = 0
        self.mstop = 0
        self.ehour = 0.0  # EHOUR: every group's EHOURT, group after group (solids, then shells)

    def _hex_forces(self, dt1, dt2t, ne=None):
        """SFORC3 over the groups of the first ne hexes (all by default)."""
        L = lib()
        ne = self.m.ne_hex if ne is None else ne
        for e0 in range(0, ne, self.GROUP):
            e1 = min(ne, e0 + self.GROUP)
            sig = np.ascontiguousarray(self.hsig[:, e0:e1])
            arrs = (self.heint, self.hrho, self.hvol0, self.hpla, self.hepsd, self.hoffg, self.hqvis)
            grp = [np.ascontiguousarray(a[e0:e1]) for a in arrs]
            ms, eh = C.c_int(0), C.c_double(0.0)
            L.ref_hex_forces(C.c_int(e1 - e0), C.c_int(self.m.nn), _p(self.x), _p(self.v),
                             _p(np.ascontiguousarray(self.hconn1[e0:e1])), _p(sig), *[_p(a) for a in grp],
                             C.c_int(self.m.lsky), _p(self.fsky), _p(np.ascontiguousarray(self.hiads1[e0:e1])),
                             _p(self.hprm), _p(self.hiprm), C.c_double(dt1), C.byref(dt2t), C.byref(ms),
                             C.byref(eh))
            self.hsig[:, e0:e1] = sig
            for dst, src in zip(arrs, grp):
                dst[e0:e1] = src
            self.mstop |= ms.value
            self.ehour = self.ehour + eh.value  # EHOUR = EHOUR + EHOURT (shvis3.F:464)

    def _q4_forces(self, dt1, dt2t, ne=None):
        """CFORC3 over the groups of the first ne shells (all by default)."""
        L = lib()
        ne = self.m.ne_q4 if ne is None else ne
        q = self.q
        eh = C.c_double(0.0)
        for e0 in range(0, ne, self.GROUP):
            e1 = min(ne, e0 + self.GROUP)
            g = {}
            for k, a in q.items():
                g[k] = np.ascontiguousarray(a[..., e0:e1])
            # Fortran layouts: FOR(NEL,5) etc. are component-major = (5, nel) C arrays;
            # SIGL(NEL,5,NPT) = (npt, 5, nel) C array
            L.ref_q4_forces(C.c_int(e1 - e0), C.c_int(self.m.nn), _p(self.x), _p(self.v), _p(self.vr),
                            _p(np.ascontiguousarray(self.qconn1[e0:e1])), _p(g["offg"]), _p(g["thk"]),
                            _p(np.ascontiguousarray(self.qthke[e0:e1])), _p(g["forc"]), _p(g["mom"]),
                            _p(g["gstr"]), _p(g["hourg"]), _p(g["eint"]), _p(g["epsd"]), _p(g["sig"]),
                            _p(g["pla"]), _p(g["epsd_ip"]), C.c_int(self.m.lsky), _p(self.fsky),
                            _p(np.ascontiguousarray(self.qiadc1[e0:e1])), _p(self.qprm), _p(self.qiprm),
                            C.c_double(dt1), C.byref(dt2t), C.byref(eh))
            for k in q:
                q[k][..., e0:e1] = g[k]
            self.ehour = self.ehour + eh.value  # chvis3.F:397

    def cycle(self, n: int = 1):
        m = self.m
        L = lib()
        for _ in range(n):
            self.dt1 = self.dt2
            self.dt2 = 1.0e6
            dt2t = C.c_double(1.0e6)
            if m.hexp is not None:
                self._hex_forces(self.dt1, dt2t)
            if m.q4 is not None:
                self._q4_forces(self.dt1, dt2t)
            # ASSPAR4 (asspar4.F:164-180): rows of each node in skyline order from zero
            a = np.zeros((m.nn, 3))
            ar = np.zeros((m.nn, 3))
            counts = np.diff(m.adsky)
            for j in range(int(counts.max())):
                has = counts > j
                rows = self.fsky[m.adsky[:-1][has] + j]
                a[has] = a[has] + rows[:, 0:3]
                ar[has] = ar[has] + rows[:, 3:6]
            if dt2t.value < self.dt2:
                self.dt2 = dt2t.value
            self.dt2 = min(self.dt2, float(np.float32(1.1)) * self.dt2old, m.dtmx)
            self.dt2old = self.dt2
            self.dt12 = 0.5 * (self.dt1 + self.dt2)
            if self.dt2 <= 0.0:
                self.mstop |= 2  # resol.F:8575, 9599
            ms = np.ascontiguousarray(m.ms)
            inr = np.ascontiguousarray(m.inr)
            L.ref_accele6(C.c_int(m.nn), _p(a), _p(ar), _p(self.v), _p(ms), _p(inr))
            code = m.code
            for bit, c in ((4, 0), (2, 1), (1, 2)):
                sel = (code & bit) != 0
                a[sel, c] = 0.0
                self.v[sel, c] = 0.0
            for bit, c in ((32, 0), (16, 1), (8, 2)):
                sel = (code & bit) != 0
                ar[sel, c] = 0.0
                self.vr[sel, c] = 0.0
            for bit, c in ((256, 0), (128, 1), (64, 2)):
                sel = (code & bit) != 0
                a[sel, c] = (m.vimp[sel, c] - self.v[sel, c]) / self.dt12
            L.ref_veldep6(C.c_int(m.nn), _p(a), _p(ar), _p(self.v), _p(self.vr), _p(self.x), _p(self.d),
                          C.c_double(self.dt2), C.c_double(self.dt12))
            self.tt = self.tt + self.dt2
            self.ncycle += 1


def bilan_internal_energy(model, heint=None, hrho=None, hoffg=None, v=None, q_eint=None, q_offg=None, x=None,
                          vr=None) -> tuple[float, float]:
    """Internal energy of a state as the reference's SBILAN and CBILAN add it into PARTSAV(1) of
    the solid part and of the shell part, group of 128 after group (oracle/ref_bilan.F).
    Returns (solids, shells)."""
    L = lib()
    for f in ("ref_sbilan", "ref_cbilan"):
        getattr(L, f).restype = None
    npsav = 29
    eh = eq = 0.0
    if model.hexp is not None:
        ps = np.zeros(npsav)
        conn = model.hexp.conn
        vol0 = np.ascontiguousarray(model.hexp.vol0, np.float64)
        for e0 in range(0, model.ne_hex, 128):
            e1 = min(model.ne_hex, e0 + 128)
            vel = np.ascontiguousarray(v[conn[e0:e1]], np.float64)     # (nel, 8, 3) = VEL(3,8,NEL)
            L.ref_sbilan(C.c_int(e1 - e0), _p(vel), _p(np.ascontiguousarray(heint[e0:e1])),
                         _p(np.ascontiguousarray(hrho[e0:e1])), _p(np.ascontiguousarray(vol0[e0:e1])),
                         _p(np.ascontiguousarray(vol0[e0:e1])), _p(np.ascontiguousarray(hoffg[e0:e1])), _p(ps))
        eh = float(ps[0])
    if model.q4 is not None:
        ps = np.zeros(npsav)
        conn1 = np.ascontiguousarray(model.q4.conn + 1, np.int32)
        rho = float(model.q4.params()["rho"])
        for e0 in range(0, model.ne_q4, 128):
            e1 = min(model.ne_q4, e0 + 128)
            ge = np.ascontiguousarray(q_eint[:, e0:e1])                 # EINT(NEL,2)
            gvol = np.ones(e1 - e0)
            L.ref_cbilan(C.c_int(e1 - e0), C.c_int(model.nn), _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(v)),
                         _p(np.ascontiguousarray(vr)), _p(np.ascontiguousarray(conn1[e0:e1])), _p(ge), C.c_double(rho),
                         _p(gvol), _p(np.ascontiguousarray(q_offg[e0:e1])), _p(ps))
        eq = float(ps[0])
    return eh, eq


# ---------------------------------------------------------------------------------------------
# The reference Starter's element initialisation (oracle/_ref/libref_starter.so, oracle/ref_starter.F)
LIB_STARTER = os.path.join(HERE, "_ref", "libref_starter.so")
_lib_st = None


def starter_available() -> bool:
    return os.path.exists(LIB_STARTER)


def lib_starter():
    global _lib_st
    if _lib_st is None:
        # RTLD_LOCAL: the Starter's modules and commons share names with the Engine's
        _lib_st = C.CDLL(LIB_STARTER, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        for f in ("ref_st_mass_", "ref_st_volume_"):
            getattr(_lib_st, f).restype = None
    return _lib_st


def starter_masses(x, hex_conn=None, hex_ids=None, rho_hex=0.0, q4_conn=None, q4_ids=None, thke=None,
                   rho_q4=0.0, ihbe=1) -> dict:
    """MS and IN of a model as the reference Starter computes them: solid volumes (SDERI3) and corner
    masses (SMASS3), shell areas (CEVECI), masses and inertias (CINMAS), then the nodal sums over
    the elements in user-id order (SPMD_MSIN).  conn: 0-based node ids; ids: element user ids
    (default index + 1)."""
    L = lib_starter()
    x = np.ascontiguousarray(x, np.float64)
    nn = x.shape[0]
    nels = 0 if hex_conn is None else int(hex_conn.shape[0])
    nelc = 0 if q4_conn is None else int(q4_conn.shape[0])
    ixs = np.zeros((max(nels, 1), 11), np.int32)
    if nels:
        ixs[:, 0] = 1
        ixs[:, 1:9] = np.asarray(hex_conn) + 1
        ixs[:, 9] = 1
        ixs[:, 10] = np.arange(1, nels + 1) if hex_ids is None else np.asarray(hex_ids)
    ixc = np.zeros((max(nelc, 1), 7), np.int32)
    th = np.zeros(max(nelc, 1))
    if nelc:
        ixc[:, 0] = 1
        ixc[:, 1:5] = np.asarray(q4_conn) + 1
        ixc[:, 5] = 1
        ixc[:, 6] = np.arange(1, nelc + 1) if q4_ids is None else np.asarray(q4_ids)
        th[:nelc] = thke
    ms, inr = np.zeros(nn), np.zeros(nn)
    vols, mss = np.zeros(max(nels, 1)), np.zeros((max(nels, 1), 8))
    area, msc, inc = np.zeros(max(nelc, 1)), np.zeros(max(nelc, 1)), np.zeros(max(nelc, 1))
    i32 = lambda v: C.byref(C.c_int32(v))  # noqa: E731
    f64 = lambda v: C.byref(C.c_double(v))  # noqa: E731
    L.ref_st_mass_(i32(nn), _p(x), i32(nels), _p(ixs), f64(rho_hex), i32(nelc), _p(ixc), _p(th), f64(rho_q4),
                   i32(ihbe), _p(ms), _p(inr), _p(vols), _p(mss), _p(area), _p(msc), _p(inc))
    return dict(ms=ms, inr=inr, vol=vols[:nels], mss=mss[:nels], area=area[:nelc], msc=msc[:nelc], inc=inc[:nelc])
