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
_SRCS = [os.path.join(_HERE, "or_oracle.c"), os.path.join(_HERE, "or_shell.c")]
_HDR = os.path.join(_HERE, "or_oracle.h")
_CR_SRC = os.path.join(_HERE, "or_crmath.cpp")
_CR_DEPS = [os.path.join(_HERE, "..", "openradioss_amd", "csrc", f) for f in ("or_math.h", "or_math_tables.h")]
# "libm": pow/log/exp of the host C library, as the reference engine calls them (pinned bitwise to
# the reference's compiled routines).  "cr": the same restatement under correctly rounded
# pow/log/exp, the elementary functions of the FP64 device path.
MATH_MODES = ("libm", "cr")
# "count": the libm build that also counts, per call site, the calls where the correctly rounded
# value differs from the host library's (or_math_site_counts); a diagnostic of the libm spread.
_LIBS = {"libm": os.path.join(_HERE, "liboracle.so"), "cr": os.path.join(_HERE, "liboracle_cr.so"),
         "count": os.path.join(_HERE, "liboracle_count.so")}
MATH_SITES = ("MQVISCB VOL**THIRD", "SHVIS3 VOL**TWO_THIRD", "M2LAW EPXE**CN", "M2LAW EPXE**(CN-1)",
              "M2LAW rate LOG / TSTAR**MT", "M2CPLR EXP(CN*LOG(PLA)) and LOG")
_LIB = _LIBS["libm"]


def _stale(out, deps):
    return not os.path.exists(out) or any(os.path.getmtime(out) < os.path.getmtime(f) for f in deps)


def build(force: bool = False, math: str = "libm") -> str:
    """Compile the oracle with contraction off (matching the reference release flags)."""
    out = _LIBS[math]
    cflags = ["-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-Wall"]
    if math == "libm":
        if force or _stale(out, _SRCS + [_HDR]):
            subprocess.check_call(["gcc", *cflags, "-shared", *_SRCS, "-o", out, "-lm"])
        return out
    if force or _stale(out, _SRCS + [_HDR, _CR_SRC] + _CR_DEPS):
        tmp = out + ".tmp.o"
        subprocess.check_call(["g++", *cflags, "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                               "-c", _CR_SRC, "-o", tmp])
        flag = "-DOR_CR_MATH" if math == "cr" else "-DOR_COUNT_MATH"
        subprocess.check_call(["gcc", *cflags, flag, "-shared", *_SRCS, tmp, "-o", out, "-lm", "-lstdc++"])
        os.remove(tmp)
    return out


class HexParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "rho0", "bulk", "shear", "ca", "cb", "cn", "epmx", "sigmx", "cc", "epdr", "fisokin",
        "asrate", "m_exp")] + [(n, C.c_int) for n in ("icc", "vp", "israte", "iform")] + [
        (n, C.c_double) for n in ("qa", "qb", "qh", "vns1", "vns2", "dtfac", "dtmx")]


class TimeState(C.Structure):
    _fields_ = [("tt", C.c_double), ("dt1", C.c_double), ("dt2", C.c_double), ("dt12", C.c_double),
                ("dt2old", C.c_double), ("ncycle", C.c_int64), ("mstop", C.c_int)]


_libs = {}


def lib(math: str = "libm"):
    if math not in _libs:
        L = C.CDLL(build(math=math))
        L.or_const.restype = C.c_double
        L.or_const.argtypes = [C.c_int]
        L.or_shell_const.restype = C.c_double
        L.or_shell_const.argtypes = [C.c_int]
        L.or_shell_table.restype = C.c_double
        L.or_shell_table.argtypes = [C.c_int, C.c_int, C.c_int]
        if math == "count":
            L.or_math_site_counts.restype = None
            L.or_math_site_reset.restype = None
        _libs[math] = L
    return _libs[math]


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

    def __init__(self, model, nthreads: int = 1, track_hourglass: bool = False, math: str = "libm"):
        self.m = model
        self.math = math
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
        L = lib(self.math)
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
        lib(self.math).or_energy(C.c_int64(self.m.nn), C.c_int64(self.m.ne), _p(self.eint), _p(self.vol0), _p(self.ms),
                        _p(self.v), _p(out))
        return out


# ---------------------------------------------------------------------------
# General (hex + Q4, 6-DOF) model
# ---------------------------------------------------------------------------
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)


class Q4Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rho", "young", "nu", "shear", "ssp", "ca", "cb", "cn", "epmx", "ymax",
                                          "cc", "epdr", "m_exp", "asrate")] + \
               [(n, C.c_int) for n in ("icc", "vp", "israte", "iform", "npt", "ipla", "ithk", "ishear")] + \
               [(n, C.c_double) for n in ("hm", "hf", "hr", "fsh", "dtfac", "dtmin")]


class Q4State(C.Structure):
    _fields_ = [("ne", C.c_int64), ("offg", _D), ("thk", _D), ("thke", _D), ("forc", _D), ("mom", _D),
                ("gstr", _D), ("hourg", _D), ("eint", _D), ("epsd", _D), ("sig", _D), ("pla", _D),
                ("epsd_ip", _D)]


class ModelC(C.Structure):
    _fields_ = [("nn", C.c_int64)] + [(n, _D) for n in ("x", "v", "a", "d", "vr", "ar", "stifn", "stifr", "ms",
                                                         "inr")] + \
               [("code", _I32), ("vimp", _D), ("adsky", _I64), ("fsky", _D), ("dtmx", C.c_double),
                ("ne_hex", C.c_int64), ("hconn", _I32), ("hiads", _I64)] + \
               [(n, _D) for n in ("hsig", "heint", "hrho", "hpla", "hepsd", "hoffg", "hqvis", "hvol0")] + \
               [("hp", C.POINTER(HexParams)), ("ne_q4", C.c_int64), ("qconn", _I32), ("qiadc", _I64),
                ("q", Q4State), ("qp", C.POINTER(Q4Params)), ("ecrit", _D), ("wfext", C.c_double), ("wpend", _D),
                ("dtel_h", _D), ("dtel_q", _D), ("itypts", C.c_int32), ("nelts", C.c_int64), ("th", _D),
                ("ehou_h", _D), ("ehou_q", _D), ("ehour", C.c_double)]


def q4_params(part) -> Q4Params:
    p = part.params()
    qp = Q4Params()
    for name, _ in Q4Params._fields_:
        setattr(qp, name, p[name])
    return qp


def _pd(a):
    return None if a is None else a.ctypes.data_as(_D)


class ModelOracle:
    """Runs a general openradioss_amd.model.Model on the CPU (hex and Q4 chains)."""

    def __init__(self, model, nthreads: int = 1, math: str = "libm"):
        m = self.m = model
        self.math = math
        nn = m.nn
        self.x = m.x.copy()
        self.v = m.v.copy()
        self.vr = m.vr.copy()
        self.a = np.zeros((nn, 3))
        self.ar = np.zeros((nn, 3))
        self.d = np.zeros((nn, 3))
        self.stifn = np.zeros(nn)
        self.stifr = np.zeros(nn)
        self.ms = np.ascontiguousarray(m.ms, np.float64)
        self.inr = np.ascontiguousarray(m.inr, np.float64)
        self.code = np.ascontiguousarray(m.code, np.int32)
        self.vimp = np.ascontiguousarray(m.vimp, np.float64)
        self.adsky = np.ascontiguousarray(m.adsky, np.int64)
        self.fsky = np.zeros((m.lsky, 8))
        self.nthreads = nthreads
        c = self.c = ModelC()
        c.nn = nn
        for n in ("x", "v", "a", "d", "vr", "ar", "stifn", "stifr", "ms", "vimp", "fsky"):
            setattr(c, n, _pd(getattr(self, n)))
        c.inr = _pd(self.inr) if m.rotational else None
        c.code = self.code.ctypes.data_as(_I32)
        c.adsky = self.adsky.ctypes.data_as(_I64)
        c.dtmx = m.dtmx
        c.ne_hex = m.ne_hex
        if m.hexp is not None:
            h = m.hexp
            st = h.initial_state()
            self.hsig = np.ascontiguousarray(st["sig"])
            self.heint, self.hrho, self.hpla = st["eint"], st["rho"], st["pla"]
            self.hepsd, self.hoffg, self.hqvis = st["epsd"], st["offg"], st["qvis"]
            self.hvol0 = np.ascontiguousarray(h.vol0)
            self.hconn = np.ascontiguousarray(h.conn, np.int32)
            self.hiads = np.ascontiguousarray(h.iads, np.int64)
            hp = HexParams()
            p = h.params(m.dtmx)
            for name, _ in HexParams._fields_:
                setattr(hp, name, p[name])
            self.hp = hp
            c.hconn = self.hconn.ctypes.data_as(_I32)
            c.hiads = self.hiads.ctypes.data_as(_I64)
            for n in ("hsig", "heint", "hrho", "hpla", "hepsd", "hoffg", "hqvis", "hvol0"):
                setattr(c, n, _pd(getattr(self, n)))
            c.hp = C.pointer(self.hp)
        c.ne_q4 = m.ne_q4
        if m.q4 is not None:
            q = m.q4
            st = q.initial_state()
            self.q = {k: np.ascontiguousarray(v) for k, v in st.items()}
            self.qthke = np.ascontiguousarray(q.thke, np.float64)
            self.qconn = np.ascontiguousarray(q.conn, np.int32)
            self.qiadc = np.ascontiguousarray(q.iadc, np.int64)
            self.qp = q4_params(q)
            c.qconn = self.qconn.ctypes.data_as(_I32)
            c.qiadc = self.qiadc.ctypes.data_as(_I64)
            c.q.ne = q.ne
            c.q.thke = _pd(self.qthke)
            for k in ("offg", "thk", "forc", "mom", "gstr", "hourg", "eint", "epsd", "sig", "pla", "epsd_ip"):
                setattr(c.q, k, _pd(self.q[k]))
            c.qp = C.pointer(self.qp)
        self.wpend = np.zeros(3 * nn)
        c.wpend = _pd(self.wpend)
        self.ts = initial_time_state()

    def cycle(self, n: int = 1):
        L = lib(self.math)
        for _ in range(n):
            L.or_cycle_model(C.byref(self.c), C.byref(self.ts), C.c_int(self.nthreads))

    def energy(self):
        out = np.zeros(3)
        lib(self.math).or_energy_model(C.byref(self.c), _pd(out))
        return out

    def cycle_output(self) -> np.ndarray:
        """One cycle that also returns ECRIT's energies of the cycle (internal, translational and
        rotational kinetic at the output point between ACCELE and VELOCITY)."""
        return self.cycle_ecrit()["energies"]

    def cycle_ecrit(self) -> dict:
        """One output cycle: ECRIT's energies, WFEXT and the controlling element (ITYPTS 1 solid, 3
        shell, 0 none; NELTS its 0-based index in its family, -1 none)."""
        out = np.zeros(8)
        self._dtel_h = np.zeros(max(self.m.ne_hex, 1))
        self._dtel_q = np.zeros(max(self.m.ne_q4, 1))
        self.c.ecrit = _pd(out)
        self.c.dtel_h = _pd(self._dtel_h) if self.m.ne_hex else None
        self.c.dtel_q = _pd(self._dtel_q) if self.m.ne_q4 else None
        try:
            lib(self.math).or_cycle_model(C.byref(self.c), C.byref(self.ts), C.c_int(self.nthreads))
        finally:
            self.c.ecrit = None
            self.c.dtel_h = None
            self.c.dtel_q = None
        return dict(energies=out[:3].copy(), enint=out[0], encin=out[1], enrot=out[2], wfext=out[3],
                    xmomt=out[4], ymomt=out[5], zmomt=out[6], itypts=int(self.c.itypts), nelts=int(self.c.nelts))

    def track_ehour(self):
        """Accumulate EHOUR (the hourglass energy of every cycle, group sums in group order) from now on."""
        self._ehou_h = np.zeros(max(self.m.ne_hex, 1))
        self._ehou_q = np.zeros(max(self.m.ne_q4, 1))
        self.c.ehou_h = _pd(self._ehou_h) if self.m.ne_hex else None
        self.c.ehou_q = _pd(self._ehou_q) if self.m.ne_q4 else None

    @property
    def ehour(self) -> float:
        return float(self.c.ehour)

    def cycle_th(self) -> dict:
        """One output cycle (cycle_ecrit) that also returns every node's D, V, A, VR, AR, X at the
        output point ("nodes", [nn][18]) and the time and DT2 the time-history record carries."""
        th = np.zeros((self.m.nn, 18))
        tt = self.ts.tt
        self.c.th = _pd(th)
        try:
            r = self.cycle_ecrit()
        finally:
            self.c.th = None
        r.update(nodes=th, tt=tt, dt2=self.ts.dt2, ehour=self.ehour)
        return r

    @property
    def wfext(self) -> float:
        """External work of the imposed velocities so far (FIXVEL's WFEXT)."""
        return float(self.c.wfext)


def math_site_counts() -> dict:
    """Per call site of the elementary functions: (calls, calls whose correctly rounded value
    differs from the host library's), accumulated by runs with math="count" since the last reset."""
    out = np.zeros(32, np.int64)
    lib("count").or_math_site_counts(out.ctypes.data_as(C.c_void_p))
    return {name: (int(out[2 * k]), int(out[2 * k + 1])) for k, name in enumerate(MATH_SITES)}


def math_site_reset():
    lib("count").or_math_site_reset()
