"""Host side of the GPU path: a resident model driven through the C ABI.

``GpuModel`` uploads an :class:`openradioss_amd.model.Model` (hexes and/or Q4
shells) once, then advances it cycle by cycle on the GPU with all state
resident in HBM, in the call order of RESOL for this path (resol.F:2721
DT1=DT2 ... 8605 TT+=DT2).  ``GpuHexModel`` keeps the hex-only interface.
``sforc3_group`` / ``cforc3_group`` are the group-level drop-ins of SFORC3 and
CFORC3 with the reference's host array layouts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L
from .model import HexModel, Model, hex_model_to_model


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def hex_params(model) -> L.OrHexParams:
    """or_hex_params of a HexModel or of the hex part of a Model."""
    p = model.params() if isinstance(model, HexModel) else model.hexp.params(model.dtmx)
    hp = L.OrHexParams()
    for name, _ in L.OrHexParams._fields_:
        setattr(hp, name, p[name])
    return hp


def q4_params(model: Model) -> L.OrQ4Params:
    p = model.q4.params()
    qp = L.OrQ4Params()
    for name, _ in L.OrQ4Params._fields_:
        setattr(qp, name, p[name])
    return qp


PRECISIONS = {"fp64": L.OR_PREC_FP64, "split": L.OR_PREC_SPLIT}


class GpuModel:
    """Resident hex / Q4 model on one GPU (precision 'fp64' or 'split')."""

    def __init__(self, model: Model, precision: str = "fp64", device: int = 0, hex_state: dict | None = None,
                 q4_state: dict | None = None, nodes: tuple | None = None):
        """nodes: (x, v, vr, d) to start from instead of the model's x, v, vr and zero d (a restart)."""
        lib = L.load()
        self.model = model
        self.prec = PRECISIONS[precision]
        npt = model.q4.prop.npt if model.q4 is not None else 0
        h = C.c_void_p()
        L.check(lib.or_model_create(C.byref(h), device, model.nn, model.ne_hex, model.ne_q4, npt, self.prec))
        self._h = h
        self._lib = lib
        try:
            if model.hexp is not None:
                self._hp = hex_params(model)
                L.check(lib.or_model_set_hex_params(h, C.byref(self._hp)))
            if model.q4 is not None:
                self._qp = q4_params(model)
                L.check(lib.or_model_set_q4_params(h, C.byref(self._qp)))
            x0, v0, vr0 = (model.x, model.v, model.vr) if nodes is None else nodes[:3]
            L.check(lib.or_model_upload_nodes(h, _p(_f64(x0)), _p(_f64(v0)), _p(_f64(vr0)),
                                              _p(_f64(model.ms)), _p(_f64(model.inr)),
                                              _p(np.ascontiguousarray(model.code, np.int32)),
                                              _p(_f64(model.vimp))))
            if model.hexp is not None:
                st = model.hexp.initial_state() if hex_state is None else hex_state
                L.check(lib.or_model_upload_hex(h, _p(np.ascontiguousarray(model.hexp.conn, np.int32)),
                                                _p(_f64(st["sig"])), _p(_f64(st["eint"])), _p(_f64(st["rho"])),
                                                _p(_f64(model.hexp.vol0)), _p(_f64(st["pla"])), _p(_f64(st["epsd"])),
                                                _p(_f64(st["offg"])), _p(_f64(st["qvis"]))))
            if model.q4 is not None:
                st = model.q4.initial_state() if q4_state is None else q4_state
                L.check(lib.or_model_upload_q4(h, _p(np.ascontiguousarray(model.q4.conn, np.int32)),
                                               _p(_f64(model.q4.thke)), *[_p(_f64(st[k])) for k in (
                                                   "offg", "thk", "forc", "mom", "gstr", "hourg", "eint", "epsd",
                                                   "sig", "pla", "epsd_ip")]))
            iads = None if model.hexp is None else np.ascontiguousarray(model.hexp.iads, np.int64)
            iadc = None if model.q4 is None else np.ascontiguousarray(model.q4.iadc, np.int64)
            L.check(lib.or_model_upload_skyline(h, _p(np.ascontiguousarray(model.adsky, np.int64)), _p(iads),
                                                _p(iadc)))
            if nodes is not None:
                L.check(lib.or_model_upload_displacement(h, _p(_f64(nodes[3]))))
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
    def nodes6(self):
        """(x, v, vr, d) as (nn,3) arrays."""
        nn = self.model.nn
        x, v, vr, d = (np.empty((nn, 3)) for _ in range(4))
        L.check(self._lib.or_model_download_nodes(self._h, _p(x), _p(v), _p(vr), _p(d)))
        return x, v, vr, d

    def nodes(self):
        x, v, _, d = self.nodes6()
        return x, v, d

    def hex_state(self) -> dict:
        ne = self.model.ne_hex
        out = dict(sig=np.empty((6, ne)), eint=np.empty(ne), rho=np.empty(ne), pla=np.empty(ne),
                   epsd=np.empty(ne), offg=np.empty(ne), qvis=np.empty(ne))
        L.check(self._lib.or_model_download_hex(self._h, _p(out["sig"]), _p(out["eint"]), _p(out["rho"]),
                                                _p(out["pla"]), _p(out["epsd"]), _p(out["offg"]),
                                                _p(out["qvis"])))
        return out

    def q4_state(self) -> dict:
        ne, npt = self.model.ne_q4, self.model.q4.prop.npt
        out = dict(offg=np.empty(ne), thk=np.empty(ne), forc=np.empty((5, ne)), mom=np.empty((3, ne)),
                   gstr=np.empty((8, ne)), hourg=np.empty((5, ne)), eint=np.empty((2, ne)), epsd=np.empty(ne),
                   sig=np.empty((npt, 5, ne)), pla=np.empty((npt, ne)), epsd_ip=np.empty((npt, ne)))
        L.check(self._lib.or_model_download_q4(self._h, *[_p(out[k]) for k in (
            "offg", "thk", "forc", "mom", "gstr", "hourg", "eint", "epsd", "sig", "pla", "epsd_ip")]))
        return out

    def energy(self) -> np.ndarray:
        """[internal, translational kinetic, rotational kinetic]."""
        out = np.zeros(3)
        L.check(self._lib.or_model_energy(self._h, _p(out)))
        return out

    def run_output(self) -> np.ndarray:
        """One cycle with ECRIT's energy output: [internal, translational kinetic, rotational
        kinetic] at the output point of that cycle (or_model_run_output)."""
        out = np.zeros(3)
        L.check(self._lib.or_model_run_output(self._h, _p(out)))
        return out

    def run_ecrit(self) -> dict:
        """One output cycle: what ECRIT prints for it (or_model_run_ecrit): the energies, the external
        work of the imposed velocities and the controlling element (itypts 1 solid, 3 shell, 0 none;
        nelts its 0-based index in its family, -1 none)."""
        e = L.OrEcrit()
        L.check(self._lib.or_model_run_ecrit(self._h, C.byref(e)))
        return dict(enint=e.enint, encin=e.encin, enrot=e.enrot, wfext=e.wfext + self.wfext_offset,
                    itypts=e.itypts, nelts=e.nelts)

    # external work carried over from a checkpoint (the device sums it from the node upload on;
    # the DT2*DW half of the checkpointed cycle's imposed work is not carried)
    wfext_offset = 0.0

    def wfext(self) -> float:
        """External work of the imposed velocities (since the node upload, plus wfext_offset)."""
        w = C.c_double()
        L.check(self._lib.or_model_wfext(self._h, C.byref(w)))
        return w.value + self.wfext_offset

    # -- time history ------------------------------------------------------
    def th_setup(self, nodes=(), hexes=(), q4s=()):
        """Entities of the time history (0-based node, hex and shell indices); starts the EHOUR
        accumulation (or_model_th_setup)."""
        self._th = [np.ascontiguousarray(np.asarray(a, np.int64).astype(np.int32)) for a in (nodes, hexes, q4s)]
        n, h, q = self._th
        L.check(self._lib.or_model_th_setup(self._h, C.c_int64(len(n)), n.ctypes.data_as(C.c_void_p),
                                            C.c_int64(len(h)), h.ctypes.data_as(C.c_void_p),
                                            C.c_int64(len(q)), q.ctypes.data_as(C.c_void_p)))
        if getattr(self, "_ehour_restart", None) is not None:  # EHOUR of a checkpoint
            self.set_ehour(self._ehour_restart)
            self._ehour_restart = None

    def run_th(self) -> dict:
        """One output cycle with the time-history values (or_model_run_th): the run_ecrit record plus
        xmomt, ymomt, zmomt, ehour, the time and the new DT2 of the record, "nodes" [n][18] (D, V, A,
        VR, AR, X), "hex" [n][13] and "q4" [n][12 + NPT] element values."""
        if getattr(self, "_th", None) is None:
            raise RuntimeError("th_setup first")
        n, h, q = self._th
        npt = self.model.q4.prop.npt if self.model.q4 is not None else 0
        no, ho, qo = np.zeros((len(n), 18)), np.zeros((len(h), 13)), np.zeros((len(q), 12 + npt))
        t0 = self.time_state()
        e = L.OrEcrit()
        g = np.zeros(4)
        L.check(self._lib.or_model_run_th(self._h, C.byref(e), g.ctypes.data_as(C.c_void_p),
                                          no.ctypes.data_as(C.c_void_p), ho.ctypes.data_as(C.c_void_p),
                                          qo.ctypes.data_as(C.c_void_p)))
        t1 = self.time_state()
        return dict(enint=e.enint, encin=e.encin, enrot=e.enrot, wfext=e.wfext + self.wfext_offset,
                    itypts=e.itypts, nelts=e.nelts, xmomt=g[0], ymomt=g[1], zmomt=g[2], ehour=g[3],
                    tt=t0.tt, ncycle=t0.ncycle, dt2=t1.dt2, nodes=no, hex=ho, q4=qo)

    def ehour(self) -> float:
        """EHOUR so far (after th_setup)."""
        v = C.c_double()
        L.check(self._lib.or_model_ehour(self._h, C.byref(v)))
        return v.value

    def set_ehour(self, value: float):
        L.check(self._lib.or_model_set_ehour(self._h, C.c_double(value)))

    def run_time_history(self, header, dthis: float, tstop: float, path=None, this: float = 0.0,
                         max_cycles: int | None = None):
        """Run until TT reaches tstop (or max_cycles cycles), writing a time-history record at every
        cycle whose TT has reached the next output time THIS (hist2.F:261-276: then THIS =
        min(TSTOP, max(TT, THIS + DTHIS))).  header: an openradioss_amd.th.ThHeader whose groups
        list this model's entities; returns the file's bytes (written to path when given) and the
        next output time."""
        from . import th as TH
        self.th_setup(*TH.group_indices(header))
        out = [TH.header_bytes(header)]
        xmass = float(np.cumsum(np.asarray(self.model.ms, np.float64))[-1])  # node after node, as ECRIT adds it
        done = 0
        while True:
            t = self.time_state()
            if t.tt >= tstop or t.mstop or (max_cycles is not None and done >= max_cycles):
                break
            if t.tt >= this:
                r = self.run_th()
                out.append(TH.time_bytes(r["tt"], TH.global_values(r, xmass), TH.group_values(header, self.model, r)))
                this = TH.next_time(this, dthis, r["tt"], tstop)
                done += 1
                continue
            # cycles that cannot reach THIS: DT2 grows by at most 10 % a cycle (resol.F:6125-6129)
            k, adv, dt = 0, 0.0, t.dt2
            while True:
                dt *= 1.1
                if t.tt + adv + dt >= this or (max_cycles is not None and done + k + 1 >= max_cycles):
                    break
                adv += dt
                k += 1
            self.run(max(k, 1))
            done += max(k, 1)
        data = b"".join(out)
        if path is not None:
            with open(path, "wb") as fh:
                fh.write(data)
        return data, this

    def run_listing(self, ncycles: int, every: int, listing=None):
        """Run ncycles cycles printing the Engine listing's cycle line (ecrit.F FORMAT 1100) at every
        cycle whose NCYCLE is a multiple of `every`; returns the openradioss_amd.listing.Listing (a
        new one unless one is passed to continue)."""
        from . import listing as LS
        m = self.model
        if listing is None:
            xmass = float(np.cumsum(np.asarray(m.ms, np.float64))[-1])  # node after node, as ECRIT adds it
            listing = LS.Listing(xmass, None if m.hexp is None else m.hexp.ids, None if m.q4 is None else m.q4.ids)
        if every <= 0:
            raise ValueError("every must be > 0")
        left = ncycles
        while left > 0:
            t0 = self.time_state()
            if t0.ncycle % every == 0:
                e = self.run_ecrit()
                t1 = self.time_state()
                listing.add(t0.ncycle, t0.tt, t1.dt2, e["enint"], e["encin"], e["enrot"], e["wfext"], e["itypts"],
                            e["nelts"])
                left -= 1
            else:
                k = min(left, every - t0.ncycle % every)
                self.run(k)
                left -= k
        return listing

    # -- restart ---------------------------------------------------------
    _TS_FIELDS = ("tt", "dt1", "dt2", "dt12", "dt2old", "ncycle", "mstop")

    def checkpoint(self) -> dict:
        """Everything a restart needs, as host arrays: X, V, VR, D, the element state and the
        cycle scalars (the content of the Engine's restart file for this path)."""
        x, v, vr, d = self.nodes6()
        ts = self.time_state()
        out = {"x": x, "v": v, "vr": vr, "d": d,
               "time": np.array([float(getattr(ts, f)) for f in self._TS_FIELDS]),
               "wfext": np.array([self.wfext()])}
        if getattr(self, "_th", None) is not None:
            out["ehour"] = np.array([self.ehour()])
        if self.model.hexp is not None:
            out.update({"hex_" + k: a for k, a in self.hex_state().items()})
        if self.model.q4 is not None:
            out.update({"q4_" + k: a for k, a in self.q4_state().items()})
        return out

    @classmethod
    def restore(cls, model: Model, ck: dict, precision: str = "fp64", device: int = 0) -> "GpuModel":
        """A resident model that continues, bit for bit, from a checkpoint() of the same mesh."""
        hs = {k[4:]: a for k, a in ck.items() if k.startswith("hex_")} or None
        qs = {k[3:]: a for k, a in ck.items() if k.startswith("q4_")} or None
        g = cls(model, precision=precision, device=device, hex_state=hs, q4_state=qs,
                nodes=(ck["x"], ck["v"], ck["vr"], ck["d"]))
        ts = L.OrTimeState()
        for f, val in zip(cls._TS_FIELDS, ck["time"]):
            setattr(ts, f, int(val) if f in ("ncycle", "mstop") else float(val))
        g.set_time_state(ts)
        if "wfext" in ck:
            g.wfext_offset = float(ck["wfext"][0])
        if "ehour" in ck:
            g._ehour_restart = float(ck["ehour"][0])
        return g

    def save(self, path: str):
        """Write checkpoint() to an .npz file."""
        np.savez(path, **self.checkpoint())

    @classmethod
    def load(cls, model: Model, path: str, precision: str = "fp64", device: int = 0) -> "GpuModel":
        with np.load(path, allow_pickle=False) as f:
            ck = {k: f[k] for k in f.files}
        return cls.restore(model, ck, precision=precision, device=device)

    def set_variant(self, family: str, variant: int):
        """Launch variant (0..3) of the 'hex' or 'q4' element kernel; results do not depend on it."""
        L.check(self._lib.or_model_set_variant(self._h, {"hex": 0, "q4": 1}[family], int(variant)))

    def autotune(self, cycles_per_variant: int = 5) -> tuple[int, int]:
        """Time every launch variant on real cycles (the model advances) and keep the fastest."""
        out = (C.c_int32 * 2)()
        L.check(self._lib.or_model_autotune(self._h, int(cycles_per_variant), out))
        return out[0], out[1]

    def set_timing(self, enable: bool):
        L.check(self._lib.or_model_set_timing(self._h, int(bool(enable))))

    def kernel_times(self, reset: bool = True):
        """(hex ms, q4 ms, node ms, cycles) averaged per launch since the last reset."""
        h, q, n, c = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        L.check(self._lib.or_model_kernel_times(self._h, int(reset), C.byref(h), C.byref(q), C.byref(n),
                                                C.byref(c)))
        return h.value, q.value, n.value, c.value

    def assembly_stats(self) -> dict:
        """The assembly plan: prefixes (a node's leading rows summed in its first tile's LDS) and
        rest rows (stored, summed by the node kernel) of each family."""
        out = np.zeros(4, np.int64)
        L.check(self._lib.or_model_assembly_stats(self._h, _p(out)))
        return dict(hex_prefixes=int(out[0]), hex_rest_rows=int(out[1]), q4_prefixes=int(out[2]),
                    q4_rest_rows=int(out[3]))

    def footprint(self) -> int:
        b = C.c_int64()
        L.check(self._lib.or_model_footprint(self._h, C.byref(b)))
        return b.value


class GpuHexModel(GpuModel):
    """Resident hex-only model (a HexModel, with its BCS / imposed-velocity lists)."""

    def __init__(self, model: HexModel, precision: str = "fp64", device: int = 0, state: dict | None = None):
        self.hex_model = model
        super().__init__(hex_model_to_model(model), precision=precision, device=device, hex_state=state)

    def energy(self) -> np.ndarray:
        """[internal, kinetic] as the hex path reports them."""
        return super().energy()[:2]


class GpuBatch:
    """Many resident models advanced together (or_batch_*): each member runs its own cycles
    with its own time step until its time reaches its t_end."""

    def __init__(self, device: int = 0):
        lib = L.load()
        h = C.c_void_p()
        L.check(lib.or_batch_create(C.byref(h), device))
        self._h = h
        self._lib = lib
        self.members = []

    def add(self, gm: GpuModel, t_end: float):
        L.check(self._lib.or_batch_add(self._h, gm._h, float(t_end)))
        self.members.append(gm)

    def remove(self, gm: GpuModel):
        L.check(self._lib.or_batch_remove(self._h, gm._h))
        self.members.remove(gm)

    def run(self, ncycles: int, sync: bool = True):
        L.check(self._lib.or_batch_run(self._h, int(ncycles)))
        if sync:
            self.sync()

    def sync(self):
        L.check(self._lib.or_batch_sync(self._h))

    def autotune(self, cycles_per_variant: int = 5) -> tuple:
        """Variants per (precision, family) chosen by timing batch cycles: (fp64 hex, fp64 q4,
        split hex, split q4)."""
        out = (C.c_int32 * 4)()
        L.check(self._lib.or_batch_autotune(self._h, int(cycles_per_variant), out))
        return tuple(out)

    def active(self) -> int:
        n = C.c_int32()
        L.check(self._lib.or_batch_active(self._h, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.or_batch_destroy(self._h)
            self._h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sforc3_group(params: L.OrHexParams, x, v, ixs, gbuf: dict, fsky, iads, nft: int, nel: int, dt1: float,
                 dt2t: float, device: int = 0, static: bool = False):
    """SFORC3 on one group (reference layouts; see or_sforc3_group in include/or_gpu.h).

    x, v: X(3,NUMNOD) as (numnod,3); ixs: IXS(11,*) as (ncol,11) int32 (1-based ids);
    gbuf: dict of GBUF arrays of the group (sig as (6,nel) = SIG(NEL,6) in Fortran order);
    fsky: FSKY(8,LSKY) as (lsky,8); iads: IADS(8,*) as (ncol,8) 1-based.
    Updates gbuf and fsky in place; returns (dt2t, mstop).  static=True calls the
    allocation-free variant (or_sforc3_group_static).
    """
    lib = L.load()
    d = C.c_double(dt2t)
    ms = C.c_int32(0)
    arrs = [gbuf[k] for k in ("sig", "eint", "rho", "vol", "pla", "epsd", "off", "qvis")]
    for a in arrs + [fsky]:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    fn = lib.or_sforc3_group_static if static else lib.or_sforc3_group
    L.check(fn(device, C.byref(params), x.shape[0], nel, nft, _p(_f64(x)), _p(_f64(v)),
                                _p(np.ascontiguousarray(ixs, np.int32)), *[_p(a) for a in arrs], fsky.shape[0],
                                _p(fsky), _p(np.ascontiguousarray(iads, np.int32)), float(dt1), C.byref(d),
                                C.byref(ms)))
    return d.value, ms.value


def cforc3_group(params: L.OrQ4Params, x, v, vr, ixc, gbuf: dict, lbuf: dict, fsky, iadc, nft: int, nel: int,
                 dt1: float, dt2t: float, device: int = 0) -> float:
    """CFORC3 on one group (reference layouts; see or_cforc3_group in include/or_gpu.h).

    x, v, vr: (numnod,3); ixc: IXC(7,*) as (ncol,7) int32 1-based; gbuf: GBUF arrays of the group
    (offg, thk, thke, forc (5,nel), mom (3,nel), gstr (8,nel), hourg (5,nel), eint (2,nel), epsd);
    lbuf: sig (npt,5,nel), pla (npt,nel), epsd (npt,nel); fsky (lsky,8); iadc (ncol,4) 1-based.
    Updates gbuf, lbuf and fsky in place; returns dt2t.
    """
    lib = L.load()
    d = C.c_double(dt2t)
    g = [gbuf[k] for k in ("offg", "thk", "thke", "forc", "mom", "gstr", "hourg", "eint", "epsd")]
    lb = [lbuf[k] for k in ("sig", "pla", "epsd")]
    for a in g + lb + [fsky]:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    L.check(lib.or_cforc3_group(device, C.byref(params), x.shape[0], nel, nft, _p(_f64(x)), _p(_f64(v)),
                                _p(_f64(vr)), _p(np.ascontiguousarray(ixc, np.int32)), *[_p(a) for a in g],
                                *[_p(a) for a in lb], fsky.shape[0], _p(fsky),
                                _p(np.ascontiguousarray(iadc, np.int32)), float(dt1), C.byref(d)))
    return d.value


def math_eval(fn: str, x, y=None, device: int = 0) -> np.ndarray:
    """Device values of the FP64 elementary functions of the element kernels (or_math_eval):
    fn 'log' / 'exp' of x, or 'pow' of (x, y)."""
    code = {"log": 0, "exp": 1, "pow": 2}[fn]
    x = _f64(x).ravel()
    yy = None if y is None else _f64(np.broadcast_to(y, x.shape)).ravel()
    out = np.empty_like(x)
    L.check(L.load().or_math_eval(device, code, x.size, _p(x), _p(yy), _p(out)))
    return out


def nodal_update(adsky, fsky, ms, x, v, d, ts: L.OrTimeState, dt2t: float, dtmx: float = 1.0e20, inr=None,
                 vr=None, code=None, vimp=None, stifn=None, stifr=None, device: int = 0):
    """ASSPAR4 -> DEPLA with the reference host arrays (or_nodal_update): adsky 1-based CSR
    (numnod+1), fsky (lsky,8) = FSKY(8,LSKY); x, v, d, vr (numnod,3) updated in place; ts advanced."""
    for a in (x, v, d) + tuple(a for a in (vr, stifn, stifr) if a is not None):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    adsky = np.ascontiguousarray(adsky, np.int64)
    L.check(L.load().or_nodal_update(device, x.shape[0], fsky.shape[0], _p(adsky), _p(_f64(fsky)), _p(_f64(ms)),
                                     _p(None if inr is None else _f64(inr)),
                                     _p(None if code is None else np.ascontiguousarray(code, np.int32)),
                                     _p(None if vimp is None else _f64(vimp)), _p(x), _p(v), _p(vr), _p(d),
                                     _p(stifn), _p(stifr), float(dt2t), float(dtmx), C.byref(ts)))
