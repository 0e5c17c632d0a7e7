"""Serving loop for crash solves on one GPU: continuous batching of resident models.

Requests arrive over time (each a model, a precision and an end time).  The loop
admits a waiting request when its HBM footprint (or_model_footprint_estimate)
fits in the budget left by the requests on the GPU, uploads it and adds it to
the running batch (or_batch_add), so that its cycles run in the same launches as
every other admitted request.  It advances the batch `cycles_per_step` cycles at
a time; a request whose time reached its end leaves the batch, its results are
downloaded and its memory is freed, which makes room for the next admissions.

Latency of a request = completion time - arrival time (queueing, upload,
compute and download included).  The batch runs on its own stream.  Each pass
of the loop queues the next step of the running requests first and admits new
ones after it, so that a new request's upload (on its own stream) overlaps the
cycles of the others; the admitted request joins the launches of the next step.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import engine as E
from . import lib as L
from . import model as M


@dataclasses.dataclass
class Request:
    rid: int
    model: M.Model
    precision: str          # "fp64" or "split"
    t_end: float            # simulated end time (s)
    arrival: float          # seconds after the start of the stream
    kind: str = ""
    # filled by the server
    admitted: float | None = None
    finished: float | None = None
    ncycle: int = 0
    result: dict | None = None

    @property
    def ne(self) -> int:
        return self.model.ne_hex + self.model.ne_q4

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def footprint(model: M.Model, precision: str) -> int:
    """Bytes the model will hold on the GPU once uploaded (computed, nothing allocated)."""
    import ctypes as C
    npt = model.q4.prop.npt if model.q4 is not None else 0
    rate = int(model.q4 is not None and model.q4.params()["israte"] > 0)
    rows = [None, None]
    for k, part in enumerate((model.hexp, model.q4)):
        if part is not None:
            rows[k] = np.ascontiguousarray(np.bincount(part.conn.reshape(-1), minlength=model.nn), np.int32)
    b = C.c_int64()
    L.check(L.load().or_model_footprint_estimate(model.nn, model.ne_hex, model.ne_q4, npt, E.PRECISIONS[precision],
                                                 rate, *[None if r is None else r.ctypes.data_as(C.c_void_p)
                                                         for r in rows], C.byref(b)))
    return b.value


class Server:
    """Continuous-batching request loop over one batch of resident models."""

    def __init__(self, hbm_budget: int, cycles_per_step: int = 50, device: int = 0, keep_results: bool = True):
        self.budget = int(hbm_budget)
        self.cycles_per_step = int(cycles_per_step)
        self.device = device
        self.keep_results = keep_results
        self.batch = E.GpuBatch(device)
        self.waiting: list[Request] = []
        self.running: dict[int, tuple[Request, E.GpuModel, int]] = {}
        self.used = 0
        self.max_used = 0
        self.done: list[Request] = []
        self.element_updates = 0

    def _admit(self, now: float):
        # first come, first served; a request that does not fit blocks the ones behind it
        while self.waiting:
            req = self.waiting[0]
            need = footprint(req.model, req.precision)
            if need > self.budget:
                raise ValueError(f"request {req.rid} needs {need} bytes, more than the whole budget")
            if self.used + need > self.budget:
                break
            self.waiting.pop(0)
            g = E.GpuModel(req.model, precision=req.precision, device=self.device)
            self.batch.add(g, req.t_end)
            req.admitted = now
            self.running[req.rid] = (req, g, need)
            self.used += need
            self.max_used = max(self.max_used, self.used)

    def _retire(self, now_fn):
        if not self.running:
            return
        n_active = self.batch.active()
        if n_active == len(self.running):
            return
        for rid in list(self.running):
            req, g, need = self.running[rid]
            ts = g.time_state()
            if ts.tt < req.t_end and ts.mstop == 0:
                continue
            self.batch.remove(g)
            req.ncycle = int(ts.ncycle)
            if self.keep_results:
                x, v, vr, d = g.nodes6()
                req.result = dict(d=d, v=v, vr=vr, tt=ts.tt, ncycle=int(ts.ncycle), mstop=int(ts.mstop),
                                  energy=g.energy())
            g.close()
            req.finished = now_fn()
            self.element_updates += req.ne * req.ncycle
            self.used -= need
            del self.running[rid]
            self.done.append(req)

    def serve(self, requests: list[Request], clock=time.perf_counter, max_wall: float = 600.0) -> dict:
        """Play the request stream in real time (arrival offsets from now) until every request is
        done.  Returns latency percentiles and throughput."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        now = lambda: clock() - t0  # noqa: E731
        while pending or self.waiting or self.running:
            t = now()
            if t > max_wall:
                raise TimeoutError("serving loop exceeded its wall-clock limit")
            while pending and pending[0].arrival <= t:
                self.waiting.append(pending.pop(0))
            stepped = bool(self.running)
            if stepped:
                self.batch.run(self.cycles_per_step, sync=False)
            self._admit(t)             # uploads overlap the step just queued
            if stepped:
                self._retire(now)
            elif not self.running and pending:
                time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
        wall = now()
        lat = np.array([r.latency for r in self.done])
        return dict(requests=len(self.done), wall_s=wall, p50_s=float(np.percentile(lat, 50)),
                    p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()),
                    element_updates=int(self.element_updates),
                    element_updates_per_s=self.element_updates / wall)

    def close(self):
        for rid in list(self.running):
            _, g, _ = self.running.pop(rid)
            g.close()
        self.batch.close()


# ---------------------------------------------------------------------------
# request streams
def bar_of(ne: int, v0: float = 1.0e4) -> M.Model:
    """Cantilever bar of about ne hexes with the 1 x 1 x 4 proportions of the 20x20x80 bar."""
    k = max(2, int(round((ne / 4.0) ** (1.0 / 3.0))))
    return M.hex_model_to_model(M.cantilever_bar(k, k, 4 * k, v0=v0))


def tube_of(ne: int, v_top: float = -1.5e4) -> M.Model:
    """Square tube of about ne shells (1 mm elements, length = 5 x width like the 60 x 300 tube)."""
    nside = max(2, int(round(np.sqrt(ne / 20.0))))
    return M.square_tube(width=float(nside), length=float(5 * nside), v_top=v_top)


def request_stream(n: int, rate: float, seed: int = 20261203, ne_range=(5_000, 500_000), ne_median: float = 40_000,
                   ne_sigma: float = 1.0, t_end_range=(1.0e-3, 10.0e-3), time_scale: float = 1.0) -> list[Request]:
    """Poisson arrivals at `rate` per second; each request a bar or a tube (50/50) of lognormal size
    clipped to ne_range, FP64 or split (50/50), end time uniform in t_end_range times time_scale."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += float(rng.exponential(1.0 / rate))
        ne = int(np.clip(rng.lognormal(np.log(ne_median), ne_sigma), *ne_range))
        kind = "bar" if rng.integers(0, 2) == 0 else "tube"
        prec = "fp64" if rng.integers(0, 2) == 0 else "split"
        t_end = float(rng.uniform(*t_end_range)) * time_scale
        m = bar_of(ne) if kind == "bar" else tube_of(ne)
        out.append(Request(rid=rid, model=m, precision=prec, t_end=t_end, arrival=t, kind=kind))
    return out
