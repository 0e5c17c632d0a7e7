"""The serving loop: requests admitted by HBM footprint into one running batch; every served
result equals the same model advanced alone, bit for bit."""
import numpy as np
import pytest

from openradioss_amd import engine as E
from openradioss_amd import lib as L
from openradioss_amd import serve as S

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone_runs(gpu_lib):
    print("case: serving 8 requests, budget for about 3 at a time", flush=True)
    reqs = S.request_stream(8, rate=200.0, seed=7, ne_range=(300, 3000), ne_median=1200, time_scale=0.02)
    fps = [S.footprint(r.model, r.precision) for r in reqs]
    budget = int(2.5 * max(fps))
    srv = S.Server(hbm_budget=budget, cycles_per_step=20)
    stats = srv.serve(reqs)
    srv.close()
    print(stats)
    assert stats["requests"] == 8
    assert srv.max_used <= budget
    assert stats["p99_s"] >= stats["p50_s"] > 0
    for r in reqs:
        assert r.result["tt"] >= r.t_end
        g = E.GpuModel(r.model, precision=r.precision)
        fp = g.footprint()
        # the estimate bounds the rest rows by every row of the model (the tiles sum most of
        # them in LDS), so it is an upper bound of the allocation
        assert fp <= S.footprint(r.model, r.precision) <= 2 * fp
        g.run(r.ncycle)
        x, v, vr, d = g.nodes6()
        assert np.array_equal(d, r.result["d"]) and np.array_equal(v, r.result["v"])
        assert g.time_state().tt == r.result["tt"]
        g.close()


def test_served_crash_solves_at_full_size(gpu_lib):
    """Six requests of the served workload: bars and tubes of 5k-500k elements (lognormal),
    FP64 and split, end times of 1-2 ms, Poisson arrivals, admitted by footprint into one batch.
    The FP64 request of the fewest elements equals the oracle bit for bit; every request equals
    its model advanced alone to the same cycle."""
    import oracle as O
    print("case: serving 6 requests of 5k-500k elements, end times 1-2 ms", flush=True)
    reqs = S.request_stream(6, rate=20.0, seed=11, t_end_range=(1.0e-3, 2.0e-3))
    sizes = sorted(r.ne for r in reqs)
    print("request sizes", sizes, [r.precision for r in reqs], flush=True)
    assert sizes[0] >= 5_000 and sizes[-1] > 100_000
    free, _ = L.device_memory(0)
    srv = S.Server(hbm_budget=int(0.5 * free), cycles_per_step=200)
    stats = srv.serve(reqs)
    srv.close()
    print(stats, flush=True)
    assert stats["requests"] == 6
    fp64 = [r for r in reqs if r.precision == "fp64" and r.result["mstop"] == 0]
    small = min(fp64, key=lambda r: r.ne)
    for r in reqs:
        # a request ends at its end time, or earlier when the run stops (MSTOP: a time step <= 0
        # once the folds of a tube have passed through each other; the path has no contact)
        assert r.result["tt"] >= r.t_end or r.result["mstop"] != 0
        print(f"request {r.rid}: {r.kind} {r.ne} {r.precision}, {r.ncycle} cycles, t {r.result['tt']:.4e}, "
              f"MSTOP {r.result['mstop']}", flush=True)
        g = E.GpuModel(r.model, precision=r.precision)
        try:
            g.run(r.ncycle)
        except L.OrError as err:
            assert err.code == L.OR_ERR_MSTOP and r.result["mstop"] != 0
        x, v, vr, d = g.nodes6()
        assert np.array_equal(d, r.result["d"], equal_nan=True) and np.array_equal(v, r.result["v"], equal_nan=True)
        ts = g.time_state()
        assert ts.tt == r.result["tt"] and ts.mstop == r.result["mstop"]
        g.close()
    o = O.ModelOracle(small.model, nthreads=16, math="cr")
    o.cycle(small.ncycle)
    print(f"oracle check: request {small.rid}, {small.kind}, {small.ne} elements, {small.ncycle} cycles", flush=True)
    assert o.ts.tt == small.result["tt"] and np.array_equal(o.d, small.result["d"])
