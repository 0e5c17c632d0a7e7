"""The serving loop: requests admitted by HBM footprint into one running batch; every served
result equals the same model advanced alone, bit for bit."""
import numpy as np
import pytest

from openradioss_amd import engine as E
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
