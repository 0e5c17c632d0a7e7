"""or_nodal_update (ASSPAR4 -> DEPLA with the reference host arrays) against the oracle's nodal
half of a cycle, bit for bit, on the FSKY rows the oracle's element half just wrote."""
import numpy as np
import pytest

import oracle as O
from golden_cases import Q4_CASES, build_q4_case
from openradioss_amd import engine as E
from openradioss_amd import lib as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["mixed_c120", "tube_8x20_c150"])
def test_nodal_dropin_matches_oracle(name, gpu_lib):
    print(f"case: or_nodal_update on {name} after 40 cycles", flush=True)
    m = build_q4_case(Q4_CASES[name])
    o = O.ModelOracle(m, nthreads=4, math="cr")
    o.cycle(40)
    pre = {k: getattr(o, k).copy() for k in ("x", "v", "vr", "d")}
    ts_pre = (o.ts.tt, o.ts.dt1, o.ts.dt2, o.ts.dt12, o.ts.dt2old, o.ts.ncycle)
    o.cycle(1)                       # element half fills o.fsky, nodal half updates the nodes
    ts = L.OrTimeState()
    ts.tt, ts.dt1, ts.dt2, ts.dt12, ts.dt2old, ts.ncycle = ts_pre
    x, v, vr, d = (pre[k] for k in ("x", "v", "vr", "d"))
    stifn = np.zeros(m.nn)
    # dt2t = the cycle's DT2 reproduces it through the pipeline (DT2 <= 1.1*DT2OLD, DTMX)
    E.nodal_update(np.asarray(m.adsky) + 1, o.fsky, m.ms, x, v, d, ts, dt2t=o.ts.dt2, dtmx=m.dtmx, inr=m.inr,
                   vr=vr, code=m.code, vimp=m.vimp, stifn=stifn)
    assert (ts.tt, ts.dt2, ts.ncycle) == (o.ts.tt, o.ts.dt2, o.ts.ncycle)
    assert np.array_equal(x, o.x) and np.array_equal(v, o.v) and np.array_equal(d, o.d)
    assert np.array_equal(vr, o.vr)
    assert np.array_equal(stifn, o.stifn)


def test_nodal_dropin_zero_time_step_stops(gpu_lib):
    """An element minimum of 0 gives DT2 = 0: the drop-in sets MSTOP's time-step bit like the oracle."""
    print("case: or_nodal_update with dt2t = 0 on mixed_c120 after 5 cycles", flush=True)
    m = build_q4_case(Q4_CASES["mixed_c120"])
    o = O.ModelOracle(m, nthreads=2, math="cr")
    o.cycle(5)
    ts = L.OrTimeState()
    ts.tt, ts.dt1, ts.dt2, ts.dt12, ts.dt2old, ts.ncycle = (o.ts.tt, o.ts.dt1, o.ts.dt2, o.ts.dt12, o.ts.dt2old,
                                                            o.ts.ncycle)
    x, v, vr, d = o.x.copy(), o.v.copy(), o.vr.copy(), o.d.copy()
    E.nodal_update(np.asarray(m.adsky) + 1, o.fsky, m.ms, x, v, d, ts, dt2t=0.0, dtmx=m.dtmx, inr=m.inr,
                   vr=vr, code=m.code, vimp=m.vimp)
    assert ts.dt2 == 0.0 and ts.mstop == L.OR_MSTOP_TIME_STEP
