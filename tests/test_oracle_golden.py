"""The C oracle against golden vectors written by the reference's own compiled routines
(tests/golden/make_golden.py over oracle/_ref): bit-for-bit equality."""
import os

import numpy as np
import pytest

import oracle as O
import refchain as R
from golden_cases import CASES, build_case

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_golden_bitwise(name):
    g = np.load(os.path.join(HERE, "golden", name + ".npz"))
    m = build_case(CASES[name])
    o = O.HexOracle(m, nthreads=4)
    for _ in range(CASES[name]["cycles"]):      # a cycle that sets MSTOP is the run's last
        o.cycle(1)
        if o.ts.mstop:
            break
    assert o.ts.ncycle == int(g["ncycle"][0])
    assert o.ts.mstop == int(g["mstop"][0] if "mstop" in g else 0)
    t = g["time"]
    assert (o.ts.tt, o.ts.dt1, o.ts.dt2, o.ts.dt12, o.ts.dt2old) == tuple(t)
    for key in ("x", "v", "d", "sig", "eint", "rho", "pla", "epsd", "offg", "qvis"):
        assert np.array_equal(getattr(o, key), g[key]), key


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs the reference tree)")
def test_oracle_matches_live_reference_per_cycle():
    """Cycle-by-cycle equality with the compiled reference on a plastic impact."""
    m = build_case(dict(kind="bar", nx=3, ny=3, nz=10, v0=8.0e4, cycles=0))
    o = O.HexOracle(m)
    r = R.RefHex(m)
    for _ in range(40):
        o.cycle(1)
        r.cycle(1)
        assert np.array_equal(o.x, r.x) and np.array_equal(o.sig, r.sig) and o.ts.dt2 == r.dt2
    assert o.pla.max() > 0.05


def test_golden_cases_exercise_their_branches():
    """Rate hardening raised the flow stress (same impact without it yields more plastic strain),
    and the crushed block stopped on a negative Jacobian with its element deleted."""
    base = np.load(os.path.join(HERE, "golden", "impact_4x4x16_c120.npz"))
    for name in ("impact_rate_c120", "impact_vp3_c120"):
        g = np.load(os.path.join(HERE, "golden", name + ".npz"))
        assert g["epsd"].max() > 0.0 and g["pla"].max() < base["pla"].max(), name
    g = np.load(os.path.join(HERE, "golden", "inverted_mstop.npz"))
    assert int(g["mstop"][0]) == 1 and int(g["ncycle"][0]) < CASES["inverted_mstop"]["cycles"]
    assert (g["offg"] == 0.0).sum() >= 1
