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
    o.cycle(CASES[name]["cycles"])
    assert o.ts.ncycle == int(g["ncycle"][0])
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
