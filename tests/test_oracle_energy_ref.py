"""The oracle's internal energy against the reference's own energy bookkeeping: SBILAN and CBILAN
(compiled from the reference sources into oracle/_ref) add EI = FILL*EINT*VOL per solid
(sbilan.F:140) and EINT(I,1) + EINT(I,2) per shell (cbilan.F:168) into PARTSAV(1,part), group
after group; ECRIT adds the parts (ecrit.F:344).  The oracle's sums equal theirs bit for bit on
states after many cycles (plastic bar, tube crush with rupture, hex block with a tied plate)."""
import numpy as np
import pytest

import oracle as O
import refchain as R
from golden_cases import CASES, Q4_CASES, build_case, build_q4_case
from openradioss_amd import model as M

pytestmark = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs the reference tree)")


def test_hex_internal_energy_equals_sbilan():
    c = CASES["impact_4x4x16_c120"]
    m = build_case(c)
    o = O.HexOracle(m, nthreads=2)
    o.cycle(c["cycles"])
    mm = M.hex_model_to_model(m)
    eh, eq = R.bilan_internal_energy(mm, heint=o.eint, hrho=o.rho, hoffg=o.offg, v=o.v)
    assert eq == 0.0 and eh > 0.0
    assert eh == o.energy()[0]


@pytest.mark.parametrize("name", ["tube_rupture_c150", "mixed_c120", "plate_dtmin_c120"])
def test_model_internal_energy_equals_sbilan_plus_cbilan(name):
    c = Q4_CASES[name]
    m = build_q4_case(c)
    o = O.ModelOracle(m, nthreads=2)
    o.cycle(c["cycles"])
    kw = dict(v=o.v, x=o.x, vr=o.vr)
    if m.hexp is not None:
        kw.update(heint=o.heint, hrho=o.hrho, hoffg=o.hoffg)
    if m.q4 is not None:
        kw.update(q_eint=o.q["eint"], q_offg=o.q["offg"])
    eh, eq = R.bilan_internal_energy(m, **kw)
    assert eh + eq > 0.0
    assert eh + eq == o.energy()[0]
