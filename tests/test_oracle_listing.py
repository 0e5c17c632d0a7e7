"""The listing's physics on the oracle: FIXVEL's external work (WFEXT) and the controlling element
(NELTS / ITYPTS).

* WFEXT is pinned by the energy balance ECRIT prints: with the imposed top velocity as the only
  energy source of a crushing tube, ENTOT = ENINT + ENCIN + ENROT must follow ENTOT0 + WFEXT
  (ecrit.F:501-512).  Counting only the DT1*DW half of each imposed DOF's work (fixvel.F:834) and
  not its deferred DT2*DW half (fixvel.F:837, added at fixvel.F:342-344 of the next cycle) leaves
  ERR near +100 %; the full count closes the balance to 1e-3.
* NELTS: on an undeformed uniform mesh every element has the same time step, and the reference's
  comparisons then name the last solid (DTX <= DT2T, mqviscb.F:622-630) or the first shell (DT <
  DT2T, cdt3.F:205-215); on a deformed state the scan equals its definition over the oracle's
  per-element candidates."""
import numpy as np

import oracle as O
from openradioss_amd import listing as LS
from openradioss_amd import model as M


def _balance(o, cycles, every):
    e0 = None
    out = []
    for c in range(cycles + 1):
        if c % every:
            o.cycle(1)
            continue
        r = o.cycle_ecrit()
        entot = r["enint"] + r["encin"] + r["enrot"]
        if e0 is None:
            e0 = entot - r["wfext"]
        err = entot / (e0 + r["wfext"]) - 1.0 if abs(e0 + r["wfext"]) > 1e-20 else 0.0
        out.append((c, err, r))
    return out


def test_energy_balance_closes_with_the_imposed_velocity_work():
    m = M.square_tube(width=12.0, length=24.0, h=1.0)
    o = O.ModelOracle(m, nthreads=4)
    rows = _balance(o, 600, 100)
    for c, err, r in rows[1:]:
        entot = r["enint"] + r["encin"] + r["enrot"]
        assert r["wfext"] > 0.9 * entot > 0.0, (c, r)
        assert abs(err) < 1.0e-3, (c, err)


def test_uniform_bar_controlling_element_is_the_last_solid():
    hm = M.cantilever_bar(3, 3, 6)
    o = O.ModelOracle(M.hex_model_to_model(hm), nthreads=2)
    r = o.cycle_ecrit()
    assert (r["itypts"], r["nelts"]) == (LS.ITY_SOLID, hm.ne - 1)


def test_uniform_tube_controlling_element_is_the_first_shell():
    m = M.square_tube(width=12.0, length=24.0, h=1.0)
    o = O.ModelOracle(m, nthreads=2)
    r = o.cycle_ecrit()
    assert (r["itypts"], r["nelts"]) == (LS.ITY_SHELL, 0)


def _by_definition(o):
    """ITYPTS / NELTS from the per-element candidates: the last solid at the smallest solid DTX,
    unless some shell's DT is smaller, then the first shell at the smallest shell DT."""
    best = (0, -1, 1.0e6)
    if o.m.ne_hex:
        d = o._dtel_h[: o.m.ne_hex]
        ok = d >= 0.0
        if ok.any():
            mh = d[ok].min()
            best = (1, int(np.flatnonzero(ok & (d == mh))[-1]), mh)
    if o.m.ne_q4:
        d = o._dtel_q[: o.m.ne_q4]
        ok = d >= 0.0
        if ok.any() and d[ok].min() < best[2]:
            mq = d[ok].min()
            best = (3, int(np.flatnonzero(ok & (d == mq))[0]), mq)
    return best[:2]


def test_controlling_element_equals_its_definition_on_deformed_states():
    for m, cycles in ((M.square_tube(width=12.0, length=24.0, h=1.0), 400),
                      (M.block_with_plate(n=6, plate=10, seed=7), 150),
                      (M.hex_model_to_model(M.cantilever_bar(4, 4, 10)), 300)):
        o = O.ModelOracle(m, nthreads=4)
        o.cycle(cycles)
        r = o.cycle_ecrit()
        assert r["nelts"] >= 0
        assert (r["itypts"], r["nelts"]) == _by_definition(o)


def test_oracle_listing_lines_read_like_verify_results():
    m = M.square_tube(width=12.0, length=24.0, h=1.0)
    o = O.ModelOracle(m, nthreads=4)
    L = LS.Listing(float(np.cumsum(m.ms)[-1]))
    for c, err, r in _balance(o, 300, 100):
        ts = o.ts
        s = L.add(c, ts.tt - ts.dt2, ts.dt2, r["enint"], r["encin"], r["enrot"], r["wfext"], r["itypts"], r["nelts"])
        f = " ".join(s.split()).split(" ")
        assert len(f) == 13 and f[5].endswith("%") and f[3] == "SHELL"
    assert LS.parse_cycle_line(L.lines[-1])["ncycle"] == 300
