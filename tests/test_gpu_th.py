"""The time history on the GPU (GpuModel.run_time_history through or_model_th_setup /
or_model_run_th) against the oracle's, written by the same writer:

* the same output times and the header byte for byte;
* every group record (node D, V, A, VR, AR, X; solid and shell element values) byte for byte: the
  FP64 state is equal bit for bit under correctly rounded pow/log/exp, so its REAL*4 is too;
* the global variables within 2 float32 ulps: the GPU sums the energies, momenta, external work
  and EHOUR in a fixed tree order (EHOUR per tile over the cycles), the oracle node after node and
  group after group, so their FP64 values differ in the last bits;
* EHOUR carried through a checkpoint;
* the split form's file: the FP64 run's layout and output cycles, its own element state.
"""
import numpy as np
import pytest

import oracle as O
import th_oracle as TO
from openradioss_amd import engine as E
from openradioss_amd import model as M
from openradioss_amd import th as TH

pytestmark = pytest.mark.gpu


def _groups(m):
    g = [TH.ThGroup.make("NODE", 1, "nodes", ["DEF", "A", "XYZ", "VR", "AR"],
                         [0, m.nn // 3, m.nn // 2, m.nn - 1], [1, m.nn // 3 + 1, m.nn // 2 + 1, m.nn])]
    if m.ne_hex:
        g.append(TH.ThGroup.make("BRIC", 2, "bricks", ["DEF", "BULK", "VOL", "PLSR"],
                                 [0, m.ne_hex // 2, m.ne_hex - 1], [1, m.ne_hex // 2 + 1, m.ne_hex]))
    if m.ne_q4:
        g.append(TH.ThGroup.make("SHEL", 3, "shells", ["DEF", "STRESS", "THIC"],
                                 [0, m.ne_q4 // 2, m.ne_q4 - 1], [1, m.ne_q4 // 2 + 1, m.ne_q4]))
    return TH.ThHeader("case", [(1, "part", 1, 1)], [(1, "law2")], [(1, "prop")], g,
                       date="Thu Dec  3 12:00:00 2026")


CASES = {
    "mixed": (lambda: M.block_with_plate(n=8, plate=16, seed=11), 60, 4),
    "tube": (lambda: M.square_tube(width=12.0, length=24.0, h=1.0), 80, 5),
    "bar": (lambda: M.hex_model_to_model(M.cantilever_bar(6, 5, 12)), 60, 3),
}


def _records(buf):
    out, p = [], 0
    while p < len(buf):
        n = int.from_bytes(buf[p:p + 4], "big")
        out.append(buf[p:p + 8 + n])
        p += 8 + n
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_time_history_equals_oracle(name):
    make, ncyc, every = CASES[name]
    m = make()
    h = _groups(m)
    probe = O.ModelOracle(m)
    probe.cycle(1)
    dthis = every * probe.ts.dt2   # an output every few cycles
    ref, recs, this_o = TO.oracle_time_history(m, h, dthis, 1.0, ncyc)
    with E.GpuModel(m) as g:
        data, this_g = g.run_time_history(h, dthis, 1.0, max_cycles=ncyc)
    assert this_g == this_o
    a, b = TH.read(data), TH.read(ref)
    assert len(a["times"]) == len(b["times"]) == len(recs) >= ncyc // every
    ra, rb = _records(data), _records(ref)
    k = a["header_records"]
    assert ra[:k] == rb[:k]
    for ta, tb in zip(a["times"], b["times"]):
        assert ta["time"] == tb["time"]
        for ga, gb in zip(ta["groups"], tb["groups"]):
            assert ga.tobytes() == gb.tobytes()
        x, y = ta["glob"].astype(np.float64), tb["glob"].astype(np.float64)
        assert x[5] == y[5] and x[6] == y[6]                  # XMASS, DT2
        scale = np.maximum(np.abs(y), 1e-30)
        scale[2:5] = np.abs(y[2:5]).max() + 1e-30          # momenta: relative to the largest
        assert np.all(np.abs(x - y) <= 2.5e-7 * scale), (name, x, y)
    assert b["times"][-1]["glob"][11] > 0                      # EHOUR


def test_ehour_through_a_checkpoint():
    m = M.block_with_plate(n=8, plate=16, seed=11)
    with E.GpuModel(m) as g:
        g.th_setup([0], [0], [0])
        g.run(30)
        ck = g.checkpoint()
        g.run(20)
        whole = g.ehour()
    with E.GpuModel.restore(m, ck) as r:
        r.th_setup([0], [0], [0])
        assert r.ehour() == float(ck["ehour"][0])
        r.run(20)
        assert abs(r.ehour() - whole) <= 1e-13 * whole


def test_deck_time_history_equals_oracle():
    """A deck's /TH/NODE, /TH/BRIC and /TH/SHEL groups (read by the deck reader, header built from
    its parts and titles) written on the GPU and by the oracle."""
    import os
    import test_th_run as TR
    from openradioss_amd import deck as D
    here = os.path.dirname(os.path.abspath(__file__))
    text = open(os.path.join(here, "golden", "deck_brick_shell.rad")).read().replace("/END\n", "") + TR.DECK_TH
    deck = D.parse(text)
    m = D.to_model(deck)
    h = TH.header_from_deck(deck, m, date="Thu Dec  3 12:00:00 2026")
    probe = O.ModelOracle(m)
    probe.cycle(1)
    dthis = 3 * probe.ts.dt2
    ref, recs, this_o = TO.oracle_time_history(m, h, dthis, 1.0, 40)
    with E.GpuModel(m) as g:
        data, this_g = g.run_time_history(h, dthis, 1.0, max_cycles=40)
    a, b = TH.read(data), TH.read(ref)
    assert this_g == this_o and len(a["times"]) == len(b["times"]) >= 10
    ra, rb = _records(data), _records(ref)
    assert ra[:a["header_records"]] == rb[:b["header_records"]]
    assert [g["names"] for g in a["groups"]] == [["", "second"], ["", "tip"], [""]]
    for ta, tb in zip(a["times"], b["times"]):
        assert ta["time"] == tb["time"]
        assert all(x.tobytes() == y.tobytes() for x, y in zip(ta["groups"], tb["groups"]))


def test_split_form_time_history():
    """The split form (FP32 element arithmetic and state, FP64 nodes): the same header and output
    cycles as the FP64 run, times within 1e-5, node displacements within 1e-4 of the largest
    displacement (the split form's tolerance) and velocities within 1e-3 of the largest velocity.
    Its element values are its own state: with a record every cycle, the last record's solid
    values (SIG, EINT, RHO, QVIS, PLA, EPSD) equal the split run's downloaded state after that
    cycle, converted to REAL*4."""
    m = M.block_with_plate(n=8, plate=16, seed=11)
    h = _groups(m)
    probe = O.ModelOracle(m)
    probe.cycle(1)
    dthis = 4 * probe.ts.dt2
    with E.GpuModel(m) as g:
        d64, _ = g.run_time_history(h, dthis, 1.0, max_cycles=60)
    with E.GpuModel(m, precision="split") as g:
        d32, _ = g.run_time_history(h, dthis, 1.0, max_cycles=60)
    a, b = TH.read(d32), TH.read(d64)
    k = b["header_records"]
    assert _records(d32)[:k] == _records(d64)[:k]
    assert len(a["times"]) == len(b["times"]) >= 15
    j = [g.kind for g in h.groups].index("NODE")
    for ta, tb in zip(a["times"], b["times"]):
        assert abs(float(ta["time"]) - float(tb["time"])) <= 1e-5 * float(tb["time"])
        for cols, tol in (([0, 1, 2], 1e-4), ([3, 4, 5], 1e-3)):     # D, V
            x = ta["groups"][j][:, cols].astype(np.float64)
            y = tb["groups"][j][:, cols].astype(np.float64)
            assert np.all(np.abs(x - y) <= tol * np.abs(y).max()), (cols, x, y)
    with E.GpuModel(m, precision="split") as g:
        every, _ = g.run_time_history(h, 1.0e-30, 1.0, max_cycles=30)
        hs = g.hex_state()
    last = TH.read(every)["times"][-1]
    e = h.groups[[g.kind for g in h.groups].index("BRIC")].index
    bg = last["groups"][[g.kind for g in h.groups].index("BRIC")]
    assert np.array_equal(bg[:, 1:7], hs["sig"][:, e].T.astype(np.float32))
    for col, key in ((7, "eint"), (8, "rho"), (9, "qvis"), (11, "pla"), (13, "epsd")):
        assert np.array_equal(bg[:, col], hs[key][e].astype(np.float32)), key
