"""A model read from a Starter deck (two law-2 bricks under two Q4 shells, clamped, with an
initial and an imposed velocity) runs on the GPU bit for bit with the oracle in FP64."""
import os

import numpy as np
import pytest

import oracle as O
from openradioss_amd import deck as D
from openradioss_amd import engine as E

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "deck_brick_shell.rad")


def test_deck_model_runs_bit_for_bit(gpu_lib):
    print("case: tests/golden/deck_brick_shell.rad, 400 cycles, fp64 vs oracle(cr)", flush=True)
    m = D.read(FIXTURE)
    with E.GpuModel(m) as g:
        g.run(400)
        x, v, vr, d = g.nodes6()
        ts = g.time_state()
    o = O.ModelOracle(m, nthreads=1, math="cr")
    o.cycle(400)
    assert ts.tt == o.ts.tt and ts.mstop == 0
    assert np.array_equal(d, o.d) and np.array_equal(v, o.v) and np.array_equal(vr, o.vr)
    assert np.abs(d).max() > 0.05          # the top has moved: the run did plastic work


def test_shuffled_element_ids_run_bit_for_bit(gpu_lib):
    """Element user ids out of element order (a deck's ids need not follow its card order): the
    skyline rows and nodal sums follow the ids, and the GPU's tiled assembly sums them in that
    order, bit for bit with the oracle (distorted hex block with a tied plate, 120 cycles)."""
    print("case: block 8 + plate 12 with shuffled element ids, 120 cycles, fp64 vs oracle(cr)", flush=True)
    from openradioss_amd import model as M
    base = M.block_with_plate(n=8, plate=12, seed=21)
    rng = np.random.default_rng(22)
    hid = rng.permutation(base.ne_hex) * 2 + 5
    qid = rng.permutation(base.ne_q4) * 3 + 100000
    m = M.build_model(base.x, hex_conn=base.hexp.conn, hex_mat=base.hexp.mat, hex_prop=base.hexp.prop,
                      q4_conn=base.q4.conn, q4_mat=base.q4.mat, q4_prop=base.q4.prop, thke=base.q4.thke,
                      v0=base.v, code=base.code, vimp=base.vimp, hex_ids=hid, q4_ids=qid)
    assert not np.array_equal(m.hexp.iads, base.hexp.iads)
    with E.GpuModel(m) as g:
        g.run(120)
        x, v, vr, d = g.nodes6()
        ts = g.time_state()
    o = O.ModelOracle(m, nthreads=4, math="cr")
    o.cycle(120)
    assert ts.tt == o.ts.tt and ts.mstop == 0
    assert np.array_equal(d, o.d) and np.array_equal(v, o.v) and np.array_equal(vr, o.vr)
