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
