"""Batches of resident models (or_batch_*): every member advances by its own cycles and time
step, and its result is bit-for-bit the result of the same model advanced alone."""
import numpy as np
import pytest

from openradioss_amd import engine as E
from openradioss_amd import model as M

pytestmark = pytest.mark.gpu


def sweep(n, seed):
    # small tubes (12 x 24 mm) so that every element size of the sweep divides the section
    return M.tube_sweep(n, seed=seed, width=12.0, length=24.0)


def alone(model, precision, ncycles):
    g = E.GpuModel(model, precision=precision)
    g.run(ncycles)
    return g


def same_state(a: E.GpuModel, b: E.GpuModel):
    ta, tb = a.time_state(), b.time_state()
    assert (ta.ncycle, ta.tt, ta.dt2) == (tb.ncycle, tb.tt, tb.dt2)
    for xa, xb in zip(a.nodes6(), b.nodes6()):
        assert np.array_equal(xa, xb)
    qa, qb = a.q4_state(), b.q4_state()
    for k in qa:
        assert np.array_equal(qa[k], qb[k]), k


def test_members_stop_at_their_end_time_and_match_alone_runs(gpu_lib):
    print("case: batch of 6 tubes, fp64 and split, own end times", flush=True)
    models = sweep(6, seed=5)
    precs = ["fp64", "split", "fp64", "split", "fp64", "split"]
    b = E.GpuBatch()
    gms, tends = [], []
    for k, ((m, p), prec) in enumerate(zip(models, precs)):
        g = E.GpuModel(m, precision=prec)
        tend = (20 + 15 * k) * 1.0e-7 * p["h"]          # a few tens of cycles each, all different
        b.add(g, tend)
        gms.append(g)
        tends.append(tend)
    nrun = 0
    while b.active() > 0:
        b.run(25)
        nrun += 25
        assert nrun < 5000
    for g, tend, (m, p), prec in zip(gms, tends, models, precs):
        ts = g.time_state()
        assert ts.tt >= tend and ts.tt - ts.dt2 < tend      # stopped at the first cycle reaching tend
        ref = alone(m, prec, ts.ncycle)
        same_state(g, ref)
        ref.close()
    b.close()
    for g in gms:
        g.close()


def test_admission_while_others_run(gpu_lib):
    """Continuous batching: a model joins after others have run; every result equals its alone run."""
    print("case: batch admission mid-run, 3 tubes", flush=True)
    models = sweep(3, seed=11)
    b = E.GpuBatch()
    g0 = E.GpuModel(models[0][0], precision="fp64")
    g1 = E.GpuModel(models[1][0], precision="split")
    b.add(g0, 1.0)
    b.add(g1, 1.0)
    b.run(37)
    g2 = E.GpuModel(models[2][0], precision="fp64")
    g2.run(5)                       # a few cycles alone first, then it joins the batch
    b.add(g2, 1.0)
    b.run(41)
    b.remove(g1)                    # leaves early
    b.run(10)
    assert b.active() == 2
    for g, (m, _), prec in ((g0, models[0], "fp64"), (g1, models[1], "split"), (g2, models[2], "fp64")):
        ref = alone(m, prec, g.time_state().ncycle)
        same_state(g, ref)
        ref.close()
    assert g0.time_state().ncycle == 88 and g1.time_state().ncycle == 78 and g2.time_state().ncycle == 56
    b.close()
    for g in (g0, g1, g2):
        g.close()


def test_removed_member_runs_alone_again(gpu_lib):
    print("case: batch remove then run alone", flush=True)
    (m, _), = sweep(1, seed=3)
    b = E.GpuBatch()
    g = E.GpuModel(m)
    b.add(g, 1.0)
    b.run(30)
    b.remove(g)
    g.run(20)
    ref = alone(m, "fp64", 50)
    same_state(g, ref)
    b.close()
    g.close()
    ref.close()


def test_full_tube_sweep_members_equal_the_oracle(gpu_lib):
    """The 256-tube parametric sweep (60 x 300 mm tubes, element size 1-4 mm, 4.5k-72k shells each,
    6,032,500 shells) as one batch for 200 batch cycles; 8 members spanning every element size,
    each at its own time step, equal the oracle (correctly rounded pow / log / exp) bit for bit."""
    import oracle as O
    print("case: 256-tube sweep, one fp64 batch, 200 cycles; 8 members against the oracle", flush=True)
    sw = M.tube_sweep(256, seed=20261203)
    b = E.GpuBatch()
    gms = []
    for m, _ in sw:
        g = E.GpuModel(m, precision="fp64")
        b.add(g, 5.0e-3)
        gms.append(g)
    b.run(200)
    assert b.active() == 256
    picks = []
    for h in M.SWEEP_SIZES:                 # two members of every element size but the largest
        idx = [i for i, (_, p) in enumerate(sw) if p["h"] == h]
        picks += idx[:2] if h != 1.0 else idx[:1]
    picks = picks[:8]
    dts = set()
    for i in picks:
        m, p = sw[i]
        ts = gms[i].time_state()
        assert ts.ncycle == 200 and ts.mstop == 0
        o = O.ModelOracle(m, nthreads=16, math="cr")
        o.cycle(200)
        x, v, vr, d = gms[i].nodes6()
        print(f"tube {i}: h {p['h']} mm, {m.ne_q4} shells, dt {ts.dt2:.4e} s", flush=True)
        assert (ts.tt, ts.dt2) == (o.ts.tt, o.ts.dt2)
        assert np.array_equal(d, o.d) and np.array_equal(v, o.v) and np.array_equal(vr, o.vr)
        dts.add(ts.dt2)
    assert len(dts) == len(picks)           # each model at its own time step
    b.close()
    for g in gms:
        g.close()
