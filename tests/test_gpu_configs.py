"""The reference workloads at their full sizes.

* cantilever bar 20x20x80 (32,000 hexes), law-2 steel, 1,000 cycles: FP64 GPU bit for bit equal
  to the oracle under correctly rounded pow/log/exp;
* square tube 60x300 (72,000 Q4 shells, 5 points), axial crush at 15 m/s: FP64 GPU bit for bit
  equal to the oracle over 12,000 cycles (0.75 ms, past the first fold), and over the whole
  20,000 when OR_FULL_RUNS=1; the split form against
  the FP64 form up to
  about 20,000 cycles at matched simulated times, beside the FP64 method's own sensitivity;
* hex block 160^3 with the 1000x1000 Q4 plate (4,096,000 hexes + 1,000,000 shells, 5,149,361
  nodes, the bench workload): FP64 GPU bit for bit equal to the oracle over its first cycles;
* the bar after 1,000 cycles and the tube after 8,800 cycles against fixtures written by the
  reference's own compiled routines: the tube within the 1e-12 FP64 parity bound, the bar within
  the bound the host libm sets (3.57e-12 measured between the reference algorithm under glibc
  2.35 and under correctly rounded pow; tests/test_ref_full_length.py counts its sources).
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from openradioss_amd import engine as E
from openradioss_amd import model as M

pytestmark = pytest.mark.gpu

THREADS = 16
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PARITY_TOL = 1e-12        # FP64 displacement parity against the reference
BAR_LIBM_BOUND = 5e-12    # the bar's displacement spread caused by the host libm alone (3.57e-12)


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_bar_and_tube_against_the_reference_at_full_length(gpu_lib):
    print("case: bar 20x20x80 1,000 cycles and tube 60x300 8,800 cycles, fp64 vs reference fixtures",
          flush=True)
    gb = np.load(os.path.join(GOLD, "bar_20x20x80_c1000.npz"))
    g = E.GpuModel(M.hex_model_to_model(M.cantilever_bar(20, 20, 80)))
    g.run(1000)
    ts = g.time_state()
    x, v, vr, d = g.nodes6()
    e = rel_err(d, gb["d_sub"])
    print(f"bar: displacement rel err {e:.3e} (libm bound {BAR_LIBM_BOUND:g}), time {ts.tt!r} vs "
          f"{gb['time'][0]!r}", flush=True)
    assert ts.ncycle == 1000 and e <= BAR_LIBM_BOUND
    assert abs(ts.tt - gb["time"][0]) <= PARITY_TOL * gb["time"][0]
    g.close()
    gt = np.load(os.path.join(GOLD, "tube_60x300_c8800.npz"))
    g = E.GpuModel(M.square_tube())
    g.run(8800)
    ts = g.time_state()
    x, v, vr, d = g.nodes6()
    e = rel_err(d[::int(gt["d_every"][0])], gt["d_sub"])
    same = hashlib.sha256(np.ascontiguousarray(d).tobytes()).hexdigest() == str(gt["sha_d"])
    print(f"tube: displacement rel err {e:.3e}, bit for bit: {same}", flush=True)
    assert ts.ncycle == 8800 and ts.mstop == 0 and e <= PARITY_TOL
    g.close()


def test_bar_20x20x80_1000_cycles_bitwise(gpu_lib):
    print("case: bar 20x20x80, 32,000 hexes, 1,000 cycles, fp64 vs oracle(cr)", flush=True)
    m = M.cantilever_bar(20, 20, 80)
    assert m.ne == 32_000 and m.nn == 35_721
    g = E.GpuHexModel(m)
    g.run(1000)
    o = O.HexOracle(m, nthreads=THREADS, math="cr")
    o.cycle(1000)
    x, v, d = g.nodes()
    st = g.hex_state()
    assert g.time_state().tt == o.ts.tt
    assert np.array_equal(d, o.d) and np.array_equal(v, o.v)
    assert np.array_equal(st["sig"], o.sig) and np.array_equal(st["pla"], o.pla)
    print(f"t = {o.ts.tt:.6e} s, max |d| = {np.abs(d).max():.4e} mm, max pla = {o.pla.max():.3e}")
    g.close()


def test_tube_60x300_12000_cycles_bitwise(gpu_lib):
    """The tube crush through its first fold (about 0.65 ms): bit for bit there is what the
    chaos after the fold cannot hide (a one-ulp difference grows to O(1)).  12,000 cycles keep
    the oracle's share near a minute; the whole 20,000 are the opt-in test below."""
    print("case: tube 60x300, 72,000 shells, 12,000 cycles, fp64 vs oracle(cr)", flush=True)
    m = M.square_tube()
    assert m.ne_q4 == 72_000 and m.nn == 72_240
    g = E.GpuModel(m)
    g.run(12000)
    o = O.ModelOracle(m, nthreads=THREADS, math="cr")
    o.cycle(12000)
    x, v, vr, d = g.nodes6()
    q = g.q4_state()
    assert g.time_state().tt == o.ts.tt
    assert np.array_equal(d, o.d) and np.array_equal(v, o.v) and np.array_equal(vr, o.vr)
    assert np.array_equal(q["sig"], o.q["sig"]) and np.array_equal(q["thk"], o.q["thk"])
    print(f"t = {o.ts.tt:.6e} s, max |d| = {np.abs(d).max():.4e} mm, max pla = {o.q['pla'].max():.3e}")
    g.close()


@pytest.mark.skipif(os.environ.get("OR_FULL_RUNS") != "1",
                    reason="the oracle's 20,000 cycles take about 110 s on 16 threads; set OR_FULL_RUNS=1 "
                           "(and a per-test timeout above that) to run it")
def test_tube_60x300_20000_cycles_bitwise(gpu_lib):
    """The tube's whole run, 20,000 cycles (1.246 ms): D, V, VR and the time bit for bit."""
    print("case: tube 60x300, 72,000 shells, 20,000 cycles, fp64 vs oracle(cr)", flush=True)
    m = M.square_tube()
    with E.GpuModel(m) as g:
        g.run(20000)
        x, v, vr, d = g.nodes6()
        tt = g.time_state().tt
    o = O.ModelOracle(m, nthreads=THREADS, math="cr")
    o.cycle(20000)
    assert tt == o.ts.tt
    assert np.array_equal(d, o.d) and np.array_equal(v, o.v) and np.array_equal(vr, o.vr)


def test_tube_60x300_split_vs_fp64_to_20000_cycles(gpu_lib):
    """Split form against FP64 at matched simulated times, next to the FP64 method's own
    sensitivity (the same model with one coordinate of one node moved by one ulp).

    Up to the first fold (0.6 ms, about 8,800 cycles) the split form stays within 1e-4 in
    displacement and energy (measured: 1e-6 and 4e-8).  After the first fold the crush is
    chaotic: at 1.25 ms (about 20,000 cycles) a one-ulp perturbation of the FP64 run moves the
    displacement field by O(0.1-1) of its maximum, and the split form differs from FP64 by the
    same amount.  The bound there is therefore relative to that sensitivity."""
    print("case: tube 60x300, split vs fp64 vs fp64 perturbed by 1 ulp, to 0.6 ms and 1.25 ms", flush=True)
    m = M.square_tube()
    mp = M.square_tube()
    k = mp.nn // 2 + 7
    mp.x[k, 0] = np.nextafter(mp.x[k, 0], np.inf)
    runs = {"fp64": E.GpuModel(m), "pert": E.GpuModel(mp), "split": E.GpuModel(m, precision="split")}
    b = E.GpuBatch()
    err = {}
    for t_end in (0.6e-3, 1.25e-3):
        for g in runs.values():
            b.add(g, t_end)
        while b.active() > 0:
            b.run(1000)
        res = {}
        for name, g in runs.items():
            x, v, vr, d = g.nodes6()
            res[name] = (d, g.energy().sum(), g.time_state())
            b.remove(g)
        d64, e64, t64 = res["fp64"]
        for name in ("pert", "split"):
            d, e, ts = res[name]
            err[(t_end, name)] = (np.abs(d - d64).max() / np.abs(d64).max(), abs(e - e64) / abs(e64))
        print(f"t_end {t_end * 1e3:.2f} ms, {t64.ncycle} cycles: perturbed fp64 d/e err {err[(t_end, 'pert')]}, "
              f"split d/e err {err[(t_end, 'split')]}", flush=True)
    assert err[(0.6e-3, "pert")][0] <= 1e-10
    assert err[(0.6e-3, "split")][0] <= 1e-4 and err[(0.6e-3, "split")][1] <= 1e-4
    assert err[(1.25e-3, "pert")][0] >= 1e-2          # the crush is chaotic at 20,000 cycles
    assert err[(1.25e-3, "split")][0] <= 2.0 * err[(1.25e-3, "pert")][0] + 1e-4
    assert err[(1.25e-3, "split")][1] <= 5e-2
    b.close()
    for g in runs.values():
        g.close()


def test_block_plate_full_size_first_cycles_bitwise(gpu_lib):
    print("case: block 160^3 + plate 1000^2, 5,096,000 elements, 4 cycles, fp64 vs oracle(cr)", flush=True)
    m = M.block_with_plate(n=160, plate=1000, seed=20261203)
    assert (m.ne_hex, m.ne_q4, m.nn) == (4_096_000, 1_000_000, 5_149_361)
    with E.GpuModel(m, precision="fp64") as g:
        g.run(4)
        x, v, vr, d = g.nodes6()
        ts = g.time_state()
        hs, qs = g.hex_state(), g.q4_state()
    o = O.ModelOracle(m, nthreads=THREADS, math="cr")
    o.cycle(4)
    assert (ts.tt, ts.dt2, ts.ncycle) == (o.ts.tt, o.ts.dt2, o.ts.ncycle)
    for name, a, b in (("x", x, o.x), ("v", v, o.v), ("vr", vr, o.vr), ("d", d, o.d),
                       ("hex sig", hs["sig"], o.hsig), ("q4 sig", qs["sig"], o.q["sig"]),
                       ("q4 hourg", qs["hourg"], o.q["hourg"])):
        assert np.array_equal(a, b), name
    print(f"t = {o.ts.tt:.6e} s after 4 cycles, max |d| = {np.abs(d).max():.4e} mm")
