"""GPU parity of the Q4 shell and mixed hex + Q4 paths: HIP kernels (through the C ABI)
against the C oracle, which is pinned bitwise to the reference's compiled CFORC3 chain.

FP64 runs are compared bit for bit with the oracle built under the correctly rounded
exp / log of the device path (math="cr"), and with the reference's golden vectors
(glibc libm) within the spread that the libm alone causes (below 1.1e-14 on these
cases between the two oracle builds)."""
import os

import numpy as np
import pytest

import oracle as O
from golden_cases import Q4_CASES, build_q4_case
from openradioss_amd import engine as E
from openradioss_amd import model as M

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIBM_SPREAD_TOL = 1e-13


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def report(tag, a, b):
    print(f"{tag}: bitwise-equal fraction {np.mean(a == b):.4f}, max rel err {rel_err(a, b):.3e}", flush=True)


def run_pair(model, ncycles, precision="fp64", math="cr"):
    o = O.ModelOracle(model, nthreads=8, math=math)
    g = E.GpuModel(model, precision=precision)
    o.cycle(ncycles)
    g.run(ncycles)
    return o, g


@pytest.mark.parametrize("name", sorted(Q4_CASES))
def test_fp64_bitwise_vs_cr_oracle_and_golden(name, gpu_lib):
    c = Q4_CASES[name]
    print(f"case: {name}, {c['cycles']} cycles, fp64", flush=True)
    m = build_q4_case(c)
    o, g = run_pair(m, c["cycles"])
    gold = np.load(os.path.join(HERE, "golden", name + ".npz"))
    ts = g.time_state()
    assert ts.ncycle == o.ts.ncycle and ts.mstop == 0
    assert (ts.tt, ts.dt2) == (o.ts.tt, o.ts.dt2)
    x, v, vr, d = g.nodes6()
    report("d vs oracle(cr)", d, o.d)
    report("d vs reference golden", d, gold["d"])
    assert rel_err(d, gold["d"]) <= LIBM_SPREAD_TOL
    assert rel_err(v, gold["v"]) <= LIBM_SPREAD_TOL
    assert np.array_equal(d, o.d)
    assert np.array_equal(v, o.v)
    assert np.array_equal(vr, o.vr)
    if m.q4 is not None:
        st = g.q4_state()
        for k in ("sig", "forc", "mom", "thk", "pla", "offg", "eint", "hourg"):
            assert np.array_equal(st[k], o.q[k]), k
    if m.hexp is not None:
        hs = g.hex_state()
        assert np.array_equal(hs["sig"], o.hsig)
        assert np.array_equal(hs["pla"], o.hpla)
    e_gpu, e_ora = g.energy(), o.energy()
    assert abs(e_gpu.sum() - e_ora.sum()) <= 1e-13 * abs(e_ora.sum())
    # total energy against the reference's state: internal (shell EINT, hex EINT*VOL0) + kinetic
    ie = float(np.sum(gold["q_eint"])) if "q_eint" in gold else 0.0
    if m.hexp is not None:
        ie += float(np.sum(gold["h_eint"] * m.hexp.vol0))
    ke = 0.5 * float(np.sum(m.ms * np.sum(gold["v"] ** 2, axis=1)) + np.sum(m.inr * np.sum(gold["vr"] ** 2, axis=1)))
    e_ref = ie + ke
    print(f"total energy rel err vs reference golden {abs(e_gpu.sum() - e_ref) / abs(e_ref):.3e}", flush=True)
    assert abs(e_gpu.sum() - e_ref) <= 1e-12 * abs(e_ref)
    g.close()


def test_split_precision_tube_error_bound(gpu_lib):
    print("case: tube 24x60, 400 cycles, split", flush=True)
    m = M.square_tube(width=24.0, length=60.0)
    o, g = run_pair(m, 400, precision="split", math="libm")
    x, v, vr, d = g.nodes6()
    err = rel_err(d, o.d)
    eerr = abs(g.energy().sum() - o.energy().sum()) / o.energy().sum()
    print(f"split: displacement rel err {err:.3e}, energy rel err {eerr:.3e}")
    assert err <= 1e-4
    assert eerr <= 1e-4
    g.close()


def test_tube_boundary_conditions(gpu_lib):
    print("case: tube 8x20, 200 cycles, BCs", flush=True)
    m = M.square_tube(width=8.0, length=20.0)
    g = E.GpuModel(m)
    g.run(200)
    x, v, vr, d = g.nodes6()
    top = m.x[:, 2] == 20.0
    base = m.x[:, 2] == 0.0
    assert np.all(v[top, 2] == -1.5e4)
    assert np.all(v[base] == 0.0) and np.all(vr[base] == 0.0) and np.all(d[base] == 0.0)
    g.close()


def test_cforc3_group_dropin_matches_oracle(gpu_lib):
    """One group of 128 shells through or_cforc3_group with the reference layouts."""
    print("case: or_cforc3_group, nel=128, distorted plate", flush=True)
    rng = np.random.default_rng(3)
    x, conn = M.flat_plate(16, 8, h=1.0, thick=1.0)
    x = x + 0.1 * rng.uniform(-1, 1, x.shape)
    m = M.build_model(x, q4_conn=conn, q4_mat=M.STEEL, q4_prop=M.ShellProp(thick=1.0, ipla=1))
    v = 3e3 * rng.standard_normal((m.nn, 3))
    vr = 300.0 * rng.standard_normal((m.nn, 3))
    st = m.q4.initial_state()
    st["sig"][:] = 150.0 * rng.standard_normal(st["sig"].shape)
    st["pla"][:] = np.abs(1e-3 * rng.standard_normal(st["pla"].shape))
    o = O.ModelOracle(m, math="cr")
    o.v[:] = v
    o.vr[:] = vr
    for k in ("sig", "pla"):
        o.q[k][...] = st[k]
    dt1 = 1.0e-7
    dt2t = np.array([1.0e6])
    O.lib("cr").or_q4_forces(O.C.c_int64(0), O.C.c_int64(m.ne_q4), o.qconn.ctypes.data_as(O._I32), O._pd(o.x),
                         O._pd(o.v), O._pd(o.vr), O.C.byref(o.c.q), o.qiadc.ctypes.data_as(O._I64),
                         O._pd(o.fsky), O.C.byref(o.qp), O.C.c_double(dt1), O._pd(dt2t))
    ixc = np.zeros((m.ne_q4, 7), np.int32)
    ixc[:, 1:5] = m.q4.conn + 1
    iadc = (m.q4.iadc + 1).astype(np.int32)
    gb = {k: np.ascontiguousarray(st[k]) for k in ("offg", "thk", "forc", "mom", "gstr", "hourg", "eint", "epsd")}
    gb["thke"] = np.ascontiguousarray(m.q4.thke)
    lb = dict(sig=np.ascontiguousarray(st["sig"]), pla=np.ascontiguousarray(st["pla"]),
              epsd=np.ascontiguousarray(st["epsd_ip"]))
    fsky = np.zeros((m.lsky, 8))
    dt2t_g = E.cforc3_group(E.q4_params(m), x, v, vr, ixc, gb, lb, fsky, iadc, nft=0, nel=m.ne_q4, dt1=dt1,
                            dt2t=1e6)
    report("fsky", fsky, o.fsky)
    assert dt2t_g == dt2t[0]
    assert np.array_equal(fsky, o.fsky)
    assert np.array_equal(lb["sig"], o.q["sig"])
    assert np.array_equal(gb["forc"], o.q["forc"])


@pytest.mark.parametrize("name", ["mixed_c120", "tube_8x20_c150"])
def test_ecrit_output_cycle_matches_oracle(name, gpu_lib):
    """or_model_run_output: ECRIT's energies of the cycle (fixed-order GPU sums against the
    oracle's sequential sums) and an unchanged trajectory."""
    c = Q4_CASES[name]
    print(f"case: {name}, ECRIT output on cycle {c['cycles']}", flush=True)
    m = build_q4_case(c)
    o = O.ModelOracle(m, nthreads=8, math="cr")
    g = E.GpuModel(m)
    o.cycle(c["cycles"] - 1)
    g.run(c["cycles"] - 1)
    eo = o.cycle_output()
    eg = g.run_output()
    print(f"ECRIT energies gpu {eg}, oracle {eo}")
    assert np.all(np.abs(eg - eo) <= 1e-13 * np.abs(eo).max())
    x, v, vr, d = g.nodes6()
    assert np.array_equal(d, o.d) and np.array_equal(v, o.v) and np.array_equal(vr, o.vr)
    g.run(5)
    o.cycle(5)
    assert np.array_equal(g.nodes6()[3], o.d)
    g.close()
