"""GPU parity of the Q4 shell and mixed hex + Q4 paths: HIP kernels (through the C ABI)
against the C oracle, which is pinned bitwise to the reference's compiled CFORC3 chain."""
import numpy as np
import pytest

import oracle as O
from golden_cases import Q4_CASES, build_q4_case
from openradioss_amd import engine as E
from openradioss_amd import model as M

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def run_pair(model, ncycles, precision="fp64"):
    o = O.ModelOracle(model, nthreads=8)
    g = E.GpuModel(model, precision=precision)
    o.cycle(ncycles)
    g.run(ncycles)
    return o, g


# FP64 tolerance: the element arithmetic follows the reference operation for operation; the
# hardening law's exp/log (and the hex pow) come from the device math library, whose last bit
# can differ from glibc's, so trajectories agree to a relative 1e-11 over hundreds of cycles.
FP64_TOL = 1e-11


@pytest.mark.parametrize("name", sorted(Q4_CASES))
def test_fp64_matches_oracle_on_golden_cases(name, gpu_lib):
    c = Q4_CASES[name]
    print(f"case: {name}, {c['cycles']} cycles, fp64", flush=True)
    m = build_q4_case(c)
    o, g = run_pair(m, c["cycles"])
    ts = g.time_state()
    assert ts.ncycle == o.ts.ncycle and ts.mstop == 0
    assert ts.dt2 == pytest.approx(o.ts.dt2, rel=1e-12)
    x, v, vr, d = g.nodes6()
    assert rel_err(d, o.d) <= FP64_TOL
    assert rel_err(v, o.v) <= FP64_TOL
    if m.q4 is not None:
        st = g.q4_state()
        assert rel_err(vr, o.vr) <= FP64_TOL
        assert rel_err(st["sig"], o.q["sig"]) <= FP64_TOL
        assert rel_err(st["forc"], o.q["forc"]) <= FP64_TOL
        assert rel_err(st["thk"], o.q["thk"]) <= FP64_TOL
        assert np.array_equal(st["offg"] == 0.0, o.q["offg"] == 0.0)
        e_gpu, e_ora = g.energy(), o.energy()
        assert abs(e_gpu.sum() - e_ora.sum()) <= FP64_TOL * abs(e_ora.sum())
    print(f"bitwise-equal displacement fraction {np.mean(d == o.d):.4f}; max rel err {rel_err(d, o.d):.3e}")
    g.close()


def test_split_precision_tube_error_bound(gpu_lib):
    print("case: tube 24x60, 400 cycles, split", flush=True)
    m = M.square_tube(width=24.0, length=60.0)
    o, g = run_pair(m, 400, precision="split")
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
    o = O.ModelOracle(m)
    o.v[:] = v
    o.vr[:] = vr
    for k in ("sig", "pla"):
        o.q[k][...] = st[k]
    dt1 = 1.0e-7
    dt2t = np.array([1.0e6])
    O.lib().or_q4_forces(O.C.c_int64(0), O.C.c_int64(m.ne_q4), o.qconn.ctypes.data_as(O._I32), O._pd(o.x),
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
    assert dt2t_g == pytest.approx(dt2t[0], rel=1e-15)
    assert rel_err(fsky, o.fsky) <= 1e-12
    assert rel_err(lb["sig"], o.q["sig"]) <= 1e-12
    assert rel_err(gb["forc"], o.q["forc"]) <= 1e-12
    print(f"group: fsky bitwise fraction {np.mean(fsky == o.fsky):.4f}")
