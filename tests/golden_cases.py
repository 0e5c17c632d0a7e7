"""Definitions of the golden cases (shared by tests/golden/make_golden.py and the tests)."""
import numpy as np

from openradioss_amd import model as M

CASES = {
    # elastic-plastic cantilever bar (a reduced 20x20x80 bar), 10 m/s
    "bar_6x5x12_c120": dict(kind="bar", nx=6, ny=5, nz=12, v0=1.0e4, cycles=120),
    # strongly plastic impact (pow branch of M2LAW every cycle), 60 m/s
    "impact_4x4x16_c120": dict(kind="bar", nx=4, ny=4, nz=16, v0=6.0e4, cycles=120),
    # distorted free block with random velocities and linear hardening (CN=1 branch)
    "distorted_linhard_c80": dict(kind="distorted", n=4, seed=9, cycles=80),
    # plastic impact with Johnson-Cook rate hardening (CC != 0, ISRATE = 1): total strain rate
    # (VP = 2) filtered at 100 kHz (m2law.F:227-245, mstrain_rate.F)
    "impact_rate_c120": dict(kind="bar", nx=4, ny=4, nz=16, v0=6.0e4, cycles=120,
                             mat=dict(c=0.05, eps0=1.0, fcut=1.0e5)),
    # the same with the deviatoric strain rate (VP = 3, IDEV = 1) and no filter
    "impact_vp3_c120": dict(kind="bar", nx=4, ny=4, nz=16, v0=6.0e4, cycles=120,
                            mat=dict(c=0.05, eps0=1.0, vp=3)),
    # a block crushed at 1 km/s until an element's Jacobian turns negative: SCHKJABT3 deletes it
    # and sets MSTOP, which ends the run after that cycle (sderi3.F:147-153, schkjabt3.F)
    "inverted_mstop": dict(kind="crush", n=2, v_top=-1.0e6, cycles=400),
}


def build_case(c):
    if c["kind"] == "bar":
        mat = M.STEEL if "mat" not in c else M.Law2(rho=7.85e-9, young=210000.0, nu=0.3, a=350.0, b=275.0,
                                                     n=0.36, **c["mat"])
        return M.cantilever_bar(c["nx"], c["ny"], c["nz"], v0=c["v0"], mat=mat)
    if c["kind"] == "crush":
        n = c["n"]
        x, conn = M.hex_block(n, n, n, 1.0)
        top = np.nonzero(x[:, 2] == float(n))[0]
        bcs = np.where(x[:, 2] == 0.0, 7, 0).astype(np.int32)
        # one top corner driven down faster than the rest, so that one element inverts first
        vals = np.full(top.size, c["v_top"])
        vals[0] = 3.0 * c["v_top"]
        imp = (top.astype(np.int32), np.full(top.size, 2, np.int32), vals)
        return M.build_hex_model(x, conn, M.STEEL, bcs=bcs, imp=imp)
    rng = np.random.default_rng(c["seed"])
    x, conn = M.hex_block(c["n"], c["n"], c["n"], 1.0)
    x = x + 0.08 * rng.uniform(-1.0, 1.0, x.shape)
    mat = M.Law2(rho=7.85e-9, young=210000.0, nu=0.3, a=300.0, b=800.0, n=1.0)
    v = 3.0e3 * rng.standard_normal(x.shape)
    return M.build_hex_model(x, conn, mat, v0=v)


# ---------------------------------------------------------------- Q4 shells and mixed models
Q4_CASES = {
    # axial crush of a small square tube (a reduced 60x300 tube): Newton return, thickness change
    "tube_8x20_c150": dict(kind="tube", width=8.0, length=20.0, ipla=1, ithk=1, v_top=-1.5e4, cycles=150),
    # strain-rate hardening with filtering and progressive rupture at 1% plastic strain
    "tube_rupture_c150": dict(kind="tube_rupture", width=8.0, length=20.0, v_top=-4.0e4, cycles=150),
    # distorted plate with random translational and rotational velocities, one triangle,
    # radial-return plasticity (IPLA=0) and constant thickness (ITHK=0)
    "plate_bend_c120": dict(kind="plate", seed=5, ipla=0, ithk=0, cycles=120),
    # the same plate with a shell time-step floor DTMIN1(3) just above its smallest element time
    # step: CDT3 with IDTMIN(3) = 2 deletes the elements whose time step falls below it
    # (cdt3.F:147-165), and the run continues without them
    "plate_dtmin_c120": dict(kind="plate", seed=5, ipla=0, ithk=0, cycles=120, dtmin=1.4e-7),
    # hex block with a shell plate tied on its top face (a reduced 160^3 block with its plate)
    "mixed_c120": dict(kind="mixed", seed=7, cycles=120),
}


def build_q4_case(c):
    if c["kind"] == "tube":
        prop = M.ShellProp(thick=1.5, ipla=c["ipla"], ithk=c["ithk"])
        return M.square_tube(width=c["width"], length=c["length"], prop=prop, v_top=c["v_top"])
    if c["kind"] == "tube_rupture":
        mat = M.Law2(rho=2.7e-9, young=70000.0, nu=0.33, a=200.0, b=150.0, n=0.3, c=0.02, eps0=1.0,
                     fcut=1000.0, epsmax=0.01)
        return M.square_tube(width=c["width"], length=c["length"], mat=mat, v_top=c["v_top"])
    rng = np.random.default_rng(c["seed"])
    if c["kind"] == "plate":
        x, conn = M.flat_plate(6, 5, h=1.0, thick=1.2)
        x = x + 0.15 * rng.uniform(-1, 1, x.shape)
        conn[3, 3] = conn[3, 2]
        nn = x.shape[0]
        v = 2.0e3 * rng.standard_normal((nn, 3))
        vr = 500.0 * rng.standard_normal((nn, 3))
        prop = M.ShellProp(thick=1.2, ipla=c["ipla"], ithk=c["ithk"], dtmin=c.get("dtmin", 0.0))
        return M.build_model(x, q4_conn=conn, q4_mat=M.STEEL, q4_prop=prop, v0=v, vr0=vr)
    # mixed: 3x3x3 hex block, 5x5 plate on top sharing the block's top nodes
    xh, ch = M.hex_block(3, 3, 3, 1.0)
    xp, cp = M.flat_plate(5, 5, h=1.0, z=3.0, origin=(-1.0, -1.0))
    return M.tie_q4_on_hex(xh, ch, xp, cp, rng)
