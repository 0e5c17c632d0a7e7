"""Pins of the C oracle (oracle/or_oracle.c) for the hex path by what the
mathematics fixes: closed forms, invariants and special cases.  CPU only."""
import numpy as np
import pytest

import oracle as O
from openradioss_amd import model as M


def small_block(nx=3, ny=2, nz=2, h=1.0, mat=M.STEEL, v=None, distort=0.0, seed=0):
    x, conn = M.hex_block(nx, ny, nz, h)
    if distort:
        rng = np.random.default_rng(seed)
        x = x + distort * h * rng.uniform(-1, 1, x.shape)
    return M.build_hex_model(x, conn, mat, v0=v)


def elem_forces(model, v, sig=None, dt1=1e-7, pla=None):
    """Run the SFORC3 chain of all elements once; return fsky and per-element outputs."""
    o = O.HexOracle(model)
    o.v[:] = v
    if sig is not None:
        o.sig[:] = sig
    if pla is not None:
        o.pla[:] = pla
    dt2t = np.array([1.0e6])
    ms = np.zeros(1, np.int32)
    O.lib().or_hex_forces(O.C.c_int64(0), O.C.c_int64(model.ne), O.C.c_int64(model.ne), O._p(o.conn), O._p(o.x),
                          O._p(o.v), O._p(o.sig), O._p(o.eint), O._p(o.rho), O._p(o.vol0), O._p(o.pla),
                          O._p(o.epsd), O._p(o.offg), O._p(o.qvis), None, O._p(o.iads), O._p(o.fsky),
                          O.C.byref(o.params), O.C.c_double(dt1), O._p(dt2t), O._p(ms))
    return o, dt2t[0], ms[0]


def test_constants():
    L = O.lib()
    assert L.or_const(0) == ((1.0 + 0.3) + 0.03) + 0.003     # ONEP333 arithmetic chain
    assert L.or_const(2) == float(np.float32(1.1))            # single-precision literal of resol.F:6352
    assert L.or_const(1) == 1.0 + 0.1


def test_volume_unit_cube_and_affine():
    x, conn = M.hex_block(1, 1, 1, 2.0)
    assert M.hex_volume(x[conn])[0] == 8.0
    rng = np.random.default_rng(3)
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    xa = x @ A.T
    assert M.hex_volume(xa[conn])[0] == pytest.approx(8.0 * np.linalg.det(A), rel=1e-13)


def test_rigid_translation_gives_zero_force():
    m = small_block(distort=0.1)
    v = np.tile([3.0e3, -2.0e3, 1.0e3], (m.nn, 1))
    o, dt2t, ms = elem_forces(m, v)
    assert ms == 0
    # a 1e3 mm/s shear across 1 mm gives forces of order 10 N; the rigid motion leaves only roundoff
    assert np.abs(o.fsky[:, :3]).max() < 1e-12
    assert np.abs(o.sig).max() < 1e-10


def test_linear_velocity_patch_constant_strain_rate():
    """A linear velocity field on a distorted hex gives the exact uniform D (SDERI3/SDEFO3 patch test)."""
    m = small_block(1, 1, 1, distort=0.15, seed=5)
    L = np.array([[1.0, 0.3, -0.2], [0.1, -0.5, 0.4], [0.0, 0.2, 0.7]])
    v = m.x @ L.T
    o, _, _ = elem_forces(m, v, dt1=1e-9)
    D = 0.5 * (L + L.T)
    trD = np.trace(D)
    G = m.params()["shear"]
    K = m.params()["bulk"]
    # after one step from zero stress: sigma = 2G dt (D - tr/3 I) - K*amu I, amu = rho/rho0-1 = 0 at t=0
    dt = 1e-9
    expect = 2 * G * dt * (D - trD / 3 * np.eye(3))
    s = o.sig[:, 0]
    got = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    # SSP and return are elastic here; RHO changes with the current volume (amu from this cycle's volume)
    amu = o.rho[0] / m.params()["rho0"] - 1.0
    expect = expect - K * amu * np.eye(3)
    assert np.allclose(got, expect, rtol=1e-9, atol=1e-12 * np.abs(expect).max())


def test_element_forces_self_equilibrated():
    rng = np.random.default_rng(7)
    m = small_block(distort=0.1, seed=2)
    v = 1e3 * rng.standard_normal((m.nn, 3))
    sig = 100.0 * rng.standard_normal((6, m.ne))
    o, _, _ = elem_forces(m, v, sig=np.ascontiguousarray(sig))
    for e in range(m.ne):
        rows = o.fsky[m.iads[e], :3]
        scale = np.abs(rows).max()
        assert np.abs(rows.sum(axis=0)).max() <= 1e-12 * scale


def test_cube_deltax_and_stable_dt_closed_form():
    """Uniform cube at rest: DELTAX = h and DTX = DTFAC*h/(c*(QB+sqrt(1+QB^2))), c^2=(4/3 G + K)/rho."""
    h = 1.0
    m = small_block(1, 1, 1, h=h)
    o, dt2t, _ = elem_forces(m, np.zeros((m.nn, 3)))
    p = m.params()
    c = np.sqrt((O.lib().or_const(0) * p["shear"] + p["bulk"]) / p["rho0"])
    qb = p["qb"]
    expect = p["dtfac"] * h / (c * qb + np.sqrt((c * qb) ** 2 + c * c))
    assert dt2t == pytest.approx(expect, rel=1e-15)
    # stiffness per node: STI/4 with STI = rho*V/dtx^2 (mqviscb.F: TRHO*TVOL)
    dtx = expect / p["dtfac"]
    assert o.fsky[0, 6] == pytest.approx(0.25 * p["rho0"] * h ** 3 / dtx ** 2, rel=1e-13)


def test_linear_hardening_radial_return():
    """CN=1: one radial return gives DPLA = (AJ2-AK)/(3G+CB) exactly for AJ2 > AK."""
    mat = M.Law2(rho=7.85e-9, young=210000.0, nu=0.3, a=350.0, b=1000.0, n=1.0)
    m = small_block(1, 1, 1, mat=mat)
    sig = np.zeros((6, 1))
    sig[3, 0] = 400.0  # pure shear above yield: AJ2 = sqrt(3)*400
    o, _, _ = elem_forces(m, np.zeros((m.nn, 3)), sig=sig, dt1=0.0)
    p = m.params()
    aj2 = np.sqrt(3.0) * 400.0
    dpla = (aj2 - p["ca"]) / (3 * p["shear"] + p["cb"])
    assert o.pla[0] == pytest.approx(dpla, rel=1e-14)
    ak = p["ca"] + p["cb"] * dpla
    assert o.sig[3, 0] == pytest.approx(400.0 * ak / aj2, rel=1e-14)


def test_thread_count_invariance_bitwise():
    m = M.cantilever_bar(6, 5, 12)
    a = O.HexOracle(m, nthreads=1)
    b = O.HexOracle(m, nthreads=8)
    a.cycle(30)
    b.cycle(30)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.v, b.v) and np.array_equal(a.sig, b.sig)
    assert a.ts.dt2 == b.ts.dt2 and a.ts.tt == b.ts.tt


def test_free_flight_energy_and_momentum():
    """Free block, smooth elastic vibration without hourglass viscosity: IE + KE stays within the
    O((w dt)^2) oscillation of the staggered kinetic energy; linear momentum is conserved to roundoff.
    (With rough fields the explicit update of viscous hourglass forces adds 0.5*lambda*dt*EHOU,
    so that case is not an invariant of the scheme.)"""
    x, conn = M.hex_block(8, 8, 8, 1.0)
    m = M.build_hex_model(x, conn, M.STEEL, prop=M.SolidProp(qh=0.0))
    m.v[:, 0] = 20.0 * np.sin(np.pi * m.x[:, 0] / 8.0)
    m.v[:, 1] = 5.0 * np.cos(np.pi * m.x[:, 2] / 8.0)
    v = m.v.copy()
    o = O.HexOracle(m)
    o.cycle(1)
    e0 = o.energy().sum()
    worst = 0.0
    for _ in range(6):
        o.cycle(100)
        worst = max(worst, abs(o.energy().sum() / e0 - 1.0))
    assert worst < 0.02
    p0 = (m.ms[:, None] * v).sum(axis=0)
    p1 = (m.ms[:, None] * o.v).sum(axis=0)
    assert np.abs(p1 - p0).max() <= 1e-12 * np.abs(m.ms[:, None] * v).sum()


def test_skyline_order_is_fillcne():
    """Rows of a node follow ascending element id (domdec2.F FILLCNE)."""
    m = small_block(3, 3, 3)
    for n in range(m.nn):
        rows = np.arange(m.adsky[n], m.adsky[n + 1])
        elems = [e for r in rows for e in np.argwhere(m.iads == r)[:, 0]]
        assert elems == sorted(elems)


def test_first_cycle_dt_of_the_bar():
    m = M.cantilever_bar()
    o = O.HexOracle(m, nthreads=8)
    o.cycle(1)
    p = m.params()
    c = np.sqrt((O.lib().or_const(0) * p["shear"] + p["bulk"]) / p["rho0"])
    expect = p["dtfac"] * 1.0 / (c * p["qb"] + np.sqrt((c * p["qb"]) ** 2 + c * c))
    assert o.ts.dt2 == pytest.approx(expect, rel=1e-14)
    assert o.ts.dt12 == 0.5 * o.ts.dt2
