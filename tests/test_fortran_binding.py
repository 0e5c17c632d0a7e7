"""The Fortran bind(c) module (openradioss_amd/fortran/or_gpu_mod.F90): a Fortran program calls
OR_SFORC3_GROUP with the Engine's arrays; its results must equal the same call made from Python."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from openradioss_amd import engine as E
from openradioss_amd import lib as L
from openradioss_amd import model as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "openradioss_amd", "lib", "or_fortran_check")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran check not built (no amdflang)")


def group_case():
    """One group of 64 distorted hexes with random velocities and a plastic state."""
    rng = np.random.default_rng(21)
    x, conn = M.hex_block(4, 4, 4, 1.0)
    x = x + 0.1 * rng.uniform(-1, 1, x.shape)
    m = M.build_hex_model(x, conn, M.STEEL)
    v = 2e3 * rng.standard_normal((m.nn, 3))
    ixs = np.zeros((m.ne, 11), np.int32)
    ixs[:, 1:9] = m.conn + 1
    iads = (m.iads + 1).astype(np.int32)
    gb = dict(sig=150.0 * rng.standard_normal((6, m.ne)), eint=np.zeros(m.ne),
              rho=np.full(m.ne, m.params()["rho0"]), vol=m.vol0.copy(),
              pla=np.abs(1e-3 * rng.standard_normal(m.ne)), epsd=np.zeros(m.ne), off=np.ones(m.ne),
              qvis=np.zeros(m.ne))
    return m, x, v, ixs, iads, gb, 1.1e-7


def write_input(path, m, x, v, ixs, iads, gb, dt1):
    hp = E.hex_params(m)
    with open(path, "wb") as f:
        f.write(np.array([m.nn, m.ne, m.lsky], np.int32).tobytes())
        f.write(bytes(hp))
        f.write(np.ascontiguousarray(x, np.float64).tobytes())      # X(3,NUMNOD)
        f.write(np.ascontiguousarray(v, np.float64).tobytes())
        f.write(ixs.tobytes())                                       # IXS(11,NEL)
        f.write(iads.tobytes())                                      # IADS(8,NEL)
        for k in ("sig", "eint", "rho", "vol", "pla", "epsd", "off", "qvis"):
            f.write(np.ascontiguousarray(gb[k], np.float64).tobytes())
        f.write(np.array([dt1]).tobytes())


def read_output(path, m):
    b = open(path, "rb").read()
    st, mstop = np.frombuffer(b[:8], np.int32)
    off = 8
    dt2t = np.frombuffer(b[off:off + 8], np.float64)[0]
    off += 8
    fsky = np.frombuffer(b[off:off + 64 * m.lsky], np.float64).reshape(m.lsky, 8)
    off += 64 * m.lsky
    out = {}
    for k, n in (("sig", 6 * m.ne), ("eint", m.ne), ("rho", m.ne), ("pla", m.ne), ("epsd", m.ne),
                 ("off", m.ne), ("qvis", m.ne)):
        out[k] = np.frombuffer(b[off:off + 8 * n], np.float64)
        off += 8 * n
    assert off == len(b)
    return int(st), int(mstop), dt2t, fsky, out


@needs_exe
def test_fortran_program_links_and_reports_version():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert f"or_version {L.load().or_version()}" in r.stdout


@needs_exe
def test_fortran_call_returns_status_without_a_device(tmp_path):
    if L.device_count() > 0:
        pytest.skip("a device is present")
    m, x, v, ixs, iads, gb, dt1 = group_case()
    write_input(tmp_path / "in.bin", m, x, v, ixs, iads, gb, dt1)
    r = subprocess.run([EXE, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0                       # the error comes back as a status, not an abort
    st, *_ = read_output(tmp_path / "out.bin", m)
    assert st != L.OR_OK
    assert "or_sforc3_group status" in r.stdout


@needs_exe
@pytest.mark.gpu
def test_fortran_group_call_equals_python_call(tmp_path, gpu_lib):
    print("case: Fortran OR_SFORC3_GROUP, 64 hexes", flush=True)
    m, x, v, ixs, iads, gb, dt1 = group_case()
    write_input(tmp_path / "in.bin", m, x, v, ixs, iads, gb, dt1)
    r = subprocess.run([EXE, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    st, mstop, dt2t_f, fsky_f, out_f = read_output(tmp_path / "out.bin", m)
    assert st == L.OR_OK and mstop == 0
    gp = {k: np.ascontiguousarray(a, np.float64).copy() for k, a in gb.items()}
    fsky = np.zeros((m.lsky, 8))
    dt2t, ms = E.sforc3_group(E.hex_params(m), x, v, ixs, gp, fsky, iads, nft=0, nel=m.ne, dt1=dt1, dt2t=1e6)
    assert dt2t_f == dt2t and ms == 0
    assert np.array_equal(fsky_f, fsky)
    for k in ("sig", "eint", "rho", "pla", "epsd", "off", "qvis"):
        assert np.array_equal(out_f[k], gp[k].ravel()), k


# ---------------------------------------------------------------------------------------------
# Engine glue: eligibility predicates and parameter packing (or_gpu_mod), the shell group, the
# nodal half and the resident model called from Fortran.

def time_state_of(b: bytes) -> L.OrTimeState:
    """or_time_state_t as a Fortran stream write leaves it: its components without the C
    struct's trailing padding."""
    return L.OrTimeState.from_buffer_copy(b + bytes(C.sizeof(L.OrTimeState) - len(b)))


def hex_engine_words(m, nel, nft=0):
    """The words an Engine holds for a law-2 solid group: IPARG column, mat_param iparam / rho /
    bulk / shear / uparam, GEO column of /PROP/SOLID."""
    p = m.params()
    iparg = np.zeros(100, np.int32)
    iparg[[0, 1, 2, 4, 8, 13, 22]] = [2, nel, nft, 1, 4, 1, 1]   # MLW, NEL, NFT, ITY, ISMSTR, JLAG, JHBE
    iparam = np.array([p["iform"], p["icc"], p["vp"], p["israte"]], np.int32)
    mat = np.array([p["rho0"], p["bulk"], p["shear"]])
    uparam = np.array([p[k] for k in ("ca", "cb", "cn", "epmx", "sigmx", "cc", "epdr", "fisokin", "asrate",
                                      "m_exp")])
    geo = np.zeros(40)
    geo[12:17] = [p["qh"], p["qa"], p["qb"], p["vns1"], p["vns2"]]
    return iparg, iparam, mat, uparam, geo, np.array([p["dtfac"], p["dtmx"]])


def q4_engine_words(m, nel, nft=0):
    """The words an Engine holds for a law-2 shell group: IPARG column, PM words, the law's
    iparam / uparam, GEO column of /PROP/SHELL."""
    p = m.q4.params()
    iparg = np.zeros(100, np.int32)
    iparg[[0, 1, 2, 4, 5, 8, 13, 22, 27, 28, 37]] = [2, nel, nft, 3, p["npt"], 2, 1, 1, p["ithk"], p["ipla"], 1]
    iparam = np.array([p["iform"], p["icc"], p["vp"], p["israte"]], np.int32)
    pm = np.zeros(30)
    pm[[0, 19, 20, 21, 26]] = [p["rho"], p["young"], p["nu"], p["shear"], p["ssp"]]
    uparam = np.array([p["ca"], p["cb"], p["cn"], p["epmx"], p["ymax"], p["cc"], p["epdr"], 0.0, p["asrate"],
                       p["m_exp"]])
    geo = np.zeros(40)
    geo[[12, 13, 14, 36, 37]] = [p["hm"], p["hf"], p["hr"], p["ishear"], p["fsh"]]
    return iparg, iparam, pm, uparam, geo, np.array([p["dtfac"], p["dtmin"]])


def q4_group_case():
    rng = np.random.default_rng(3)
    x, conn = M.flat_plate(16, 8, h=1.0, thick=1.0)
    x = x + 0.1 * rng.uniform(-1, 1, x.shape)
    m = M.build_model(x, q4_conn=conn, q4_mat=M.STEEL, q4_prop=M.ShellProp(thick=1.0, ipla=1))
    v = 3e3 * rng.standard_normal((m.nn, 3))
    vr = 300.0 * rng.standard_normal((m.nn, 3))
    st = m.q4.initial_state()
    st["sig"][:] = 150.0 * rng.standard_normal(st["sig"].shape)
    st["pla"][:] = np.abs(1e-3 * rng.standard_normal(st["pla"].shape))
    ixc = np.zeros((m.ne_q4, 7), np.int32)
    ixc[:, 1:5] = m.q4.conn + 1
    iadc = (m.q4.iadc + 1).astype(np.int32)
    gb = {k: np.ascontiguousarray(st[k]) for k in ("offg", "thk", "forc", "mom", "gstr", "hourg", "eint", "epsd")}
    gb["thke"] = np.ascontiguousarray(m.q4.thke)
    lb = dict(sig=np.ascontiguousarray(st["sig"]), pla=np.ascontiguousarray(st["pla"]),
              epsd=np.ascontiguousarray(st["epsd_ip"]))
    return m, x, v, vr, ixc, iadc, gb, lb, 1.0e-7


def write_q4group(path, m, x, v, vr, ixc, iadc, gb, lb, dt1, iparg=None):
    words = q4_engine_words(m, m.ne_q4)
    if iparg is not None:
        words = (iparg,) + words[1:]
    with open(path, "wb") as f:
        f.write(np.array([m.nn, m.ne_q4, m.lsky], np.int32).tobytes())
        for a in words:
            f.write(np.ascontiguousarray(a).tobytes())
        for a in (x, v, vr):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(ixc.tobytes())
        f.write(iadc.tobytes())
        for k in ("offg", "thk", "thke", "forc", "mom", "gstr", "hourg", "eint", "epsd"):
            f.write(np.ascontiguousarray(gb[k], np.float64).tobytes())
        for k in ("sig", "pla", "epsd"):
            f.write(np.ascontiguousarray(lb[k], np.float64).tobytes())
        f.write(np.array([dt1]).tobytes())


def read_q4group(path, m):
    b = open(path, "rb").read()
    st, ok = np.frombuffer(b[:8], np.int32)
    off = 8
    out = {"dt2t": np.frombuffer(b[off:off + 8], np.float64)[0]}
    off += 8
    npt, ne = m.q4.prop.npt, m.ne_q4
    for k, n in (("fsky", 8 * m.lsky), ("offg", ne), ("thk", ne), ("forc", 5 * ne), ("mom", 3 * ne),
                 ("gstr", 8 * ne), ("hourg", 5 * ne), ("eint", 2 * ne), ("epsd", ne), ("sig", npt * 5 * ne),
                 ("pla", npt * ne), ("epsd_ip", npt * ne)):
        out[k] = np.frombuffer(b[off:off + 8 * n], np.float64)
        off += 8 * n
    assert off == len(b)
    return int(st), int(ok), out


@needs_exe
def test_fortran_eligibility_predicates(tmp_path):
    """Groups off the path are refused by the Fortran predicate before any call (no device needed):
    another element type, another law, a failure model, a sub-stack."""
    case = q4_group_case()
    m = case[0]
    base = q4_engine_words(m, m.ne_q4)[0]
    for idx, val in ((4, 1), (0, 36), (42, 1), (70, 1), (8, 1), (22, 11)):
        iparg = base.copy()
        iparg[idx] = val
        write_q4group(tmp_path / "in.bin", *case, iparg=iparg)
        r = subprocess.run([EXE, "q4group", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        st, ok, _ = read_q4group(tmp_path / "out.bin", m)
        assert ok == 0 and st == L.OR_ERR_ARG, (idx, val)


@needs_exe
@pytest.mark.gpu
def test_fortran_shell_group_through_the_engine_glue_equals_python_call(tmp_path, gpu_lib):
    print("case: Fortran eligibility + OR_PACK_Q4_PARAMS + OR_CFORC3_GROUP, 128 shells", flush=True)
    m, x, v, vr, ixc, iadc, gb, lb, dt1 = q4_group_case()
    write_q4group(tmp_path / "in.bin", m, x, v, vr, ixc, iadc, gb, lb, dt1)
    r = subprocess.run([EXE, "q4group", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    st, ok, out = read_q4group(tmp_path / "out.bin", m)
    assert ok == 1 and st == L.OR_OK
    gp = {k: a.copy() for k, a in gb.items()}
    lp = {k: a.copy() for k, a in lb.items()}
    fsky = np.zeros((m.lsky, 8))
    dt2t = E.cforc3_group(E.q4_params(m), x, v, vr, ixc, gp, lp, fsky, iadc, nft=0, nel=m.ne_q4, dt1=dt1, dt2t=1e6)
    assert out["dt2t"] == dt2t
    assert np.array_equal(out["fsky"], fsky.ravel())
    for k in ("offg", "thk", "forc", "mom", "gstr", "hourg", "eint", "epsd"):
        assert np.array_equal(out[k], gp[k].ravel()), k
    assert np.array_equal(out["sig"], lp["sig"].ravel()) and np.array_equal(out["pla"], lp["pla"].ravel())


@needs_exe
@pytest.mark.gpu
def test_fortran_solid_group_through_the_engine_glue_equals_python_call(tmp_path, gpu_lib):
    print("case: Fortran eligibility + OR_PACK_HEX_PARAMS + OR_SFORC3_GROUP, 64 hexes", flush=True)
    m, x, v, ixs, iads, gb, dt1 = group_case()
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([m.nn, m.ne, m.lsky], np.int32).tobytes())
        for a in hex_engine_words(m, m.ne):
            f.write(np.ascontiguousarray(a).tobytes())
        f.write(np.ascontiguousarray(x, np.float64).tobytes())
        f.write(np.ascontiguousarray(v, np.float64).tobytes())
        f.write(ixs.tobytes())
        f.write(iads.tobytes())
        for k in ("sig", "eint", "rho", "vol", "pla", "epsd", "off", "qvis"):
            f.write(np.ascontiguousarray(gb[k], np.float64).tobytes())
        f.write(np.array([dt1]).tobytes())
    r = subprocess.run([EXE, "hexgroup", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    b = open(tmp_path / "out.bin", "rb").read()
    st, ok, mstop = np.frombuffer(b[:12], np.int32)
    dt2t_f = np.frombuffer(b[12:20], np.float64)[0]
    fsky_f = np.frombuffer(b[20:20 + 64 * m.lsky], np.float64)
    assert st == L.OR_OK and ok == 1 and mstop == 0
    gp = {k: np.ascontiguousarray(a, np.float64).copy() for k, a in gb.items()}
    fsky = np.zeros((m.lsky, 8))
    dt2t, _ = E.sforc3_group(E.hex_params(m), x, v, ixs, gp, fsky, iads, nft=0, nel=m.ne, dt1=dt1, dt2t=1e6)
    assert dt2t_f == dt2t and np.array_equal(fsky_f, fsky.ravel())
    sig_f = np.frombuffer(b[20 + 64 * m.lsky:20 + 64 * m.lsky + 48 * m.ne], np.float64)
    assert np.array_equal(sig_f, gp["sig"].ravel())


@needs_exe
@pytest.mark.gpu
def test_fortran_nodal_update_equals_python_call(tmp_path, gpu_lib):
    print("case: Fortran OR_NODAL_UPDATE on the mixed block + plate after 10 cycles", flush=True)
    import oracle as O
    from golden_cases import Q4_CASES, build_q4_case
    m = build_q4_case(Q4_CASES["mixed_c120"])
    o = O.ModelOracle(m, nthreads=2, math="cr")
    o.cycle(10)
    pre = {k: getattr(o, k).copy() for k in ("x", "v", "vr", "d")}
    ts0 = (o.ts.tt, o.ts.dt1, o.ts.dt2, o.ts.dt12, o.ts.dt2old, o.ts.ncycle)
    o.cycle(1)
    fsky = o.fsky.copy()
    ts = L.OrTimeState()
    ts.tt, ts.dt1, ts.dt2, ts.dt12, ts.dt2old, ts.ncycle = ts0
    adsky = np.asarray(m.adsky, np.int64) + 1
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([m.nn, m.lsky], np.int64).tobytes())
        f.write(adsky.tobytes())
        f.write(np.ascontiguousarray(fsky).tobytes())
        f.write(np.ascontiguousarray(m.ms, np.float64).tobytes())
        f.write(np.ascontiguousarray(m.inr, np.float64).tobytes())
        f.write(np.ascontiguousarray(m.code, np.int32).tobytes())
        for a in (m.vimp, pre["x"], pre["v"], pre["vr"], pre["d"]):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(np.array([o.ts.dt2, m.dtmx]).tobytes())
        f.write(bytes(ts))
    r = subprocess.run([EXE, "nodal", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    b = open(tmp_path / "out.bin", "rb").read()
    assert np.frombuffer(b[:4], np.int32)[0] == L.OR_OK
    n3 = 24 * m.nn
    xf, vf, vrf, df = (np.frombuffer(b[4 + k * n3:4 + (k + 1) * n3], np.float64).reshape(m.nn, 3) for k in range(4))
    tsf = time_state_of(b[4 + 4 * n3 + 16 * m.nn:])
    x, v, vr, d = (pre[k].copy() for k in ("x", "v", "vr", "d"))
    E.nodal_update(adsky, fsky, m.ms, x, v, d, ts, dt2t=o.ts.dt2, dtmx=m.dtmx, inr=m.inr, vr=vr, code=m.code,
                   vimp=m.vimp)
    assert np.array_equal(xf, x) and np.array_equal(vf, v) and np.array_equal(vrf, vr) and np.array_equal(df, d)
    assert (tsf.tt, tsf.dt2, tsf.ncycle) == (ts.tt, ts.dt2, ts.ncycle)
    assert np.array_equal(xf, o.x)


@needs_exe
@pytest.mark.gpu
def test_fortran_resident_model_equals_python_model(tmp_path, gpu_lib):
    print("case: Fortran resident model (create, upload, time-history setup, 59 cycles and an output cycle, "
          "download), mixed block + plate", flush=True)
    from golden_cases import Q4_CASES, build_q4_case
    m = build_q4_case(Q4_CASES["mixed_c120"])
    hp, qp = E.hex_params(m), E.q4_params(m)
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([m.nn, m.ne_hex, m.ne_q4], np.int64).tobytes())
        f.write(np.array([m.q4.prop.npt, L.OR_PREC_FP64, 60], np.int32).tobytes())
        f.write(bytes(hp))
        f.write(bytes(qp))
        for a in (m.x, m.v, m.vr, m.ms, m.inr):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(np.ascontiguousarray(m.code, np.int32).tobytes())
        f.write(np.ascontiguousarray(m.vimp, np.float64).tobytes())
        f.write(np.ascontiguousarray(m.hexp.conn, np.int32).tobytes())
        f.write(np.ascontiguousarray(m.hexp.vol0, np.float64).tobytes())
        f.write(np.full(m.ne_hex, m.hexp.params(m.dtmx)["rho0"]).tobytes())
        f.write(np.ascontiguousarray(m.q4.conn, np.int32).tobytes())
        f.write(np.ascontiguousarray(m.q4.thke, np.float64).tobytes())
        f.write(np.ascontiguousarray(m.adsky, np.int64).tobytes())
        f.write(np.ascontiguousarray(m.hexp.iads, np.int64).tobytes())
        f.write(np.ascontiguousarray(m.q4.iadc, np.int64).tobytes())
    r = subprocess.run([EXE, "model", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    b = open(tmp_path / "out.bin", "rb").read()
    assert np.frombuffer(b[:4], np.int32)[0] == L.OR_OK, r.stdout
    n3 = 24 * m.nn
    xf, vf, vrf, df = (np.frombuffer(b[4 + k * n3:4 + (k + 1) * n3], np.float64).reshape(m.nn, 3) for k in range(4))
    tsf = time_state_of(b[4 + 4 * n3:4 + 4 * n3 + 52])
    o = 4 + 4 * n3 + 52
    ecf = L.OrEcrit.from_buffer_copy(b[o:o + C.sizeof(L.OrEcrit)])
    o += C.sizeof(L.OrEcrit)
    thf = np.frombuffer(b[o:], np.float64)
    g = E.GpuModel(m)
    g.th_setup([0], [0], [0])
    g.run(59)
    ec = g.run_th()
    x, v, vr, d = g.nodes6()
    ts = g.time_state()
    g.close()
    # the time-history record of the output cycle, as OR_MODEL_RUN_TH returned it to Fortran
    npt = m.q4.prop.npt
    assert thf.size == 4 + 18 + 13 + 12 + npt
    assert np.array_equal(thf[:4], [ec["xmomt"], ec["ymomt"], ec["zmomt"], ec["ehour"]])
    assert np.array_equal(thf[4:22], ec["nodes"][0]) and np.array_equal(thf[22:35], ec["hex"][0])
    assert np.array_equal(thf[35:], ec["q4"][0]) and ec["ehour"] > 0
    assert tsf.ncycle == ts.ncycle == 60 and tsf.tt == ts.tt
    assert np.array_equal(xf, x) and np.array_equal(vf, v) and np.array_equal(vrf, vr) and np.array_equal(df, d)
    # the listing record of the output cycle, as OR_MODEL_RUN_ECRIT returned it to Fortran
    assert (ecf.enint, ecf.encin, ecf.enrot, ecf.wfext) == (ec["enint"], ec["encin"], ec["enrot"], ec["wfext"])
    assert (ecf.itypts, ecf.nelts) == (ec["itypts"], ec["nelts"]) and ec["nelts"] >= 0
