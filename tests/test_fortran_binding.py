"""The Fortran bind(c) module (openradioss_amd/fortran/or_gpu_mod.F90): a Fortran program calls
OR_SFORC3_GROUP with the Engine's arrays; its results must equal the same call made from Python."""
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
