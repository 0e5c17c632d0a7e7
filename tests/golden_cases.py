"""Definitions of the golden cases (shared by tests/golden/make_golden.py and the tests)."""
import numpy as np

from openradioss_amd import model as M

CASES = {
    # elastic-plastic cantilever bar (configs[0] family, reduced), 10 m/s
    "bar_6x5x12_c120": dict(kind="bar", nx=6, ny=5, nz=12, v0=1.0e4, cycles=120),
    # strongly plastic impact (pow branch of M2LAW every cycle), 60 m/s
    "impact_4x4x16_c120": dict(kind="bar", nx=4, ny=4, nz=16, v0=6.0e4, cycles=120),
    # distorted free block with random velocities and linear hardening (CN=1 branch)
    "distorted_linhard_c80": dict(kind="distorted", n=4, seed=9, cycles=80),
}


def build_case(c):
    if c["kind"] == "bar":
        return M.cantilever_bar(c["nx"], c["ny"], c["nz"], v0=c["v0"])
    rng = np.random.default_rng(c["seed"])
    x, conn = M.hex_block(c["n"], c["n"], c["n"], 1.0)
    x = x + 0.08 * rng.uniform(-1.0, 1.0, x.shape)
    mat = M.Law2(rho=7.85e-9, young=210000.0, nu=0.3, a=300.0, b=800.0, n=1.0)
    v = 3.0e3 * rng.standard_normal(x.shape)
    return M.build_hex_model(x, conn, mat, v0=v)
