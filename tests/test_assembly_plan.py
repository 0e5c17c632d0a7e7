"""The tiled assembly plan (openradioss_amd/csrc/or_assembly.h) on the CPU: the LDS prefixes and
rest rows of the element tiles, summed as the node kernel sums them, give the plain ASSPAR4 sum
of the skyline rows (asspar4.F:164-180) bit for bit; and on a structured block most rows are
summed in LDS.  The plan is simulated by tests/asm_check.cpp, built here with g++."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from openradioss_amd import model as M

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("asm") / "libasm_check.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                           os.path.join(HERE, "asm_check.cpp"), "-o", out])
    lib = C.CDLL(out)
    lib.asm_simulate.restype = C.c_int
    lib.asm_simulate.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64] + [C.c_void_p] * 10
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def family_rows(m, fam):
    """Rows of one family per node in skyline order: (rptr, element, corner)."""
    part = m.hexp if fam == "hex" else m.q4
    iad = part.iads if fam == "hex" else part.iadc
    ne, nk = iad.shape
    row_e = np.empty(m.lsky, np.int32)
    row_k = np.empty(m.lsky, np.int8)
    valid = np.zeros(m.lsky, bool)
    row_e[iad.reshape(-1)] = np.repeat(np.arange(ne, dtype=np.int32), nk)
    row_k[iad.reshape(-1)] = np.tile(np.arange(nk, dtype=np.int8), ne)
    valid[iad.reshape(-1)] = True
    node = np.repeat(np.arange(m.nn), np.diff(m.adsky))
    keep = valid
    counts = np.bincount(node[keep], minlength=m.nn)
    rptr = np.zeros(m.nn + 1, np.int64)
    np.cumsum(counts, out=rptr[1:])
    return rptr, np.ascontiguousarray(row_e[keep]), np.ascontiguousarray(row_k[keep]), counts


def simulate(sim, m, fam, vals, init, prefix_ok):
    part = m.hexp if fam == "hex" else m.q4
    conn = np.ascontiguousarray(part.conn, np.int32)
    ne, nk = conn.shape
    W = vals.shape[2]
    rptr, re, rk, _ = family_rows(m, fam)
    out = np.zeros((m.nn, W))
    stats = np.zeros(3, np.int64)
    x = np.ascontiguousarray(m.x, np.float64)
    sim.asm_simulate(nk, W, ne, m.nn, _p(conn), _p(x), _p(rptr), _p(re), _p(rk),
                     _p(np.ascontiguousarray(prefix_ok, np.int8)), _p(np.ascontiguousarray(vals)),
                     _p(None if init is None else np.ascontiguousarray(init)), _p(out), _p(stats))
    return out, stats


def asspar4(m, vals_h, vals_q):
    """Plain ASSPAR4: every node sums its rows in skyline order from zero (forces of hex and shell
    rows in one running sum, moments from the shell rows)."""
    rows = np.zeros((m.lsky, 6))
    if m.hexp is not None:
        rows[m.hexp.iads.reshape(-1), :3] = vals_h.reshape(-1, 3)
    if m.q4 is not None:
        rows[m.q4.iadc.reshape(-1), :] = vals_q.reshape(-1, 6)
    out = np.zeros((m.nn, 6))
    counts = np.diff(m.adsky)
    for j in range(int(counts.max())):
        has = counts > j
        out[has] = out[has] + rows[m.adsky[:-1][has] + j]
    return out


def corner_values(rng, ne, nk, W):
    """Corner words whose sum depends on the order: mixed magnitudes, signs and signed zeros."""
    v = rng.standard_normal((ne, nk, W)) * 10.0 ** rng.integers(-6, 7, (ne, nk, W))
    v[rng.random((ne, nk, W)) < 0.02] = -0.0
    return v


def check_model(sim, m, seed=0):
    rng = np.random.default_rng(seed)
    vh = corner_values(rng, m.ne_hex, 8, 3) if m.hexp is not None else None
    vq = corner_values(rng, m.ne_q4, 4, 6) if m.q4 is not None else None
    ref = asspar4(m, vh, vq)
    got = np.zeros((m.nn, 6))
    nh = np.zeros(m.nn, np.int64)
    stats = {}
    if m.hexp is not None:
        sh, st = simulate(sim, m, "hex", vh, None, np.ones(m.nn, np.int8))
        got[:, :3] = sh
        nh = family_rows(m, "hex")[3]
        stats["hex"] = st
    if m.q4 is not None:
        init = np.zeros((m.nn, 6))
        init[:, :3] = got[:, :3]
        sq, st = simulate(sim, m, "q4", vq, init, (nh == 0).astype(np.int8))
        has_q = family_rows(m, "q4")[3] > 0
        got[has_q] = sq[has_q]
        stats["q4"] = st
    assert np.array_equal(got.view(np.int64), ref.view(np.int64)), "tiled sums differ from the ASSPAR4 sum"
    return stats


def test_block_with_plate_sums_bit_for_bit(sim):
    st = check_model(sim, M.block_with_plate(n=16, plate=40, seed=3), seed=1)
    # a 16^3 block in 8 x 8 x 4 bricks: about 3.3 stored words-triples per node instead of 8
    nodes = 17 ** 3
    assert (st["hex"][0] + st["hex"][1]) / nodes < 4.0
    assert st["hex"][2] == 8 * 16 ** 3


def test_bar_tube_and_distorted_block_sum_bit_for_bit(sim):
    check_model(sim, M.hex_model_to_model(M.cantilever_bar(6, 5, 40)), seed=2)
    check_model(sim, M.square_tube(width=20.0, length=40.0), seed=3)
    x, conn = M.hex_block(12, 9, 7, 1.0)
    rng = np.random.default_rng(4)
    x = x + 0.2 * rng.standard_normal(x.shape)
    m = M.build_model(x, hex_conn=conn, hex_mat=M.Law2(rho=7.85e-9, young=2.1e5, nu=0.3, a=350.0, b=275.0, n=0.36))
    check_model(sim, m, seed=5)


def test_shuffled_element_ids_sum_bit_for_bit(sim):
    """Element numbering unrelated to the geometry: the rows of a node come in a scattered order,
    so the runs summed in LDS are short, and the sums must still be exact."""
    x, conn = M.hex_block(10, 10, 10, 1.0)
    rng = np.random.default_rng(6)
    conn = conn[rng.permutation(conn.shape[0])]
    m = M.build_model(x, hex_conn=conn, hex_mat=M.Law2(rho=7.85e-9, young=2.1e5, nu=0.3, a=350.0, b=275.0, n=0.36))
    check_model(sim, m, seed=7)


def test_fewer_elements_than_a_tile(sim):
    check_model(sim, M.block_with_plate(n=2, plate=3, seed=8), seed=9)
