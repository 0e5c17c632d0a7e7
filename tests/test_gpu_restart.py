"""Restart of a resident model: checkpoint after k cycles, restore into a new model (through an
.npz file), run the remaining cycles, and compare with the uninterrupted run bit for bit --
nodes, element state and cycle scalars, in both precisions."""
import numpy as np
import pytest

from golden_cases import Q4_CASES, build_q4_case
from openradioss_amd import engine as E

pytestmark = pytest.mark.gpu


def _state(g):
    ts = g.time_state()
    out = dict(zip(("x", "v", "vr", "d"), g.nodes6()))
    if g.model.hexp is not None:
        out.update({"hex_" + k: a for k, a in g.hex_state().items()})
    if g.model.q4 is not None:
        out.update({"q4_" + k: a for k, a in g.q4_state().items()})
    return out, (ts.tt, ts.dt1, ts.dt2, ts.dt12, ts.dt2old, ts.ncycle, ts.mstop)


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("name", ["mixed_c120", "tube_8x20_c150"])
def test_restart_continues_bit_for_bit(name, precision, gpu_lib, tmp_path):
    print(f"case: restart of {name} ({precision}) after 37 of 90 cycles", flush=True)
    m = build_q4_case(Q4_CASES[name])
    with E.GpuModel(m, precision=precision) as ref:
        ref.run(90)
        want, want_ts = _state(ref)
    with E.GpuModel(m, precision=precision) as a:
        a.run(37)
        path = str(tmp_path / "ck.npz")
        a.save(path)
    with E.GpuModel.load(m, path, precision=precision) as b:
        assert b.time_state().ncycle == 37
        b.run(53)
        got, got_ts = _state(b)
    assert got_ts == want_ts
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_restore_rejects_displacements_before_nodes(gpu_lib):
    import ctypes as C
    from openradioss_amd import lib as L
    lib = L.load()
    h = C.c_void_p()
    L.check(lib.or_model_create(C.byref(h), 0, 4, 0, 0, 0, L.OR_PREC_FP64))
    try:
        d = np.zeros(12)
        assert lib.or_model_upload_displacement(h, d.ctypes.data_as(C.c_void_p)) != L.OR_OK
    finally:
        lib.or_model_destroy(h)
