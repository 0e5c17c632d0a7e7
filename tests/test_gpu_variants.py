"""Launch variants of the element kernels (register budget, instruction scheduler)
and their autotuners:
every variant computes bit-identical results, and tuned runs equal untuned ones."""
import numpy as np
import pytest

import oracle as O
from golden_cases import Q4_CASES, build_q4_case
from openradioss_amd import engine as E
from openradioss_amd import model as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_every_variant_gives_the_same_bits(precision, gpu_lib):
    print(f"case: mixed_c120, 4 variants per element kernel, {precision}", flush=True)
    c = Q4_CASES["mixed_c120"]
    m = build_q4_case(c)
    ref = None
    for v in range(4):
        g = E.GpuModel(m, precision=precision)
        g.set_variant("hex", v)
        g.set_variant("q4", 3 - v)
        g.run(c["cycles"])
        out = g.nodes6() + (g.q4_state()["sig"], g.hex_state()["sig"])
        if ref is None:
            ref = out
            if precision == "fp64":
                o = O.ModelOracle(m, nthreads=8, math="cr")
                o.cycle(c["cycles"])
                assert np.array_equal(out[3], o.d)
        else:
            for a, b in zip(out, ref):
                assert np.array_equal(a, b)
        g.close()


def test_model_autotune_keeps_results(gpu_lib):
    print("case: autotune on the 8x20 tube", flush=True)
    m = M.square_tube(width=8.0, length=20.0)
    g = E.GpuModel(m)
    chosen = g.autotune(2)                 # 4 variants x 2 cycles of the shell kernel
    assert 0 <= chosen[1] < 4
    g.run(12)
    n = g.time_state().ncycle
    assert n == 4 * 2 + 12
    o = O.ModelOracle(m, nthreads=8, math="cr")
    o.cycle(n)
    assert np.array_equal(g.nodes6()[3], o.d)
    g.close()


def test_batch_autotune_keeps_results(gpu_lib):
    print("case: batch autotune, 4 tubes fp64 / split", flush=True)
    sw = M.tube_sweep(4, seed=13, width=12.0, length=24.0)
    b = E.GpuBatch()
    gms = []
    for k, (m, _) in enumerate(sw):
        g = E.GpuModel(m, precision="fp64" if k % 2 == 0 else "split")
        b.add(g, 1.0)
        gms.append(g)
    chosen = b.autotune(2)
    assert all(0 <= v < 4 for v in chosen)
    b.run(10)
    for g, (m, _), k in zip(gms, sw, range(4)):
        n = g.time_state().ncycle
        ref = E.GpuModel(m, precision="fp64" if k % 2 == 0 else "split")
        ref.run(n)
        assert np.array_equal(ref.nodes6()[3], g.nodes6()[3])
        ref.close()
    b.close()
    for g in gms:
        g.close()
