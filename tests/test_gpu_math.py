"""Device evaluation of the FP64 pow / log / exp (or_math_eval through the C ABI) against
the host build of the same routines: bit-for-bit equal on the argument ranges of the path."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from openradioss_amd import engine as E

pytestmark = pytest.mark.gpu


def host(fn, x, y=None):
    L = O.lib("cr")
    L.or_cr_eval.argtypes = [C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), np.float64)
    out = np.empty_like(x)
    L.or_cr_eval({"log": 0, "exp": 1, "pow": 2}[fn], x.size, x.ctypes.data, None if yy is None else yy.ctypes.data,
                 out.ctypes.data)
    return out


def args(fn, n, seed):
    rng = np.random.default_rng(seed)
    if fn == "log":
        return np.exp(rng.uniform(-40.0, 25.0, n)), None
    if fn == "exp":
        return rng.uniform(-60.0, 12.0, n), None
    x = np.exp(rng.uniform(-30.0, 14.0, n))
    y = np.where(rng.integers(0, 2, n) == 0, rng.choice([1.0 / 3.0, 2.0 / 3.0], n), rng.uniform(-1.0, 1.0, n))
    return x, y


@pytest.mark.parametrize("fn", ["log", "exp", "pow"])
def test_device_equals_host_bitwise(fn, gpu_lib):
    print(f"case: or_math_eval {fn}, 1e6 random arguments", flush=True)
    x, y = args(fn, 1_000_000, 11)
    d = E.math_eval(fn, x, y)
    h = host(fn, x, y)
    assert np.array_equal(d, h), f"{np.sum(d != h)} of {x.size} differ"


def test_special_arguments(gpu_lib):
    print("case: or_math_eval special arguments", flush=True)
    x = np.array([0.0, 1.0, 2.0, 8.0, 1e-300, 1e300, 0.5, 3.0])
    y = np.array([0.3, 0.37, 10.0, 1.0 / 3.0, 0.5, 0.25, -0.7, 0.0])
    assert np.array_equal(E.math_eval("pow", x, y), host("pow", x, y))
    z = np.array([0.0, 1.0, -1.0, 700.0, -700.0, 1e-20])
    assert np.array_equal(E.math_eval("exp", z), host("exp", z))
    w = np.array([1.0, 2.0, 1e-300, 1e300, 0.75, 1.5 - 2.0 ** -52])
    assert np.array_equal(E.math_eval("log", w), host("log", w))
