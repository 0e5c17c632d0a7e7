"""The correctly rounded pow / log / exp of the FP64 path (openradioss_amd/csrc/or_math.h,
host build in the oracle's OR_CR_MATH library) against 80-bit long double, and the
two oracle builds against each other."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from golden_cases import Q4_CASES, build_q4_case


def selftest(fn, n, seed):
    L = O.lib("cr")
    L.or_cr_selftest.argtypes = [C.c_int, C.c_int64, C.c_uint64, C.POINTER(C.c_int64)]
    c = (C.c_int64 * 4)()
    L.or_cr_selftest(fn, n, seed, c)
    return list(c)


@pytest.mark.parametrize("fn", [0, 1, 2], ids=["log", "exp", "pow"])
def test_correct_rounding_against_long_double(fn):
    differ, host_closer, ours_closer, beyond_1ulp = selftest(fn, 1_000_000, 777 + fn)
    # never farther from the 64-bit-significand value than the host libm, never beyond 1 ulp
    assert host_closer == 0
    assert beyond_1ulp == 0
    # the host libm (glibc) is not correctly rounded: the two differ on a small fraction of calls,
    # and in the decided cases it is always the host result that is off
    assert differ < 5e-3 * 1_000_000
    assert ours_closer >= 0.5 * differ


def test_shared_logarithm_powers_equal_two_powers():
    """or_pow2 (one logarithm for two exponents of the same base, used for VOLN**(1/3) with
    VOLN**(2/3) and EPXE**n with EPXE**(n-1)) gives exactly the two or_pow results, special
    bases included."""
    L = O.lib("cr")
    L.or_cr_pow2_check.restype = C.c_int64
    L.or_cr_pow2_check.argtypes = [C.c_int64, C.c_uint64]
    assert L.or_cr_pow2_check(200_000, 4242) == 0


def cr(name):
    L = O.lib("cr")
    f = getattr(L, "or_cr_" + name)
    f.restype = C.c_double
    f.argtypes = [C.c_double] * (2 if name == "pow" else 1)
    return f


def test_exact_cases():
    pw, lg, ex = cr("pow"), cr("log"), cr("exp")
    # results representable exactly must come out exactly
    assert pw(2.0, 10.0) == 1024.0
    assert pw(0.25, 0.5) == 0.5
    assert pw(8.0, 1.0 / 3.0) == 2.0       # 1/3 rounds below 1/3: 8^(1/3-d) still rounds to 2
    assert pw(9.0, 0.5) == 3.0
    assert pw(1.0, 0.37) == 1.0 and pw(5.0, 0.0) == 1.0
    assert pw(0.0, 0.3) == 0.0             # EPXE = 0 in the hardening law
    assert lg(1.0) == 0.0 and ex(0.0) == 1.0
    assert lg(2.0) == np.log(2.0)           # ln 2 to the nearest double
    assert ex(1.0) == np.e
    # identities within rounding
    for x in (1e-12, 3.7e-5, 0.61, 1.0 + 2.0 ** -40, 17.0, 4.4e8):
        assert ex(lg(x)) == pytest.approx(x, rel=4e-16 * max(1.0, abs(np.log(x))))


def test_builds_agree_where_no_elementary_function_is_called():
    """Radial return (IPLA=0) with no rate effects calls no pow/log/exp: both builds agree bitwise."""
    c = Q4_CASES["plate_bend_c120"]
    m = build_q4_case(c)
    a, b = O.ModelOracle(m, 4), O.ModelOracle(m, 4, math="cr")
    a.cycle(c["cycles"])
    b.cycle(c["cycles"])
    assert np.array_equal(a.d, b.d) and np.array_equal(a.q["sig"], b.q["sig"])


def test_builds_differ_only_by_libm_rounding():
    """Where the hardening law runs, the two builds stay within the libm spread."""
    c = Q4_CASES["tube_8x20_c150"]
    m = build_q4_case(c)
    a, b = O.ModelOracle(m, 4), O.ModelOracle(m, 4, math="cr")
    a.cycle(c["cycles"])
    b.cycle(c["cycles"])
    rel = np.abs(a.d - b.d).max() / np.abs(a.d).max()
    assert 0.0 < rel < 1e-13
