"""The allocation-free SFORC3 group drop-in (or_sforc3_group_static) against or_sforc3_group,
bit for bit, on two calls in a row (the second with a smaller group at an offset), both
through the C ABI; and its refusal of a group larger than its staging array."""
import numpy as np
import pytest

from openradioss_amd import engine as E
from openradioss_amd import lib as L
from openradioss_amd import model as M

pytestmark = pytest.mark.gpu


def group(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    x, conn = M.hex_block(nx, ny, nz, 1.0)
    x = x + 0.1 * rng.uniform(-1, 1, x.shape)
    m = M.build_hex_model(x, conn, M.STEEL)
    v = 3e3 * rng.standard_normal((m.nn, 3))
    ixs = np.zeros((m.ne, 11), np.int32)
    ixs[:, 0] = 1
    ixs[:, 1:9] = m.conn + 1
    ixs[:, 10] = np.arange(1, m.ne + 1)
    iads = (m.iads + 1).astype(np.int32)
    st = dict(sig=200.0 * rng.standard_normal((6, m.ne)), eint=np.abs(rng.standard_normal(m.ne)),
              rho=np.full(m.ne, m.params()["rho0"]), vol=m.vol0.copy(),
              pla=np.abs(1e-3 * rng.standard_normal(m.ne)), epsd=np.zeros(m.ne), off=np.ones(m.ne),
              qvis=np.zeros(m.ne))
    fsky = rng.standard_normal((m.lsky, 8))       # words the kernel leaves alone must survive
    return m, x, v, ixs, iads, st, fsky


def run(m, x, v, ixs, iads, st, fsky, nft, nel, static):
    gb = {k: np.ascontiguousarray(a[..., nft:nft + nel]).copy() for k, a in st.items()}
    f = fsky.copy()
    dt2t, ms = E.sforc3_group(E.hex_params(m), x, v, ixs, gb, f, iads, nft=nft, nel=nel, dt1=1.1e-7, dt2t=1e6,
                              static=static)
    return dt2t, ms, gb, f


def test_static_group_equals_allocating_group_twice(gpu_lib):
    for (nx, ny, nz, seed, nft, nel) in ((8, 4, 4, 1, 0, 128), (5, 4, 3, 2, 13, 37)):
        print(f"case: or_sforc3_group_static, {nx}x{ny}x{nz} block, elements {nft}..{nft + nel - 1}", flush=True)
        case = group(nx, ny, nz, seed)
        a = run(*case, nft=nft, nel=nel, static=False)
        b = run(*case, nft=nft, nel=nel, static=True)
        assert a[0] == b[0] and a[1] == b[1]
        for k in a[2]:
            assert np.array_equal(a[2][k], b[2][k]), k
        assert np.array_equal(a[3], b[3])
        assert not np.array_equal(a[3], case[6])   # the group did write its rows


def test_static_group_refuses_more_than_128_elements(gpu_lib):
    m, x, v, ixs, iads, st, fsky = group(9, 4, 4, 3)
    with pytest.raises(L.OrError, match="1..128"):
        run(m, x, v, ixs, iads, st, fsky, nft=0, nel=129, static=True)
