"""End-to-end parity of the measured query shapes, driven through the drop-in gdb:: surface the
way the engine's planner drives it (segments of 2^20 rows, dense/hash choice, overflow retry,
two-phase COUNT DISTINCT, GPU top-N, join filter + hash join)."""
import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi, queries as Q

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dense", [True, False])
def test_c1_query_full_size(gpu, dense):
    H.say(f"case query c1 dense={dense}")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    k, v, p = H.gen_c1(1_000_000, seed=1)
    t = Q.DeviceTable([k, v, p], [abi.T_I32, abi.T_I64, abi.T_F64])
    oi, od = Q.c1(t, dense)
    order = np.argsort(oi[:, 0])
    oi, od = oi[order], od[order]
    m = (k < 500) & (v > 1000)
    keys = np.unique(k[m])
    assert np.array_equal(oi[:, 0], keys)
    assert np.array_equal(oi[:, 1], np.bincount(k[m], minlength=1000)[keys])
    exact = np.zeros(1000, dtype=object)
    np.add.at(exact, k[m], v[m].astype(object))
    assert all(int(a) == int(b) for a, b in zip(oi[:, 2], exact[keys]))
    mn = np.full(1000, 2 ** 62, np.int64)
    mx = np.full(1000, -1, np.int64)
    np.minimum.at(mn, k[m], v[m])
    np.maximum.at(mx, k[m], v[m])
    assert np.array_equal(oi[:, 3], mn[keys]) and np.array_equal(oi[:, 4], mx[keys])
    psum = np.bincount(k[m], weights=p[m], minlength=1000)[keys]
    np.testing.assert_allclose(od, psum / oi[:, 1], rtol=1e-12)


@pytest.mark.parametrize("n", [1_000_000, 10_000_000])
def test_c2_query(gpu, n):
    H.say(f"case query c2 n={n}")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    user, region, date, url, dur = H.gen_c2(n, seed=2)
    t = Q.DeviceTable([user, region, date, url, dur], [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32],
                      [0, 1, 0, 0, 0])
    top, st = Q.c2(t, 10)
    full = H.oracle_c2(user, region, date, url, dur)
    H.check_topn(top, full)
    assert st[1] == len(full)  # groups after phase 2 = distinct users passing the filter
    H.say(f"  pairs={st[0]} groups={st[1]} cap1={st[2]} retries={st[3]}")


def test_c3_query_scaled(gpu):
    H.say("case query c3 (fact 6M, dim 150k)")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    fk, qty, price, flag, dkey, dattr = H.gen_c3(6_000_000, 150_000)
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    rows, st = Q.c3(fact, dim)
    exp = H.oracle_c3(fk, qty, price, flag, dkey, dattr)
    rows = rows[np.argsort(rows[:, 0])]
    assert np.array_equal(rows, exp)
    H.say(f"  dim={st[0]} fact_filtered={st[1]} pairs={st[2]}")
