"""Serving forms: the many-small-aggregations workload batched into one launch equals the plain
per-query path query by query, and in the continuous-batching serving loop every request's result
is the same as when it runs alone (small aggregates, web COUNT DISTINCT queries, star joins)."""
import numpy as np
import pytest

import helpers as H
from warpstore_amd import abi, queries as Q

pytestmark = pytest.mark.gpu


def test_c4_batched_equals_per_query(gpu):
    H.say("case C4 batched vs per-query (48 queries, lognormal sizes)")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    s = Q.C4Set(seed=4, nq=48)
    try:
        db, gb = s.run(batched=True)
        dp, gp = s.run(batched=False)
    finally:
        s.free()
    assert np.array_equal(gb, gp)
    assert np.array_equal(db, dp)
    assert (gb > 0).sum() > 40


def test_serving_mix_results_equal_alone(gpu):
    H.say("case serving mix vs alone")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    pool = Q.C4Set(seed=5, nq=32, rows_override=65536)
    user, region, date, url, dur = H.gen_c2(1_000_000)
    web = Q.DeviceTable([user, region, date, url, dur], [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32],
                        [0, 1, 0, 0, 0])
    fk, qty, price, flag, dkey, dattr = H.gen_c3(2_000_000, 50_000)
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    try:
        n = 90
        st1, d1, ty = Q.serve(pool, web, fact, dim, rate=400.0, count=n, mode=1, workers=2, max_batch=64, seed=9)
        st0, d0, ty0 = Q.serve(pool, web, fact, dim, rate=400.0, count=n, mode=0, workers=4, seed=9)
        assert st1["completed"] == n and st0["completed"] == n
        assert np.array_equal(ty, ty0) and set(ty.tolist()) == {0, 1, 2}
        assert np.array_equal(d1, d0)
        for i in range(n):
            assert Q.serve_alone(pool, web, fact, dim, rate=400.0, count=n, seed=9, index=i) == int(d1[i]), i
        H.say(f"batched {st1}")
        H.say(f"per-query {st0}")
    finally:
        for t in (web, fact, dim):
            t.free()
        pool.free()
