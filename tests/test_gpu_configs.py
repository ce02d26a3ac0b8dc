"""Parity at the measured configurations' own sizes, through the engine-side drivers:
  - the star join at fact 60M x dimension 1.5M against the oracle's query restatement;
  - all 256 queries of the batched workload (rows lognormal in [4096, 4M], 10 to 1e5 distinct
    keys), batched and one by one, against numpy per query (and the oracle's aggregation on a
    sample of them);
  - the serving mix on the 10M-row web table, the 60M x 1.5M star join and 256 segments of 64k
    rows: every request's result in the continuously batched stream (learned sizing models on,
    HBM-budget admission) equals its result alone, in the per-query stream, and the CPU's (numpy /
    oracle) result.
Results are compared through the drivers' order-independent digests, recomputed here on the CPU
from the independent results."""
import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi, queries as Q
from test_gpu_batch import c4_plan

pytestmark = pytest.mark.gpu

U64 = np.uint64
MASK = (1 << 64) - 1


def c4_digest(key, cnt, s, mn, mx) -> int:
    """The drivers' digest of one query's rows (key, count, sum, min, max): a sum over rows of a
    per-row mix, plus the row count (order-independent)."""
    g = len(key)
    h = np.full(g, 0x9e3779b97f4a7c15, U64)
    with np.errstate(over="ignore"):
        for v in (key, cnt, s, mn, mx):
            h ^= np.asarray(v, np.int64).view(U64)
            h *= U64(0xbf58476d1ce4e5b9)
            h ^= h >> U64(31)
        total = int(h.sum(dtype=U64)) if g else 0
    return (0x12345678 + total + g) & MASK


def c4_numpy(key, val, theta):
    """SELECT key, COUNT(*), SUM(val), MIN(val), MAX(val) WHERE val > theta GROUP BY key."""
    m = val > theta
    k, v = key[m], val[m]
    if len(k) == 0:
        z = np.zeros(0, np.int64)
        return z, z, z, z, z
    order = np.argsort(k, kind="stable")
    k, v = k[order], v[order]
    starts = np.concatenate([[0], np.nonzero(np.diff(k))[0] + 1])
    cnt = np.diff(np.concatenate([starts, [len(k)]]))
    return (k[starts].astype(np.int64), cnt.astype(np.int64), np.add.reduceat(v, starts),
            np.minimum.reduceat(v, starts), np.maximum.reduceat(v, starts))


def c4_shapes(seed, nq):
    g = ws.datagen()
    rows, distinct, theta = (np.zeros(nq, np.int64) for _ in range(3))
    g.wsg_c4_shapes(seed, nq, ws.ptr(rows), ws.ptr(distinct), ws.ptr(theta))
    return rows, distinct, theta


def c4_data(seed, q, rows, distinct):
    g = ws.datagen()
    key, val = np.empty(rows, np.int32), np.empty(rows, np.int64)
    g.wsg_c4_data(seed, q, rows, distinct, ws.ptr(key), ws.ptr(val))
    return key, val


def test_c3_star_join_full_size(gpu):
    H.say("case config star join, fact 60M x dim 1.5M")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    fk, qty, price, flag, dkey, dattr = H.gen_c3(60_000_000, 1_500_000)
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    try:
        rows, st = Q.c3(fact, dim)  # distinct dimension keys: the join runs as the fused filter
        Q.set_join_rewrite(False)
        try:
            jrows, jst = Q.c3(fact, dim)  # the hash-join plan
        finally:
            Q.set_join_rewrite(True)
    finally:
        fact.free()
        dim.free()
    exp = H.oracle_c3(fk, qty, price, flag, dkey, dattr)
    assert np.array_equal(rows[np.argsort(rows[:, 0])], exp)
    assert np.array_equal(jrows[np.argsort(jrows[:, 0])], exp)
    sel = dattr < 50
    assert st[0] == jst[0] == int(sel.sum())
    assert st[2] == jst[2] == int(exp[:, 1].sum())  # every join pair lands in one group
    assert st[1] == jst[1]
    H.say(f"  dim={st[0]} fact_filtered={st[1]} pairs={st[2]} groups={len(exp)}")


def test_c3_filtered_keys_distinct_without_key_column(gpu):
    """Repeated dimension keys whose extra copies all fail the filter: the key column is not a key,
    but the filtered keys are distinct (as many bitmap bits as filtered rows), so the join still
    runs as the fused filter, over the materialized build side."""
    H.say("case config star join, repeated keys filtered out")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    fk, qty, price, flag, dkey, dattr = H.gen_c3(3_000_000, 200_000)
    dkey = dkey.copy()
    dattr = dattr.copy()
    dkey[1::7] = dkey[0:-1:7][: len(dkey[1::7])]
    dattr[1::7] = 999  # the copies never pass attr < 50
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    try:
        gpu.ws_prof_reset()
        gpu.ws_prof_only(b"")
        gpu.ws_prof_enable(1)
        try:
            rows, st = Q.c3(fact, dim)
        finally:
            prof = ws.prof_rows()
            gpu.ws_prof_enable(0)
    finally:
        fact.free()
        dim.free()
    exp = H.oracle_c3(fk, qty, price, flag, dkey, dattr)
    assert np.array_equal(rows[np.argsort(rows[:, 0])], exp)
    assert "bitmap_count" in prof and not any(k.startswith("join_") for k in prof), sorted(prof)
    assert st[2] == st[1] == int(exp[:, 1].sum())


def test_c3_duplicate_dimension_keys_keep_the_join(gpu):
    """A dimension whose filtered keys repeat: each fact row joins every copy, so the driver must
    keep the hash join (the bitmap filter alone would count each row once)."""
    H.say("case config star join, duplicate dimension keys")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    fk, qty, price, flag, dkey, dattr = H.gen_c3(3_000_000, 200_000)
    dkey = dkey.copy()
    dkey[1::7] = dkey[0:-1:7][: len(dkey[1::7])]  # every 7th key repeats its neighbour's
    dattr = dattr.copy()
    dattr[1::7] = dattr[0:-1:7][: len(dattr[1::7])]  # with the same attribute, so both copies pass
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    try:
        rows, st = Q.c3(fact, dim)
    finally:
        fact.free()
        dim.free()
    exp = H.oracle_c3(fk, qty, price, flag, dkey, dattr)
    assert np.array_equal(rows[np.argsort(rows[:, 0])], exp)
    assert st[2] == int(exp[:, 1].sum()) and st[2] > st[1]  # more pairs than fact rows: the join ran


def test_c4_all_256_queries(gpu):
    H.say("case config 256 batched GROUP BY queries vs numpy per query")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    seed, nq = 4, 256
    rows, distinct, theta = c4_shapes(seed, nq)
    s = Q.C4Set(seed=seed, nq=nq)
    try:
        assert s.rows == int(rows.sum())
        db, gb = s.run(batched=True)
        dp, gp = s.run(batched=False)
    finally:
        s.free()
    checked_by_oracle = 0
    for q in range(nq):
        key, val = c4_data(seed, q, int(rows[q]), int(distinct[q]))
        res = c4_numpy(key, val, int(theta[q]))
        want = c4_digest(*res)
        assert int(gb[q]) == len(res[0]) and int(gp[q]) == len(res[0]), (q, gb[q], gp[q], len(res[0]))
        assert int(db[q]) == want, f"query {q}: batched digest differs ({rows[q]} rows, {distinct[q]} keys)"
        assert int(dp[q]) == want, f"query {q}: per-query digest differs"
        if q % 32 == 0 and rows[q] <= 1_000_000:  # the oracle's aggregation on a sample
            t = H.Table({"key": key, "val": val}, int(rows[q]))
            b = abi.EnvBuilder()
            sk, sv = b.slot(abi.T_I32), b.slot(abi.T_I64)
            plan = c4_plan(b, sk, sv, int(theta[q]))
            ok, oa, _ = H.oracle_aggregate(plan, b.env, H.Bound(t, [("key", abi.T_I32, 0), ("val", abi.T_I64, 0)], False),
                                           max_groups=200_000)
            bits = oa.view(np.int64)
            osum = bits[:, 2] * (1 << 32) + bits[:, 1]
            assert c4_digest(ok.view(np.int64)[:, 0], oa[:, 0].astype(np.int64), osum, bits[:, 3], bits[:, 4]) == want
            checked_by_oracle += 1
    assert checked_by_oracle >= 4
    H.say(f"  rows={int(rows.sum())} groups={int(gb.sum())} oracle-checked={checked_by_oracle}")


def web_digest(full) -> int:
    """The web query's digest: groups, distinct (user, url) pairs, then the top-10 counts."""
    d = (len(full) * 1000003 + int(full[:, 3].sum())) & MASK
    for c in full[:10, 1]:
        d = (d * 31 + int(c)) & MASK
    return d


def join_digest(rows) -> int:
    d = len(rows)
    for v in rows.reshape(-1):
        d = (d * 1099511628211 + (int(v) & MASK)) & MASK
    return d


def test_c5_serving_full_tables(gpu):
    H.say("case config serving mix on the 10M web table and the 60M x 1.5M star join")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    seed, nq, seg = 5, 256, 65536
    pool = Q.C4Set(seed=seed, nq=nq, rows_override=seg)
    user, region, date, url, dur = H.gen_c2(10_000_000, seed=2)
    web = Q.DeviceTable([user, region, date, url, dur], [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32],
                        [0, 1, 0, 0, 0])
    fk, qty, price, flag, dkey, dattr = H.gen_c3(60_000_000, 1_500_000)
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    from warpstore_amd import sizing
    try:
        # the serving form with learned sizing on: group-count and join-size models trained on the
        # GPU from the same queries under other filters, and a tight HBM budget for admission
        sizing.log_clear("groups")
        sizing.log_clear("joins")
        for d, r in ((19723 + 90, 64), (19723 + 200, 128), (19723 + 365, 256), (19723 + 30, 200)):
            Q.c2(web, 10, date_lt=d, region_lt=r)
        Q.set_join_rewrite(False)  # the join sizes come from the hash-join plan
        try:
            for a in (10, 30, 100, 300):
                Q.c3(fact, dim, attr_lt=a)
        finally:
            Q.set_join_rewrite(True)
        gm, jm = sizing.train_from_log("groups", epochs=300), sizing.train_from_log("joins", epochs=300)
        n, rate = 120, 300.0
        Q.set_sizers(gm.h, jm.h)
        st1, d1, ty = Q.serve(pool, web, fact, dim, rate=rate, count=n, mode=1, workers=2, max_batch=256, seed=11,
                              budget_bytes=1 << 30)
        Q.set_sizers(None, None)
        st0, d0, ty0 = Q.serve(pool, web, fact, dim, rate=rate, count=n, mode=0, workers=4, seed=11)
        assert st1["completed"] == n and st0["completed"] == n
        assert np.array_equal(ty, ty0) and set(ty.tolist()) == {0, 1, 2}
        assert np.array_equal(d1, d0)
        for i in range(n):
            assert Q.serve_alone(pool, web, fact, dim, rate=rate, count=n, seed=11, index=i) == int(d1[i]), i
        # the heavy requests against the CPU: the web query through the oracle, the star join too
        full = H.oracle_c2(user, region, date, url, dur)
        jrows = H.oracle_c3(fk, qty, price, flag, dkey, dattr)
        assert all(int(d1[i]) == web_digest(full) for i in range(n) if ty[i] == 1)
        assert all(int(d1[i]) == join_digest(jrows) for i in range(n) if ty[i] == 2)
        H.say(f"  batched {st1}")
        H.say(f"  per-query {st0}")
    finally:
        Q.set_sizers(None, None)
        for t in (web, fact, dim):
            t.free()
        pool.free()
    # the small requests against numpy: the stream's (segment, theta) picks, recomputed on the CPU
    _, distinct, _ = c4_shapes(seed, nq)
    picks = serve_picks(n, rate, 11, nq)
    for i in range(n):
        if ty[i] != 0:
            continue
        q, theta = picks[i]
        key, val = c4_data(seed, q, seg, int(distinct[q]))
        assert int(d1[i]) == c4_digest(*c4_numpy(key, val, theta)), i


def _u01(seed: int, i: int) -> float:
    x = (seed * 0x9e3779b97f4a7c15 + i * 0xd1b54a32d192ed03 + 1) & MASK
    x ^= x >> 30
    x = (x * 0xbf58476d1ce4e5b9) & MASK
    x ^= x >> 27
    x = (x * 0x94d049bb133111eb) & MASK
    x ^= x >> 31
    return (x >> 11) * (1.0 / 9007199254740992.0)


def serve_picks(count, rate, seed, pool_size):
    """(pool index, theta) of each request of the serving stream (the stream's own counter-based
    draws: index 4i+2 picks the segment, 4i+3 the filter constant)."""
    return [(int(_u01(seed, 4 * i + 2) * pool_size), 100000 + int(_u01(seed, 4 * i + 3) * 800000.0)) for i in range(count)]
