"""GPU parity of the partitioned hash GROUP BY (count / scatter / merge passes) against the CPU oracle.

The partitioned plan must give the same groups as the reference's shared-table plan: keys, COUNT,
integer slots and MIN / MAX bit-exact, double sums within 1e-12 relative. Cases: the web table's
phase-1 key pair (compiled and interpreted), NULLs / NaN / deleted rows / double keys with -0.0,
heavy skew (one key for most rows), partitions that outgrow their LDS table (spill table and
overflow retry), an empty result, a wide plan, and a partitioned result that more batches follow."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu

STRAT_AUTO, STRAT_TABLE, STRAT_PART = 0, 1, 2


def jit_for(L, env, ncols, plan):
    L.ws_jit_set_mode(abi.JIT_SYNC)
    out = C.c_void_p()
    ws.check(L.ws_jit_get(C.byref(env), ncols, C.byref(plan), abi.JF_GROUPED, b"test", C.byref(out)))
    assert out.value, "plan did not compile"
    return out


def pair_plan(b, su, sr, sd, sl, sdu):
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sd), b.ilit(19723 + 300), (abi.OP_LT,), (abi.OP_COL, sr), b.ilit(224), (abi.OP_LT,),
                         (abi.OP_AND,)])
    plan.nkeys = 2
    plan.keys[0], plan.keys[1] = b.col(su), b.col(sl)
    plan.key_int[0] = plan.key_int[1] = 1
    plan.naggs = 2
    plan.ops[0], plan.ops[1] = abi.AGG_COUNT, abi.AGG_SUM
    plan.agg_int[0] = plan.agg_int[1] = 1
    plan.aggs[1] = b.col(sdu)
    return plan


def web_setup(n, batch_rows=1 << 18, n_users=50_000, n_urls=20_000, seed=2):
    user, region, date, url, dur = H.gen_c2(n, seed=seed, n_users=n_users, n_urls=n_urls)
    t = H.Table({"user": user, "region": region, "date": date, "url": url, "dur": dur}, batch_rows)
    t.encode("region")
    b = abi.EnvBuilder()
    su, sr, sd, sl, sdu = (b.slot(abi.T_I64), b.slot(abi.T_I32, enc=2), b.slot(abi.T_I32), b.slot(abi.T_I32),
                           b.slot(abi.T_I32))
    plan = pair_plan(b, su, sr, sd, sl, sdu)
    slots = [("user", abi.T_I64, 0), ("region", abi.T_I32, 0), ("date", abi.T_I32, 0), ("url", abi.T_I32, 0),
             ("dur", abi.T_I32, 0)]
    return t, b, plan, slots


@pytest.mark.parametrize("mode", ["interp", "jit"])
def test_partitioned_pair_keys(gpu, mode):
    """The web table's (user_id, url) pairs, 600k rows in ragged batches, forced partitioned."""
    H.say(f"case partitioned pair keys {mode}")
    t, b, plan, slots = web_setup(600_000, batch_rows=[250_000, 1, 0, 99_999, 250_000])
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    jit = jit_for(gpu, b.env, len(slots), plan) if mode == "jit" else None
    gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 19, jit=jit, strategy=STRAT_PART)
    assert H.LAST_RUN["partitioned"], "the partitioned plan did not run"
    H.assert_groups_equal(plan, gk, ga, ok, oa)
    # the shared-table plan on the same input agrees too
    tk, ta, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 19, jit=jit, strategy=STRAT_TABLE)
    assert not H.LAST_RUN["partitioned"]
    H.assert_groups_equal(plan, tk, ta, ok, oa)


def test_partitioned_nulls_deleted_double_keys(gpu):
    """Nullable int slots, NaN doubles, deleted rows, a double key with -0.0 / +0.0 and NaN, split sums."""
    H.say("case partitioned nulls/deleted/double keys")
    n = 400_000
    rng = np.random.default_rng(21)
    k1 = rng.integers(0, 40_000, n).astype(np.int64)
    k2 = np.round(rng.random(n) * 8, 0) - 4.0  # -4 .. 4 with both signs of zero
    k2[rng.random(n) < 0.01] = -0.0
    k2[rng.random(n) < 0.01] = np.nan
    val = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    val[rng.random(n) < 0.1] = abi.NULL_I64
    dd = rng.random(n) * 100 - 50
    dd[rng.random(n) < 0.05] = np.nan
    deleted = rng.random(n) < 0.15
    t = H.Table({"k1": k1, "k2": k2, "val": val, "dd": dd}, [131_072, 100_000, 0, 168_928], deleted)
    b = abi.EnvBuilder()
    s1, s2, sv, sd = b.slot(abi.T_I64), b.slot(abi.T_F64), b.slot(abi.T_I64N), b.slot(abi.T_F64)
    plan = abi.AggPlan()
    plan.nkeys = 2
    plan.keys[0], plan.keys[1] = b.col(s1), b.col(s2)
    plan.key_int[0], plan.key_int[1] = 1, 0
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX, abi.AGG_SUM, abi.AGG_MIN,
           abi.AGG_MAX]
    plan.naggs = len(ops)
    for a, o in enumerate(ops):
        plan.ops[a] = o
    plan.agg_int[0] = 1
    plan.aggs[1] = b.prog([(abi.OP_COL, sv), (abi.OP_ISNULL,), (abi.OP_NOT,)])
    for a in (2, 3, 4, 5):
        plan.aggs[a] = b.col(sv)
        plan.agg_int[a] = 1
        plan.agg_bits[a] = 1
        plan.agg_null[a] = abi.T_I64N
    plan.agg_part[2], plan.agg_part[3] = 1, 2
    for a in (6, 7, 8):
        plan.aggs[a] = b.col(sd)
    slots = [("k1", abi.T_I64, 0), ("k2", abi.T_F64, 0), ("val", abi.T_I64N, 0), ("dd", abi.T_F64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    for mode in ("interp", "jit"):
        jit = jit_for(gpu, b.env, len(slots), plan) if mode == "jit" else None
        gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 20, jit=jit, strategy=STRAT_PART)
        assert H.LAST_RUN["partitioned"]
        H.assert_groups_equal(plan, gk, ga, ok, oa)


@pytest.mark.parametrize("mode", ["interp", "jit"])
def test_partitioned_heavy_skew(gpu, mode):
    """90 % of the rows in one group (pre-aggregation and wave folding), the rest spread; the
    interpreted run takes the column-only evaluator (every program is one column)."""
    H.say(f"case partitioned heavy skew {mode}")
    n = 1_500_000
    rng = np.random.default_rng(5)
    key = rng.integers(0, 200_000, n).astype(np.int64)
    key[rng.random(n) < 0.9] = 7
    val = rng.integers(0, 1 << 20, n).astype(np.int64)
    p = rng.random(n)
    t = H.Table({"key": key, "val": val, "p": p}, 1 << 20)
    b = abi.EnvBuilder()
    sk, sv, sp = b.slot(abi.T_I64), b.slot(abi.T_I64), b.slot(abi.T_F64)
    plan = abi.AggPlan()
    plan.nkeys = 1
    plan.keys[0] = b.col(sk)
    plan.key_int[0] = 1
    plan.naggs = 4
    plan.ops[0], plan.ops[1], plan.ops[2], plan.ops[3] = abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_MAX, abi.AGG_SUM
    plan.agg_int[0] = plan.agg_int[1] = plan.agg_int[2] = 1
    plan.agg_bits[1] = plan.agg_bits[2] = 1
    plan.aggs[1] = plan.aggs[2] = b.col(sv)
    plan.aggs[3] = b.col(sp)
    slots = [("key", abi.T_I64, 0), ("val", abi.T_I64, 0), ("p", abi.T_F64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    jit = jit_for(gpu, b.env, len(slots), plan) if mode == "jit" else None
    gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 19, jit=jit, strategy=STRAT_PART)
    assert H.LAST_RUN["partitioned"]
    H.assert_groups_equal(plan, gk, ga, ok, oa)
    assert int(ga[:, 0].sum()) == n
    tk, ta, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 19, jit=jit, strategy=STRAT_TABLE)
    H.assert_groups_equal(plan, tk, ta, ok, oa)


def test_partitioned_spill_and_retry(gpu):
    """A wrong estimate (2 partitions for ~150k groups): partitions spill to the shared table,
    which overflows and is retried bigger until it holds them; the result stays exact."""
    H.say("case partitioned spill/retry")
    t, b, plan, slots = web_setup(400_000, n_users=100_000, n_urls=50_000, seed=9)
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    gk, ga, _, cap = H.gpu_aggregate(plan, b.env, dev, cap=1 << 12, strategy=STRAT_PART, expected=1)
    assert H.LAST_RUN["partitioned"] and H.LAST_RUN["retries"] >= 1
    H.assert_groups_equal(plan, gk, ga, ok, oa)


def test_partitioned_empty_and_wide(gpu):
    """No row passes the filter (0 groups); then a plan of 3 keys and 12 slots (16-word records)."""
    H.say("case partitioned empty/wide")
    n = 200_000
    rng = np.random.default_rng(8)
    a = rng.integers(0, 50, n).astype(np.int32)
    c = rng.integers(0, 40, n).astype(np.int64)
    d = rng.integers(0, 30, n).astype(np.int32)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    t = H.Table({"a": a, "c": c, "d": d, "v": v}, 1 << 16)
    b = abi.EnvBuilder()
    sa, sc, sd, sv = b.slot(abi.T_I32), b.slot(abi.T_I64), b.slot(abi.T_I32), b.slot(abi.T_I64)
    slots = [("a", abi.T_I32, 0), ("c", abi.T_I64, 0), ("d", abi.T_I32, 0), ("v", abi.T_I64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sa), b.ilit(-1), (abi.OP_LT,)])  # never true
    plan.nkeys = 1
    plan.keys[0] = b.col(sc)
    plan.key_int[0] = 1
    plan.naggs = 1
    plan.ops[0] = abi.AGG_COUNT
    plan.agg_int[0] = 1
    gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 18, strategy=STRAT_PART)
    assert H.LAST_RUN["partitioned"] and len(gk) == 0
    wide = abi.AggPlan()
    wide.nkeys = 3
    wide.keys[0], wide.keys[1], wide.keys[2] = b.col(sa), b.col(sc), b.col(sd)
    wide.key_int[0] = wide.key_int[1] = wide.key_int[2] = 1
    wide.naggs = 12
    for s in range(12):  # COUNT, SUM, MIN, MAX repeated; int64 bits on the SUM / MIN / MAX slots
        wide.ops[s] = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX][s % 4]
        wide.agg_int[s] = 1
        if s % 4:
            wide.aggs[s] = b.col(sv)
            wide.agg_bits[s] = 1
    ok, oa, _ = H.oracle_aggregate(wide, b.env, host)
    jit = jit_for(gpu, b.env, len(slots), wide)
    gk, ga, _, _ = H.gpu_aggregate(wide, b.env, dev, cap=1 << 17, jit=jit, strategy=STRAT_PART)
    assert H.LAST_RUN["partitioned"]
    H.assert_groups_equal(wide, gk, ga, ok, oa)


def _add_batches(L, h, b, dev, t, which):
    for bi in which:
        e = H.env_with_batch(b.env, dev, bi)
        ws.check(L.ws_agg_add_batch(h, C.byref(e), dev.nslots, t.sizes[bi], dev.dels[bi]))


def _compact(L, h, nk, na):
    kp, ap, g = C.c_void_p(), C.c_void_p(), C.c_int64()
    ws.check(L.ws_agg_compact(h, C.byref(kp), C.byref(ap), C.byref(g)))
    G = g.value
    keys, accs = np.zeros(max(G * nk, 1)), np.zeros(max(G * na, 1))
    if G:
        ws.check(L.ws_d2h(ws.ptr(keys), kp, G * nk * 8))
        ws.check(L.ws_d2h(ws.ptr(accs), ap, G * na * 8))
    L.ws_tmp_free(kp)
    L.ws_tmp_free(ap)
    return (np.ascontiguousarray(keys[:G * nk].reshape(nk, G).T), np.ascontiguousarray(accs[:G * na].reshape(na, G).T))


def test_partitioned_compact_and_windows(gpu):
    """compact() of a partitioned run (column-major, in device memory); windows never partition
    and accumulate into the shared table."""
    H.say("case partitioned compact / windows")
    t, b, plan, slots = web_setup(300_000, batch_rows=100_000, n_users=30_000, n_urls=10_000, seed=4)
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    L = ws.lib()
    h = C.c_void_p()
    ws.check(L.ws_agg_create(C.byref(plan), 1 << 19, None, C.byref(h)))
    try:
        ws.check(L.ws_agg_set_strategy(h, STRAT_PART, 0))
        _add_batches(L, h, b, dev, t, range(3))
        ws.check(L.ws_agg_run(h, C.byref(b.env), None))
        assert L.ws_agg_partitioned(h)
        gc = C.c_int64()
        ws.check(L.ws_agg_group_count(h, C.byref(gc)))
        assert gc.value == len(ok)
        gk, ga = _compact(L, h, 2, 2)
    finally:
        L.ws_agg_destroy(h)
    H.assert_groups_equal(plan, gk, ga, ok, oa)
    h = C.c_void_p()
    ws.check(L.ws_agg_create(C.byref(plan), 1 << 19, None, C.byref(h)))
    try:
        ws.check(L.ws_agg_set_strategy(h, STRAT_PART, 0))
        _add_batches(L, h, b, dev, t, (0, 1))
        ws.check(L.ws_agg_run_window(h, C.byref(b.env), None))
        assert not L.ws_agg_partitioned(h)
        _add_batches(L, h, b, dev, t, (2,))
        ws.check(L.ws_agg_run_window(h, C.byref(b.env), None))
        ov = C.c_int()
        ws.check(L.ws_agg_overflowed(h, C.byref(ov)))
        assert ov.value == 0
        gk, ga = _compact(L, h, 2, 2)
    finally:
        L.ws_agg_destroy(h)
    H.assert_groups_equal(plan, gk, ga, ok, oa)


def test_partitioned_run_twice_migrates(gpu):
    """A second run() after a partitioned one (the batches stay added, so every row counts twice):
    the partitioned groups move into the shared table first, then the table run adds to them."""
    H.say("case partitioned run twice (migration)")
    t, b, plan, slots = web_setup(300_000, batch_rows=100_000, n_users=30_000, n_urls=10_000, seed=6)
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    L = ws.lib()
    h = C.c_void_p()
    ws.check(L.ws_agg_create(C.byref(plan), 1 << 19, None, C.byref(h)))
    try:
        ws.check(L.ws_agg_set_strategy(h, STRAT_PART, 0))
        _add_batches(L, h, b, dev, t, range(3))
        ws.check(L.ws_agg_run(h, C.byref(b.env), None))
        assert L.ws_agg_partitioned(h)
        ws.check(L.ws_agg_run(h, C.byref(b.env), None))
        assert not L.ws_agg_partitioned(h)
        ov = C.c_int()
        ws.check(L.ws_agg_overflowed(h, C.byref(ov)))
        assert ov.value == 0
        g = C.c_int64()
        ws.check(L.ws_agg_fetch_begin(h, C.byref(g)))
        keys, accs = np.zeros(g.value * 2), np.zeros(g.value * 2)
        ws.check(L.ws_agg_fetch_copy(h, ws.ptr(keys), ws.ptr(accs)))
    finally:
        L.ws_agg_destroy(h)
    H.assert_groups_equal(plan, keys.reshape(-1, 2), accs.reshape(-1, 2), ok, oa * 2)


@pytest.mark.parametrize("case", ["narrow", "negative_wide", "full_range", "nullable"])
def test_partitioned_packed_records(gpu, case):
    """Integer keys: packed partial-group records (the keys' ranges fit 63 bits: narrow, negative
    and wide) or the full layout (full-range int64 pairs, a nullable key whose NULL sentinel spans
    the range), each against the oracle and against the run with packing switched off."""
    H.say(f"case partitioned packed records {case}")
    n = 500_000
    rng = np.random.default_rng(len(case))
    if case == "narrow":
        k1, k2 = rng.integers(0, 3000, n), rng.integers(0, 200, n)
    elif case == "negative_wide":
        k1 = rng.integers(-(1 << 40), 1 << 40, n) // 100_003 * 100_003  # ~22M distinct over 41 bits
        k1[rng.random(n) < 0.5] = -(1 << 40)
        k2 = rng.integers(-7, 9, n)
    elif case == "full_range":
        k1 = rng.integers(-(1 << 62), 1 << 62, n) * 2
        k1[:100] = np.iinfo(np.int64).min
        k1[100:200] = np.iinfo(np.int64).max
        k2 = rng.integers(-(1 << 62), 1 << 62, n)
        k1[rng.random(n) < 0.7] = 12345  # repeats, so groups merge
        k2[rng.random(n) < 0.7] = -9
    else:
        k1, k2 = rng.integers(0, 50_000, n), rng.integers(0, 10, n)
        k1[rng.random(n) < 0.05] = abi.NULL_I64
    k1, k2 = k1.astype(np.int64), k2.astype(np.int64)
    val = rng.integers(-1000, 1000, n).astype(np.int64)
    t = H.Table({"k1": k1, "k2": k2, "val": val}, [200_000, 300_000])
    b = abi.EnvBuilder()
    t1 = abi.T_I64N if case == "nullable" else abi.T_I64
    s1, s2, sv = b.slot(t1), b.slot(abi.T_I64), b.slot(abi.T_I64)
    plan = abi.AggPlan()
    plan.nkeys = 2
    plan.keys[0], plan.keys[1] = b.col(s1), b.col(s2)
    plan.key_int[0] = plan.key_int[1] = 1
    plan.naggs = 3
    plan.ops[0], plan.ops[1], plan.ops[2] = abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_MIN
    plan.agg_int[0] = plan.agg_int[1] = plan.agg_int[2] = 1
    plan.aggs[1] = plan.aggs[2] = b.col(sv)
    plan.agg_bits[1] = plan.agg_bits[2] = 1
    slots = [("k1", t1, 0), ("k2", abi.T_I64, 0), ("val", abi.T_I64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    L = ws.lib()
    for pack in (1, 0):
        assert L.ws_tune_set(b"part_pack", pack) == 0
        try:
            for mode in ("interp", "jit"):
                jit = jit_for(gpu, b.env, len(slots), plan) if mode == "jit" else None
                gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 20, jit=jit, strategy=STRAT_PART)
                assert H.LAST_RUN["partitioned"]
                H.assert_groups_equal(plan, gk, ga, ok, oa)
        finally:
            assert L.ws_tune_set(b"part_pack", 1) == 0
