"""GPU parity of the fused filter + aggregate kernels against the CPU oracle, through the C ABI.

Integer, COUNT, MIN / MAX and key results must be bit-exact; double sums within 1e-12 relative
(summation order differs)."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu


def c1_setup(n, batch_rows=1 << 20, seed=1, deleted=None, encode_k=False):
    k, v, p = H.gen_c1(n, seed)
    t = H.Table({"k": k, "v": v, "p": p}, batch_rows, deleted)
    if encode_k:
        t.encode("k")
    b = abi.EnvBuilder()
    sk, sv, sp = b.slot(abi.T_I32), b.slot(abi.T_I64), b.slot(abi.T_F64)
    plan = H.plan_c1(b, sk, sv, sp)
    slots = [("k", abi.T_I32, 0), ("v", abi.T_I64, 0), ("p", abi.T_F64, 0)]
    return t, b, plan, slots


def jit_for(L, env, ncols, plan, fn):
    L.ws_jit_set_mode(abi.JIT_SYNC)
    out = C.c_void_p()
    ws.check(L.ws_jit_get(C.byref(env), ncols, C.byref(plan), fn, b"test", C.byref(out)))
    assert out.value, "plan did not compile"
    return out


@pytest.mark.parametrize("mode", ["interp", "jit"])
@pytest.mark.parametrize("path", ["dense", "hash"])
def test_c1_filter_groupby(gpu, mode, path):
    H.say(f"case c1 {path} {mode}")
    t, b, plan, slots = c1_setup(300_000, batch_rows=[100_000, 1, 0 + 99_999, 100_000])
    dev = H.Bound(t, slots, True)
    host = H.Bound(t, slots, False)
    ds = H.dense_spec(plan, [0], [1000]) if path == "dense" else None
    jit = jit_for(gpu, b.env, len(slots), plan, abi.JF_DENSE if ds is not None else abi.JF_GROUPED) if mode == "jit" else None
    gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 11, dense=ds, jit=jit)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host, dense=ds is not None)
    assert len(ok) == 500
    H.assert_groups_equal(plan, gk, ga, ok, oa)


def test_c1_full_size_dense_jit(gpu):
    """The configuration as specified: 1M rows, one 2^20-row segment, dense GROUP BY, compiled."""
    H.say("case c1 full 1M dense jit")
    t, b, plan, slots = c1_setup(1_000_000)
    dev = H.Bound(t, slots, True)
    host = H.Bound(t, slots, False)
    ds = H.dense_spec(plan, [0], [1000])
    jit = jit_for(gpu, b.env, 3, plan, abi.JF_DENSE)
    gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, dense=ds, jit=jit)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host, dense=True)
    H.assert_groups_equal(plan, gk, ga, ok, oa)
    # Independent check of COUNT / SUM(v) through numpy: sum = hi * 2^32 + lo recombined exactly.
    k, v, _ = t.cols["k"], t.cols["v"], t.cols["p"]
    m = (k < 500) & (v > 1000)
    gk, ga = H.sort_groups(gk, ga)
    keys = gk[:, 0].view(np.int64)
    cnt = np.bincount(k[m], minlength=1000)[keys]
    assert np.array_equal(ga[:, 0].astype(np.int64), cnt)
    sums = np.bincount(k[m], weights=v[m].astype(np.float64), minlength=1000)[keys]  # exact < 2^53
    lo, hi = ga[:, 1].view(np.int64), ga[:, 2].view(np.int64)
    assert np.array_equal(hi * (1 << 32) + lo, sums.astype(np.int64))


def test_hash_overflow_retry(gpu):
    """A capacity far below the group count overflows; the retried aggregation is exact."""
    H.say("case hash overflow retry")
    t, b, plan, slots = c1_setup(200_000)
    plan.where = b.prog([b.ilit(1)])  # every row
    dev = H.Bound(t, slots, True)
    host = H.Bound(t, slots, False)
    gk, ga, _, cap = H.gpu_aggregate(plan, b.env, dev, cap=256)
    assert cap > 256
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    assert len(ok) == 1000
    H.assert_groups_equal(plan, gk, ga, ok, oa)


@pytest.mark.parametrize("mode", ["interp", "jit"])
def test_nokeys_exact_int128(gpu, mode):
    """No GROUP BY: integer sums are 128-bit exact (values near 2^62 overflow int64 quickly)."""
    H.say(f"case nokeys int128 {mode}")
    n = 100_000
    rng = np.random.default_rng(5)
    v = rng.integers(2 ** 61, 2 ** 62, n, dtype=np.int64)
    v[::7] = -v[::7]
    x = rng.random(n) * 100
    t = H.Table({"v": v, "x": x}, [40_000, 60_000])
    b = abi.EnvBuilder()
    sv, sx = b.slot(abi.T_I64), b.slot(abi.T_F64)
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sx), b.dlit(90.0), (abi.OP_LT,)])
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX]
    plan.naggs = len(ops)
    for a, o in enumerate(ops):
        plan.ops[a] = o
    plan.agg_int[0] = 1
    for a in (1, 2, 3):
        plan.aggs[a] = b.col(sv)
        plan.agg_int[a] = 1
    for a in (4, 5, 6):
        plan.aggs[a] = b.col(sx)
    slots = [("v", abi.T_I64, 0), ("x", abi.T_F64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    jit = jit_for(gpu, b.env, 2, plan, abi.JF_NOKEYS) if mode == "jit" else None
    _, ga, ghi, _ = H.gpu_aggregate(plan, b.env, dev, jit=jit)
    _, oa, ohi = H.oracle_aggregate(plan, b.env, host)
    ga, oa = ga[0], oa[0]
    assert ga[0] == oa[0]
    for a in (1, 2, 3):
        assert ga[a:a + 1].view(np.int64)[0] == oa[a:a + 1].view(np.int64)[0], a
    assert ghi[1] == ohi[1]
    m = x < 90.0
    exact = sum(int(q) for q in v[m])
    assert (int(ghi[1]) << 64) + int(ga[1:2].view(np.uint64)[0]) == exact
    np.testing.assert_allclose(ga[4], oa[4], rtol=1e-12)
    assert ga[5] == oa[5] and ga[6] == oa[6]


def test_nullable_deleted_encoded(gpu):
    """Nullable ints (sentinel NULLs), deleted rows, an encoded key column, empty batches."""
    H.say("case nullable/deleted/encoded")
    n = 150_000
    rng = np.random.default_rng(11)
    key = rng.integers(0, 200, n).astype(np.int32)
    val = rng.integers(-10 ** 6, 10 ** 6, n).astype(np.int64)
    val[rng.random(n) < 0.1] = abi.NULL_I64
    dd = rng.random(n) * 10
    dd[rng.random(n) < 0.05] = np.nan
    deleted = rng.random(n) < 0.2
    t = H.Table({"key": key, "val": val, "dd": dd}, [50_000, 0, 50_000, 50_000], deleted)
    t.encode("key")
    b = abi.EnvBuilder()
    sk, sv, sd = b.slot(abi.T_I32, enc=1), b.slot(abi.T_I64N), b.slot(abi.T_F64)
    plan = abi.AggPlan()
    plan.nkeys = 1
    plan.keys[0] = b.col(sk)
    plan.key_int[0] = 1
    # slots: COUNT(*), count(val not null), SUM(val) lo/hi, MIN(val), MAX(val), SUM(dd), MIN(dd)
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX, abi.AGG_SUM, abi.AGG_MIN]
    plan.naggs = len(ops)
    for a, o in enumerate(ops):
        plan.ops[a] = o
    plan.agg_int[0] = 1
    plan.aggs[1] = b.prog([(abi.OP_COL, sv), (abi.OP_ISNULL,), (abi.OP_NOT,)])
    plan.agg_int[1] = 0
    for a in (2, 3, 4, 5):
        plan.aggs[a] = b.col(sv)
        plan.agg_int[a] = 1
        plan.agg_bits[a] = 1
        plan.agg_null[a] = abi.T_I64N
    plan.agg_part[2], plan.agg_part[3] = 1, 2
    plan.aggs[6] = b.col(sd)
    plan.aggs[7] = b.col(sd)
    slots = [("key", abi.T_I32, 0), ("val", abi.T_I64N, 0), ("dd", abi.T_F64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    for mode in ("interp", "jit"):
        jit = jit_for(gpu, b.env, 3, plan, abi.JF_GROUPED) if mode == "jit" else None
        gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, jit=jit)
        H.assert_groups_equal(plan, gk, ga, ok, oa)
    # live non-NULL count check against numpy
    live = ~deleted
    nn = np.bincount(key[live & (val != abi.NULL_I64)], minlength=200)
    ok_keys = ok[:, 0].view(np.int64)
    assert np.array_equal(oa[:, 1].astype(np.int64), nn[ok_keys])


def test_jit_matches_interpreter_two_keys(gpu):
    """Two keys (int64 + int32, the web table's phase-1 pair shape) with an int32 SUM."""
    H.say("case two keys jit vs interp")
    user, region, date, url, dur = H.gen_c2(200_000, n_users=20_000, n_urls=5_000)
    t = H.Table({"user": user, "region": region, "date": date, "url": url, "dur": dur}, 1 << 16)
    t.encode("region")
    b = abi.EnvBuilder()
    su, sr, sd, sl, sdu = (b.slot(abi.T_I64), b.slot(abi.T_I32, enc=2), b.slot(abi.T_I32), b.slot(abi.T_I32),
                           b.slot(abi.T_I32))
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
    slots = [("user", abi.T_I64, 0), ("region", abi.T_I32, 0), ("date", abi.T_I32, 0), ("url", abi.T_I32, 0),
             ("dur", abi.T_I32, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    gi_k, gi_a, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 17)
    H.assert_groups_equal(plan, gi_k, gi_a, ok, oa)
    jit = jit_for(gpu, b.env, len(slots), plan, abi.JF_GROUPED)
    gj_k, gj_a, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 17, jit=jit)
    H.assert_groups_equal(plan, gj_k, gj_a, ok, oa)


def test_run_window_accumulates(gpu):
    """Scan windows: each batch aggregated by its own launch into the running result."""
    H.say("case run windows")
    t, b, plan, slots = c1_setup(120_000, batch_rows=30_000)
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ds = H.dense_spec(plan, [0], [1000])
    gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, dense=ds, windows=True)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host, dense=True)
    H.assert_groups_equal(plan, gk, ga, ok, oa)


def test_dense_private_and_global_modes(gpu):
    """Few groups (per-thread private accumulators) and very many groups (global atomics)."""
    H.say("case dense private/global")
    n = 200_000
    rng = np.random.default_rng(3)
    small = rng.integers(0, 3, n).astype(np.int32)
    big = rng.integers(0, 300_000, n).astype(np.int32)
    val = rng.integers(0, 1000, n).astype(np.int64)
    t = H.Table({"s": small, "g": big, "v": val}, 1 << 16)
    for col, rngv in (("s", 3), ("g", 300_000)):
        b = abi.EnvBuilder()
        sk, sv = b.slot(abi.T_I32), b.slot(abi.T_I64)
        plan = abi.AggPlan()
        plan.nkeys = 1
        plan.keys[0] = b.col(sk)
        plan.key_int[0] = 1
        plan.naggs = 3
        plan.ops[0], plan.ops[1], plan.ops[2] = abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_MAX
        plan.agg_int[0] = plan.agg_int[1] = plan.agg_int[2] = 1
        plan.aggs[1] = plan.aggs[2] = b.col(sv)
        slots = [(col, abi.T_I32, 0), ("v", abi.T_I64, 0)]
        dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
        ds = H.dense_spec(plan, [0], [rngv])
        gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, dense=ds)
        ok, oa, _ = H.oracle_aggregate(plan, b.env, host, dense=True)
        H.assert_groups_equal(plan, gk, ga, ok, oa)


def test_needs_compiled_plan_error(gpu):
    """More than 16 aggregates without a compiled plan: error 1235, as the engine expects."""
    b = abi.EnvBuilder()
    s = b.slot(abi.T_I64)
    plan = abi.AggPlan()
    plan.naggs = 20
    for a in range(20):
        plan.ops[a] = abi.AGG_SUM if a else abi.AGG_COUNT
        plan.aggs[a] = b.col(s)
    h = C.c_void_p()
    ws.check(gpu.ws_agg_create(C.byref(plan), 1, None, C.byref(h)))
    try:
        x = ws.DevBuf.from_array(np.arange(10, dtype=np.int64))
        e = abi.ExprEnv.from_buffer_copy(b.env)
        e.cols[0].p = x.p
        ws.check(gpu.ws_agg_add_batch(h, C.byref(e), 1, 10, None))
        assert gpu.ws_agg_run(h, C.byref(e), None) == ws.WS_ERR_NEEDS_PLAN
    finally:
        gpu.ws_agg_destroy(h)


@pytest.mark.parametrize("case", ["sparse", "dense_filter", "deletes", "out_of_range"])
def test_dense_staged_set_filter(gpu, case):
    """Dense GROUP BY whose WHERE is a set-bitmap test (a join filter fused into the scan): the
    kernel with the set staged in LDS (one 1024-thread block per CU, passing rows dealt out one per
    lane and added a tile later) against the row-at-a-time kernel and the oracle. The 2^21-bit set
    only partly fits the LDS; ragged batches give partial tiles. "dense_filter" passes 90 % of the
    rows (more than 64 per wave and tile: earlier rounds added at once); "deletes" takes the
    row-by-row path; "out_of_range" has keys outside the dense range, reported by both kernels."""
    H.say(f"case dense staged set filter {case}")
    n = 3_000_000
    rng = np.random.default_rng(91)
    big = (1 << 21) + 37
    k = rng.integers(-50, big + 50, n).astype(np.int32)
    g = rng.integers(0, 40, n).astype(np.int32)
    v = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    p = rng.random(n) * 100
    bits = rng.random((big + 31) // 32 * 32) < (0.9 if case == "dense_filter" else 0.05)
    words = np.packbits(bits, bitorder="little").view(np.uint32).copy()
    dset = ws.DevBuf.from_array(words)
    deleted = rng.random(n) < 0.05 if case == "deletes" else None
    t = H.Table({"k": k, "g": g, "v": v, "p": p}, [1_048_576, 5, 0, 1_200_001, 751_418], deleted)
    b = abi.EnvBuilder()
    sk, sg, sv, sp = b.slot(abi.T_I32), b.slot(abi.T_I32), b.slot(abi.T_I64), b.slot(abi.T_F64)
    s0 = b.add_set(dset.p, big)
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sk), b.ilit(3), (abi.OP_SUB,), (abi.OP_INSET, s0)])
    plan.nkeys = 1
    plan.keys[0] = b.col(sg)
    plan.key_int[0] = 1
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX, abi.AGG_SUM, abi.AGG_MIN]
    plan.naggs = len(ops)
    for a, op in enumerate(ops):
        plan.ops[a] = op
    for a in (1, 2, 3, 4):
        plan.aggs[a] = b.col(sv)
        plan.agg_int[a] = 1
    plan.agg_bits[1] = plan.agg_bits[2] = 1
    plan.agg_part[1], plan.agg_part[2] = 1, 2
    plan.agg_int[0] = 1
    plan.aggs[5] = plan.aggs[6] = b.col(sp)
    slots = [("k", abi.T_I32, 0), ("g", abi.T_I32, 0), ("v", abi.T_I64, 0), ("p", abi.T_F64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    ds = H.dense_spec(plan, [0], [30 if case == "out_of_range" else 40])
    jit = jit_for(gpu, b.env, len(slots), plan, abi.JF_DENSE)
    knob = C.c_int64()
    ws.check(gpu.ws_tune_get(b"dense_staged_min_set_words", C.byref(knob)))
    results = {}
    try:
        for kernel, words_min in (("staged", 0), ("plain", 1 << 40)):
            ws.check(gpu.ws_tune_set(b"dense_staged_min_set_words", words_min))
            gpu.ws_prof_reset()
            gpu.ws_prof_only(b"")
            gpu.ws_prof_enable(1)
            try:
                if case == "out_of_range":
                    with pytest.raises(ws.WsError, match="outside the planned key range"):
                        H.gpu_aggregate(plan, b.env, dev, dense=ds, jit=jit)
                else:
                    results[kernel] = H.gpu_aggregate(plan, b.env, dev, dense=ds, jit=jit)
            finally:
                prof = ws.prof_rows()
                gpu.ws_prof_enable(0)
            staged_ran = any(name.startswith("agg_dense_staged") for name in prof)
            assert staged_ran == (kernel == "staged"), prof
    finally:
        ws.check(gpu.ws_tune_set(b"dense_staged_min_set_words", knob.value))
    if case == "out_of_range":
        return
    b.env.sets[s0] = words.ctypes.data  # the oracle reads the set from host memory
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host, dense=True)
    assert len(ok) == 40
    for kernel in ("staged", "plain"):
        gk, ga, _, _ = results[kernel]
        H.assert_groups_equal(plan, gk, ga, ok, oa)
