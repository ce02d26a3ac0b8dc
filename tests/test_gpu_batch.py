"""Batched form: many dense filter + GROUP BY queries (the 256-query batch shape: int32 keys with
10 to 1e5 distinct, int64 values, WHERE val > theta_q) in one launch and one fetch, against each
query run alone through the plain path and against the CPU oracle. Integer slots bit-exact."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu


def c4_plan(b, sk, sv, theta):
    """SELECT key, COUNT(*), SUM(val), MIN(val), MAX(val) FROM seg WHERE val > theta GROUP BY key
    (slots: COUNT(*), SUM low / high halves, MIN, MAX; exact int64 bits)."""
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sv), b.ilit(theta), (abi.OP_GT,)])
    plan.nkeys = 1
    plan.keys[0] = b.col(sk)
    plan.key_int[0] = 1
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX]
    plan.naggs = len(ops)
    for a, o in enumerate(ops):
        plan.ops[a] = o
    plan.agg_int[0] = 1
    for a in (1, 2, 3, 4):
        plan.aggs[a] = b.col(sv)
        plan.agg_int[a] = 1
        plan.agg_bits[a] = 1
    plan.agg_part[1], plan.agg_part[2] = 1, 2
    return plan


def make_queries(nq, seed, max_rows=300_000):
    rng = np.random.default_rng(seed)
    qs = []
    for q in range(nq):
        n = int(np.clip(round(rng.lognormal(np.log(40_000), 1.25)), 4096, max_rows))
        d = int(round(10 ** rng.uniform(1, 5)))
        key = rng.integers(0, d, n).astype(np.int32)
        val = rng.integers(0, 1_000_000, n).astype(np.int64)
        theta = int(rng.uniform(1e5, 9e5)) if q % 7 else 2_000_000  # every 7th query selects nothing
        deleted = rng.random(n) < 0.05 if q % 3 == 0 else None
        t = H.Table({"key": key, "val": val}, [n // 2, n - n // 2] if q % 2 else n, deleted)
        b = abi.EnvBuilder()
        sk, sv = b.slot(abi.T_I32), b.slot(abi.T_I64)
        plan = c4_plan(b, sk, sv, theta)
        ds = H.dense_spec(plan, [0], [d])
        slots = [("key", abi.T_I32, 0), ("val", abi.T_I64, 0)]
        qs.append(dict(t=t, b=b, plan=plan, ds=ds, dev=H.Bound(t, slots, True), host=H.Bound(t, slots, False)))
    return qs


@pytest.mark.parametrize("ranges", [0, 64])
@pytest.mark.parametrize("mode", ["interp", "jit"])
def test_batched_dense_equals_alone_and_oracle(gpu, mode, ranges):
    """ranges: the batch_key_ranges knob. 0 sends the groups of queries with more groups than the
    kernel's LDS table to device-scope atomics; 64 reads those queries' rows once per table-sized
    key range (1e5 keys over 512-slot tables is 196 ranges, so the largest queries still take the
    atomics)."""
    H.say(f"case batched dense {mode} key ranges {ranges}")
    L = gpu
    saved = C.c_int64()
    ws.check(L.ws_tune_get(b"batch_key_ranges", C.byref(saved)))
    ws.check(L.ws_tune_set(b"batch_key_ranges", ranges))
    try:
        _batched_case(L, mode)
    finally:
        ws.check(L.ws_tune_set(b"batch_key_ranges", saved.value))


def _batched_case(L, mode):
    qs = make_queries(40, seed=44)
    jit = None
    if mode == "jit":
        L.ws_jit_set_mode(abi.JIT_SYNC)
        jit = C.c_void_p()
        q0 = qs[0]
        ws.check(L.ws_jit_get(C.byref(q0["b"].env), 2, C.byref(q0["plan"]), abi.JF_DENSE, b"batch", C.byref(jit)))
        assert jit.value
    aggs = (C.c_void_p * len(qs))()
    envs = (C.c_void_p * len(qs))()
    keep = []
    try:
        for i, q in enumerate(qs):
            h = C.c_void_p()
            ws.check(L.ws_agg_create(C.byref(q["plan"]), 1, C.byref(q["ds"]), C.byref(h)))
            aggs[i] = h.value
            for bi in range(len(q["t"].sizes)):
                e = H.env_with_batch(q["b"].env, q["dev"], bi)
                ws.check(L.ws_agg_add_batch(h, C.byref(e), 2, q["t"].sizes[bi], q["dev"].dels[bi]))
            env = abi.ExprEnv.from_buffer_copy(q["b"].env)
            keep.append(env)
            envs[i] = C.addressof(env)
        ws.check(L.ws_agg_run_batch(len(qs), aggs, envs, jit))
        counts = np.zeros(len(qs), np.int64)
        total = C.c_int64()
        cap = sum(int(q["ds"].G) for q in qs)
        keys = np.zeros(cap)
        accs = np.zeros(cap * 5)
        ws.check(L.ws_agg_fetch_many(len(qs), aggs, ws.ptr(counts), C.byref(total), cap, ws.ptr(keys), ws.ptr(accs)))
    finally:
        for i in range(len(qs)):
            if aggs[i]:
                L.ws_agg_destroy(C.c_void_p(aggs[i]))
    off = 0
    for i, q in enumerate(qs):
        g = int(counts[i])
        bk = keys[off:off + g].reshape(g, 1)
        ba = accs[off * 5:(off + g) * 5].reshape(g, 5)
        off += g
        ak, aa, _, _ = H.gpu_aggregate(q["plan"], q["b"].env, q["dev"], dense=q["ds"], jit=jit)  # alone, plain path
        H.assert_groups_equal(q["plan"], bk, ba, *H.sort_groups(ak, aa))
        if i % 5 == 0:
            ok, oa, _ = H.oracle_aggregate(q["plan"], q["b"].env, q["host"], dense=True)
            H.assert_groups_equal(q["plan"], bk, ba, ok, oa)
        if i % 7 == 0:
            assert g == 0
    assert off == total.value
