"""Serving forms: the many-small-aggregations workload batched into one launch equals the plain
per-query path query by query, and in the continuous-batching serving loop every request's result
is the same as when it runs alone (small aggregates, web COUNT DISTINCT queries, star joins)."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
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


def test_autotuner_choices_keep_results(gpu):
    """The autotuner times its candidates on the workload itself and records a choice per shape;
    the batched results under the chosen knobs equal the per-query results, and the web query's
    result is the same before and after tuning its partition knobs."""
    H.say("case autotuner (batched GROUP BY and partitioned web query)")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    s = Q.C4Set(seed=14, nq=24)
    try:
        dp, gp = s.run(batched=False)
        r = s.autotune(reps=1)
        assert r["trials"] > 5 and r["tuned_ms"] <= r["start_ms"] * 1.0001
        assert r["chosen"]["batch_tile_rows"] in (1024, 2048, 4096, 8192, 16384, 32768)
        db, gb = s.run(batched=True)
        assert np.array_equal(db, dp) and np.array_equal(gb, gp)
    finally:
        s.free()
    user, region, date, url, dur = H.gen_c2(1_000_000)
    web = Q.DeviceTable([user, region, date, url, dur], [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32],
                        [0, 1, 0, 0, 0])
    try:
        before, st0 = Q.c2(web, 10)
        r = Q.autotune_web(web, reps=1)
        assert r["trials"] > 3
        after, st1 = Q.c2(web, 10)
        assert np.array_equal(st0[:2], st1[:2])
        H.check_topn(after, H.oracle_c2(user, region, date, url, dur))
    finally:
        web.free()
        for k, v in (("batch_tile_rows", 4096), ("batch_lds_slots", 1024), ("batch_blocks_per_cu", 4), ("part_chunks_per_cu", 1),
                     ("part_scatter_lds_kib", 120), ("part_merge_lds_kib", 60)):
            gpu.ws_tune_set(k.encode(), v)


def test_prepared_batch_graph_replays(gpu):
    """A batch prepared once (its launches recorded in a HIP graph) and replayed three times, with
    the program's own eager work and synchronizations between the replays: every replay's results
    equal the per-query path's (the tables start from their initial values each time)."""
    H.say("case prepared batch: graph replays")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    s = Q.C4Set(seed=21, nq=40)
    pr = None
    try:
        dp, gp = s.run(batched=False)
        pr = s.prepare()
        for rep in range(3):
            dr, gr = pr.run()
            assert np.array_equal(dr, dp) and np.array_equal(gr, gp), rep
            db, gb = s.run(batched=True)  # eager work between replays
            assert np.array_equal(db, dp)
            gpu.ws_sync()
    finally:
        if pr is not None:
            pr.free()
        s.free()


def test_prepared_batch_pipeline(gpu):
    """Two, then three, prepared copies of one set take turns in a pipeline (each batch's groups
    copied to the host while the later batches' kernels run): the last batch of 1 to 7 steps equals
    the per-query path, and so does a plain replay of every copy afterwards. The halves' misuse (a finish with no
    launch, a second launch before the finish) fails with an error and leaves the plan usable."""
    H.say("case prepared batch: pipelined launch / finish")
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    s = Q.C4Set(seed=23, nq=40)
    a = b = c = None
    try:
        dp, gp = s.run(batched=False)
        a, b, c = s.prepare(), s.prepare(), s.prepare()
        for steps in (1, 2, 5):
            d, g = a.pipeline(b, steps)
            assert np.array_equal(d, dp) and np.array_equal(g, gp), steps
        for steps in (1, 2, 3, 7):  # three copies: two batches in flight during each copy-out
            d, g = a.pipeline([b, c], steps)
            assert np.array_equal(d, dp) and np.array_equal(g, gp), ("three copies", steps)
        for steps, depth in ((1, 3), (4, 3), (3, 1)):  # every batch built from scratch
            d, g = s.fresh_pipeline(steps, depth)
            assert np.array_equal(d, dp) and np.array_equal(g, gp), ("fresh", steps, depth)
        for p in (a, b, c):
            d, g = p.run()
            assert np.array_equal(d, dp) and np.array_equal(g, gp)
        counts = np.zeros(40, np.int64)
        total = C.c_int64()
        assert gpu.ws_agg_batch_finish(C.c_void_p(_plan_of(a)), ws.ptr(counts), C.byref(total), 0, None, None) != 0
        assert gpu.ws_agg_batch_launch(C.c_void_p(_plan_of(a))) == 0
        assert gpu.ws_agg_batch_launch(C.c_void_p(_plan_of(a))) != 0
        cap = int(gp.sum())
        keys, accs = np.zeros(cap), np.zeros(cap * 5)
        ws.check(gpu.ws_agg_batch_finish(C.c_void_p(_plan_of(a)), ws.ptr(counts), C.byref(total), cap, ws.ptr(keys),
                                         ws.ptr(accs)))
        assert total.value == cap and np.array_equal(counts, gp)
        d, g = a.run()
        assert np.array_equal(d, dp)
    finally:
        for p in (a, b, c):
            if p is not None:
                p.free()
        s.free()


def _plan_of(prepared) -> int:
    """The ws_agg_batch_plan of a prepared C4 set."""
    return Q.qlib().wsq_c4_plan_handle(prepared.h)
