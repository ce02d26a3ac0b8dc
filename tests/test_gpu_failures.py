"""Failure and fallback paths of the operator surface, each forced from the host:
  - persistent plan cache (reference src/jit.cu:35-92): a compiled plan is written to the cache
    directory, a fresh process-level cache loads it from disk, and a file with a flipped byte is
    discarded and the plan compiled again (jit.cu:67-72), all with identical results;
  - interpret-first mode (JIT_ON, jit.cu:523-543): the first request interprets while the plan
    compiles in the background, later requests get the compiled plan, results identical;
  - codegen failure (jit.cu:433-442): a plan hipRTC rejects (an invalid extra option) yields no
    program, the query runs interpreted with identical results, the plan cache marks it failed;
  - out of device memory (gpu.cu:23-30, 165-174): allocations the device cannot satisfy fail with
    code 1041 and leave the library usable; a VRAM budget refuses columns beyond it with 1041;
  - an aggregation handle run on one thread and fetched / destroyed on another (device blocks
    freed by a thread that did not take them)."""
import ctypes as C
import os

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu


def _query(n=200_000, seed=3, k_lt=500):
    k, v, p = H.gen_c1(n, seed)
    t = H.Table({"k": k, "v": v, "p": p}, [n // 2, n - n // 2])
    b = abi.EnvBuilder()
    sk, sv, sp = b.slot(abi.T_I32), b.slot(abi.T_I64), b.slot(abi.T_F64)
    plan = H.plan_c1(b, sk, sv, sp, k_lt=k_lt)
    slots = [("k", abi.T_I32, 0), ("v", abi.T_I64, 0), ("p", abi.T_F64, 0)]
    return t, b, plan, H.Bound(t, slots, True), H.Bound(t, slots, False)


def _get(gpu, b, plan, mode, kind=b"failure paths"):
    gpu.ws_jit_set_mode(mode)
    out = C.c_void_p()
    ws.check(gpu.ws_jit_get(C.byref(b.env), 3, C.byref(plan), abi.JF_GROUPED, kind, C.byref(out)))
    return out


def _rows(gpu):
    rows = (abi.PlanCacheRow * 256)()
    n = gpu.ws_jit_plan_cache(rows, 256)
    return [rows[i] for i in range(min(n, 256))]


def _run(plan, b, dev, jit):
    gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 11, jit=jit if jit and jit.value else None)
    return H.sort_groups(gk, ga)


@pytest.fixture
def fresh_cache(gpu, tmp_path):
    gpu.ws_jit_clear_cache()
    yield tmp_path
    gpu.ws_jit_set_cache_dir(b"")
    gpu.ws_jit_clear_cache()
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    os.environ.pop("WARPSTORE_HIPRTC_OPTS", None)


def test_plan_cache_disk_reload_and_corrupt_file(gpu, fresh_cache):
    H.say("case plan cache on disk: store, reload, corrupt file")
    t, b, plan, dev, host = _query()
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    gpu.ws_jit_set_cache_dir(str(fresh_cache).encode())
    jit = _get(gpu, b, plan, abi.JIT_SYNC)
    assert jit.value
    assert [r.origin for r in _rows(gpu)] == [b"compiled"]
    files = sorted(p for p in os.listdir(fresh_cache) if p.endswith(".wsplan"))
    assert len(files) == 1
    k1, a1 = _run(plan, b, dev, jit)
    H.assert_groups_equal(plan, k1, a1, ok, oa)
    # a new in-memory cache (as a new process would have) loads the plan from the file
    gpu.ws_jit_clear_cache()
    gpu.ws_jit_begin_query(b"SELECT reload")
    jit = _get(gpu, b, plan, abi.JIT_SYNC)
    info = abi.JitQueryInfo()
    gpu.ws_jit_end_query(0.0, C.byref(info))
    assert jit.value and info.from_disk == 1 and info.compiled == 0
    assert [r.origin for r in _rows(gpu)] == [b"disk"]
    k2, a2 = _run(plan, b, dev, jit)
    assert np.array_equal(k1, k2) and np.array_equal(a1.view(np.int64)[:, :5], a2.view(np.int64)[:, :5])
    # flip one byte of the code object: the file is discarded and the plan compiled again
    path = os.path.join(fresh_cache, files[0])
    raw = bytearray(open(path, "rb").read())
    raw[len(raw) - 100] ^= 0x40
    open(path, "wb").write(bytes(raw))
    gpu.ws_jit_clear_cache()
    jit = _get(gpu, b, plan, abi.JIT_SYNC)
    assert jit.value and [r.origin for r in _rows(gpu)] == [b"compiled"]
    assert open(path, "rb").read() != bytes(raw)  # rewritten with a good image
    k3, a3 = _run(plan, b, dev, jit)
    H.assert_groups_equal(plan, k3, a3, ok, oa)
    # a damaged header (wrong digest) is discarded the same way
    raw = bytearray(open(path, "rb").read())
    raw[20] ^= 0x01
    open(path, "wb").write(bytes(raw))
    gpu.ws_jit_clear_cache()
    jit = _get(gpu, b, plan, abi.JIT_SYNC)
    assert jit.value and [r.origin for r in _rows(gpu)] == [b"compiled"]


def test_interpret_first_then_compiled(gpu, fresh_cache):
    H.say("case JIT_ON: interpret first, compile in the background")
    t, b, plan, dev, host = _query(k_lt=321)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    jit = _get(gpu, b, plan, abi.JIT_ON)
    first_interpreted = not jit.value
    k1, a1 = _run(plan, b, dev, jit)
    H.assert_groups_equal(plan, k1, a1, ok, oa)
    import time
    t0 = time.time()
    while not jit.value and time.time() - t0 < 120:  # the background compile publishes the plan
        time.sleep(0.05)
        jit = _get(gpu, b, plan, abi.JIT_ON)
    assert jit.value, "background compile did not finish"
    k2, a2 = _run(plan, b, dev, jit)
    H.assert_groups_equal(plan, k2, a2, ok, oa)
    assert first_interpreted  # a fresh shape is never compiled inside the request in this mode


def test_codegen_failure_falls_back_to_interpreter(gpu, fresh_cache):
    H.say("case codegen failure -> interpreter")
    t, b, plan, dev, host = _query(k_lt=123)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    os.environ["WARPSTORE_HIPRTC_OPTS"] = "--no-such-hiprtc-option-for-this-test"
    jit = _get(gpu, b, plan, abi.JIT_SYNC)
    assert not jit.value
    rows = _rows(gpu)
    assert len(rows) == 1 and rows[0].kind.endswith(b"(compile failed)") and rows[0].origin == b""
    k1, a1 = _run(plan, b, dev, jit)  # interpreted
    H.assert_groups_equal(plan, k1, a1, ok, oa)
    # a later request of the same shape does not retry the failed compile
    assert not _get(gpu, b, plan, abi.JIT_SYNC).value
    os.environ.pop("WARPSTORE_HIPRTC_OPTS")
    gpu.ws_jit_clear_cache()
    jit = _get(gpu, b, plan, abi.JIT_SYNC)
    assert jit.value


def test_out_of_memory_is_1041_and_recoverable(gpu):
    H.say("case out of device memory -> 1041")
    out = C.c_void_p()
    rc = gpu.ws_tmp_alloc(1 << 50, C.byref(out))
    assert rc == ws.WS_ERR_OOM, (rc, gpu.ws_last_error())
    rc = gpu.ws_dev_alloc(1 << 50, C.byref(out))
    assert rc == ws.WS_ERR_OOM, (rc, gpu.ws_last_error())
    # a dense aggregation whose accumulator table (2^40 groups) cannot be allocated fails the same way
    b = abi.EnvBuilder()
    s = b.slot(abi.T_I32)
    plan = abi.AggPlan()
    plan.nkeys, plan.keys[0], plan.key_int[0] = 1, b.col(s), 1
    plan.naggs, plan.ops[0], plan.agg_int[0] = 1, abi.AGG_COUNT, 1
    ds = H.dense_spec(plan, [0], [1 << 40])
    h = C.c_void_p()
    rc = gpu.ws_agg_create(C.byref(plan), 1 << 11, C.byref(ds), C.byref(h))
    assert rc == ws.WS_ERR_OOM and not h.value, (rc, gpu.ws_last_error())
    # the library is still usable: the same query at a sane capacity runs (the engine's retry)
    t, b2, plan2, dev, host = _query(n=50_000)
    ok, oa, _ = H.oracle_aggregate(plan2, b2.env, host)
    gk, ga, _, _ = H.gpu_aggregate(plan2, b2.env, dev, cap=1 << 11)
    H.assert_groups_equal(plan2, gk, ga, ok, oa)


def test_vram_budget_refuses_columns(gpu):
    H.say("case VRAM budget -> 1041 on placement")
    used = gpu.ws_vram_used()
    gpu.ws_set_vram_limit(used + (1 << 20))
    try:
        a = np.arange(1 << 19, dtype=np.int64)  # 4 MiB: beyond the 1 MiB of budget left
        dev = C.c_void_p()
        on_dev = C.c_int(-1)
        rc = gpu.ws_place_column(ws.ptr(a), a.nbytes, 1, C.byref(on_dev), C.byref(dev))
        assert rc == ws.WS_ERR_OOM, (rc, gpu.ws_last_error())
        assert gpu.ws_vram_used() == used
    finally:
        gpu.ws_set_vram_limit((1 << 63) - 1)


def test_aggregation_handle_moves_between_threads(gpu):
    """A handle built and run on one thread, fetched and destroyed on another: the destroying
    thread's frees are cross-thread frees (the blocks go to the shared pool, whose reuse waits for
    the device), and the blocks are reused correctly by later queries on both threads."""
    H.say("case aggregation handle moves between threads")
    import threading
    t, b, plan, dev, host = _query(n=200_000)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    for _ in range(3):
        box = {}
        th = threading.Thread(target=lambda: box.update(r=H.gpu_aggregate(plan, b.env, dev, cap=1 << 11, fetch=False)))
        th.start()
        th.join()
        h, _cap = box["r"]
        try:
            g = C.c_int64()
            ws.check(gpu.ws_agg_fetch_begin(h, C.byref(g)))
            keys = np.zeros(max(g.value * plan.nkeys, 1))
            accs = np.zeros(max(g.value * plan.naggs, 1))
            ws.check(gpu.ws_agg_fetch_copy(h, ws.ptr(keys), ws.ptr(accs)))
        finally:
            gpu.ws_agg_destroy(h)
        gk = keys[:g.value * plan.nkeys].reshape(g.value, plan.nkeys)
        ga = accs[:g.value * plan.naggs].reshape(g.value, plan.naggs)
        H.assert_groups_equal(plan, gk, ga, ok, oa)
        # this thread takes the freed blocks from the shared pool
        gk, ga, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 11)
        H.assert_groups_equal(plan, gk, ga, ok, oa)
    # a raw block taken on one thread and freed on another keeps its contents until reused
    box = {}
    pat = np.arange(1 << 18, dtype=np.int64)

    def take():
        p = C.c_void_p()
        ws.check(gpu.ws_tmp_alloc(pat.nbytes, C.byref(p)))
        ws.check(gpu.ws_h2d(p, ws.ptr(pat), pat.nbytes))
        box["p"] = p

    th = threading.Thread(target=take)
    th.start()
    th.join()
    back = np.zeros_like(pat)
    ws.check(gpu.ws_d2h(ws.ptr(back), box["p"], pat.nbytes))
    assert np.array_equal(back, pat)
    gpu.ws_tmp_free(box["p"])
    q = C.c_void_p()
    ws.check(gpu.ws_tmp_alloc(pat.nbytes, C.byref(q)))
    rev = pat[::-1].copy()  # held in a name: ws.ptr of a temporary would point at freed memory
    ws.check(gpu.ws_h2d(q, ws.ptr(rev), pat.nbytes))
    ws.check(gpu.ws_d2h(ws.ptr(back), q, pat.nbytes))
    assert np.array_equal(back, rev)
    gpu.ws_tmp_free(q)
    gpu.ws_trim_pool()
