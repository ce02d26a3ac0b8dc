"""GPU parity of selection / compaction, hash joins, sorting, gathers and code remaps against the
CPU oracle (bit-exact: these are integer / index results)."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu


def colview(buf, t):
    v = abi.ColView()
    v.p, v.t = buf.p, t
    return v


def hcolview(arr, t):
    v = abi.ColView()
    v.p, v.t = arr.ctypes.data, t
    return v


@pytest.mark.parametrize("mode", ["interp", "jit"])
def test_select_multi(gpu, mode):
    H.say(f"case select_multi {mode}")
    n = 250_000
    rng = np.random.default_rng(7)
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    x = rng.random(n)
    x[rng.random(n) < 0.03] = np.nan
    deleted = rng.random(n) < 0.1
    t = H.Table({"a": a, "x": x}, [70_000, 0, 1, 179_999], deleted)
    b = abi.EnvBuilder()
    sa, sx = b.slot(abi.T_I32), b.slot(abi.T_F64)
    where = b.prog([(abi.OP_COL, sa), b.ilit(100), (abi.OP_GT,), (abi.OP_COL, sx), b.dlit(0.5), (abi.OP_LT,), (abi.OP_OR,)])
    slots = [("a", abi.T_I32, 0), ("x", abi.T_F64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    nb = len(t.sizes)
    bn = np.array(t.sizes, np.int64)
    jit = None
    if mode == "jit":
        gpu.ws_jit_set_mode(abi.JIT_SYNC)
        fp = abi.AggPlan()
        fp.has_where, fp.where = 1, where
        jit = C.c_void_p()
        ws.check(gpu.ws_jit_get(C.byref(b.env), 2, C.byref(fp), abi.JF_FLAGM, b"filter", C.byref(jit)))
        assert jit.value
    out = C.c_void_p()
    per = np.zeros(nb, np.int64)
    cnt = C.c_int64()
    ws.check(gpu.ws_select_multi(C.byref(b.env), 1, where, nb, ws.ptr(bn), C.cast(dev.dels, C.c_void_p),
                                 C.cast(dev.views, C.c_void_p), 2, C.byref(out), ws.ptr(per), jit, C.byref(cnt)))
    got = ws.DevBuf(cnt.value * 4, tmp=True, adopt=out.value).to_array(np.int32, cnt.value)
    exp = np.zeros(n, np.int32)
    eper = np.zeros(nb, np.int64)
    m = H.oracle().or_select(C.byref(b.env), 1, where, nb, ws.ptr(bn), C.cast(host.dels, C.c_void_p),
                             C.cast(host.views, C.c_void_p), 2, ws.ptr(exp), ws.ptr(eper))
    assert cnt.value == m
    assert np.array_equal(got, exp[:m])
    assert np.array_equal(per, eper)


@pytest.mark.parametrize("mode", ["interp", "jit"])
def test_select_multi_chunked(gpu, mode):
    """The chunked selection kernel (one look-back per chunk of tiles), forced by the tuning knobs:
    2.6M rows in ragged batches with deletes, against the oracle."""
    H.say(f"case select_multi chunked {mode}")
    n = 2_600_000
    rng = np.random.default_rng(71)
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    x = rng.random(n)
    x[rng.random(n) < 0.03] = np.nan
    deleted = rng.random(n) < 0.1
    t = H.Table({"a": a, "x": x}, [1_048_576, 3, 0, 700_001, 851_420], deleted)
    b = abi.EnvBuilder()
    sa, sx = b.slot(abi.T_I32), b.slot(abi.T_F64)
    where = b.prog([(abi.OP_COL, sa), b.ilit(990), (abi.OP_GT,), (abi.OP_COL, sx), b.dlit(0.002), (abi.OP_LT,), (abi.OP_OR,)])
    slots = [("a", abi.T_I32, 0), ("x", abi.T_F64, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    nb = len(t.sizes)
    bn = np.array(t.sizes, np.int64)
    jit = None
    if mode == "jit":
        gpu.ws_jit_set_mode(abi.JIT_SYNC)
        fp = abi.AggPlan()
        fp.has_where, fp.where = 1, where
        jit = C.c_void_p()
        ws.check(gpu.ws_jit_get(C.byref(b.env), 2, C.byref(fp), abi.JF_FLAGM, b"filter", C.byref(jit)))
        assert jit.value
    out = C.c_void_p()
    per = np.zeros(nb, np.int64)
    cnt = C.c_int64()
    # one chunk per CU: 2,540 tiles are above 2 x 256 chunks, so the chunked kernel runs (a compiled
    # plan only when the ballot pass is held off)
    knobs = {"select_chunks_per_cu": 1, "select_mask_min_rows": 1 << 40}
    saved = {}
    for key, val in knobs.items():
        v = C.c_int64()
        ws.check(gpu.ws_tune_get(key.encode(), C.byref(v)))
        saved[key] = v.value
        ws.check(gpu.ws_tune_set(key.encode(), val))
    try:
        ws.check(gpu.ws_select_multi(C.byref(b.env), 1, where, nb, ws.ptr(bn), C.cast(dev.dels, C.c_void_p),
                                     C.cast(dev.views, C.c_void_p), 2, C.byref(out), ws.ptr(per), jit, C.byref(cnt)))
    finally:
        for key, val in saved.items():
            ws.check(gpu.ws_tune_set(key.encode(), val))
    got = ws.DevBuf(cnt.value * 4, tmp=True, adopt=out.value).to_array(np.int32, cnt.value)
    exp = np.zeros(n, np.int32)
    eper = np.zeros(nb, np.int64)
    m = H.oracle().or_select(C.byref(b.env), 1, where, nb, ws.ptr(bn), C.cast(host.dels, C.c_void_p),
                             C.cast(host.views, C.c_void_p), 2, ws.ptr(exp), ws.ptr(eper))
    assert cnt.value == m and m > 1000
    assert np.array_equal(got, exp[:m])
    assert np.array_equal(per, eper)


@pytest.mark.parametrize("case", ["partial", "full", "empty", "partial_nodel"])
def test_select_staged_sets(gpu, case):
    """Compiled selections with set bitmaps: the ballot pass stages the sets' leading words in LDS
    and reads the rest from global memory. A 2^21-bit set only partly fits ("partial"), a 5000-bit
    one fully ("full"), an empty set matches nothing ("empty"); 3M rows in ragged batches with
    deletes (without: full tiles take the pipelined path), against the oracle and the single-pass
    kernel."""
    H.say(f"case select staged sets {case}")
    n = 3_000_000
    rng = np.random.default_rng(72)
    big = (1 << 21) + 37
    k = rng.integers(-50, big + 50, n).astype(np.int32)
    a = rng.integers(-10, 6000, n).astype(np.int32)
    deleted = None if case == "partial_nodel" else rng.random(n) < 0.05  # no deletes: the pipelined path
    t = H.Table({"k": k, "a": a}, [1_048_576, 5, 0, 1_200_001, 751_418], deleted)
    sizes = {"partial": (big, 5000), "full": (0, 5000), "empty": (0, 0), "partial_nodel": (big, 5000)}[case]
    bits = [rng.random((s + 31) // 32 * 32) < 0.3 for s in sizes]
    words = [np.packbits(bb, bitorder="little").view(np.uint32).copy() if bb.size else np.zeros(1, np.uint32)
             for bb in bits]
    dsets = [ws.DevBuf.from_array(w) for w in words]
    b = abi.EnvBuilder()
    sk, sa = b.slot(abi.T_I32), b.slot(abi.T_I32)
    s0 = b.add_set(dsets[0].p if sizes[0] else 0, sizes[0])
    s1 = b.add_set(dsets[1].p if sizes[1] else 0, sizes[1])
    where = b.prog([(abi.OP_COL, sk), b.ilit(3), (abi.OP_SUB,), (abi.OP_INSET, s0), (abi.OP_COL, sa), (abi.OP_INSET, s1),
                    (abi.OP_OR,)])
    slots = [("k", abi.T_I32, 0), ("a", abi.T_I32, 0)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    nb = len(t.sizes)
    bn = np.array(t.sizes, np.int64)
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    fp = abi.AggPlan()
    fp.has_where, fp.where = 1, where
    jit = C.c_void_p()
    ws.check(gpu.ws_jit_get(C.byref(b.env), 2, C.byref(fp), abi.JF_FLAGM, b"filter", C.byref(jit)))
    assert jit.value

    def run():
        out = C.c_void_p()
        per = np.zeros(nb, np.int64)
        cnt = C.c_int64()
        ws.check(gpu.ws_select_multi(C.byref(b.env), 1, where, nb, ws.ptr(bn), C.cast(dev.dels, C.c_void_p),
                                     C.cast(dev.views, C.c_void_p), 2, C.byref(out), ws.ptr(per), jit, C.byref(cnt)))
        return ws.DevBuf(cnt.value * 4, tmp=True, adopt=out.value).to_array(np.int32, cnt.value), per

    v = C.c_int64()
    ws.check(gpu.ws_tune_get(b"select_mask_min_rows", C.byref(v)))
    try:
        ws.check(gpu.ws_tune_set(b"select_mask_min_rows", 1 << 40))  # the chunked single-pass kernel
        plain, plain_per = run()
        ws.check(gpu.ws_tune_set(b"select_mask_min_rows", 0))  # the ballot pass with staged sets
        got, per = run()
    finally:
        ws.check(gpu.ws_tune_set(b"select_mask_min_rows", v.value))
    hsets = [w.ctypes.data for w in words]
    b.env.sets[s0], b.env.sets[s1] = (hsets[0] if sizes[0] else 0), (hsets[1] if sizes[1] else 0)
    exp = np.zeros(n, np.int32)
    eper = np.zeros(nb, np.int64)
    m = H.oracle().or_select(C.byref(b.env), 1, where, nb, ws.ptr(bn), C.cast(host.dels, C.c_void_p),
                             C.cast(host.views, C.c_void_p), 2, ws.ptr(exp), ws.ptr(eper))
    assert len(got) == m and (m > 1000 or case == "empty")
    assert np.array_equal(got, exp[:m]) and np.array_equal(per, eper)
    assert np.array_equal(plain, exp[:m]) and np.array_equal(plain_per, eper)


def test_select_single_env_cols(gpu):
    H.say("case select single")
    n = 100_003
    v = np.arange(n, dtype=np.int64) * 7 % 1000
    dv = ws.DevBuf.from_array(v)
    b = abi.EnvBuilder()
    s = b.slot(abi.T_I64)
    b.env.cols[s].p = dv.p
    where = b.prog([(abi.OP_COL, s), b.ilit(500), (abi.OP_GE,)])
    out = ws.DevBuf(n * 4)
    cnt = C.c_int64()
    ws.check(gpu.ws_select(C.byref(b.env), 1, where, n, None, out.p, None, C.byref(cnt)))
    got = out.to_array(np.int32, cnt.value)
    assert np.array_equal(got, np.nonzero(v >= 500)[0].astype(np.int32))


def run_join(L, bkey, nb, pkey, np_, outer=0):
    ob, op, m = C.c_void_p(), C.c_void_p(), C.c_int64()
    if outer:
        ws.check(L.ws_hash_join_outer(bkey, nb, pkey, np_, C.byref(ob), C.byref(op), 1 if outer == 2 else 0, C.byref(m)))
    else:
        ws.check(L.ws_hash_join(bkey, nb, pkey, np_, C.byref(ob), C.byref(op), C.byref(m)))
    if m.value == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    b = ws.DevBuf(m.value * 4, tmp=True, adopt=ob.value).to_array(np.int32, m.value)
    p = ws.DevBuf(m.value * 4, tmp=True, adopt=op.value).to_array(np.int32, m.value)
    return b, p


def oracle_join(bk, bt, pk, pt, outer=0):
    cap = len(bk) * len(pk) + len(pk) + len(bk) + 1
    ob = np.zeros(min(cap, 10_000_000), np.int32)
    op = np.zeros_like(ob)
    m = H.oracle().or_join(C.byref(hcolview(bk, bt)), len(bk), C.byref(hcolview(pk, pt)), len(pk), outer, ws.ptr(ob),
                           ws.ptr(op), len(ob))
    assert m >= 0
    return ob[:m], op[:m]


@pytest.mark.parametrize("outer", [0, 1, 2])
def test_hash_join(gpu, outer):
    H.say(f"case join outer={outer}")
    rng = np.random.default_rng(13)
    bk = rng.integers(0, 5000, 20_000).astype(np.int64)  # duplicate build keys
    bk[::97] = abi.NULL_I64
    pk = rng.integers(-100, 6000, 300_000).astype(np.int64)
    pk[::101] = abi.NULL_I64
    db, dp = ws.DevBuf.from_array(bk), ws.DevBuf.from_array(pk)
    gb, gp = run_join(gpu, colview(db, abi.T_I64N), len(bk), colview(dp, abi.T_I64N), len(pk), outer)
    eb, ep = oracle_join(bk, abi.T_I64N, pk, abi.T_I64N, outer)
    # probe order: the GPU output is ascending in the probe index (chain order inside a row)
    if outer < 2:
        assert np.all(np.diff(gp) >= 0)
    order = np.lexsort((gb, gp))
    order_e = np.lexsort((eb, ep))
    assert np.array_equal(gp[order], ep[order_e]) and np.array_equal(gb[order], eb[order_e])


@pytest.mark.parametrize("outer", [0, 1, 2])
def test_hash_join_unique_keys(gpu, outer):
    """Unique build keys take the single-probe form (one probe per row, no look-back)."""
    H.say(f"case join unique outer={outer}")
    rng = np.random.default_rng(14)
    bk = rng.permutation(20_000)[:15_000].astype(np.int64)
    bk[::97] = abi.NULL_I64
    pk = rng.integers(-100, 21_000, 300_000).astype(np.int64)
    pk[::101] = abi.NULL_I64
    db, dp = ws.DevBuf.from_array(bk), ws.DevBuf.from_array(pk)
    gb, gp = run_join(gpu, colview(db, abi.T_I64N), len(bk), colview(dp, abi.T_I64N), len(pk), outer)
    eb, ep = oracle_join(bk, abi.T_I64N, pk, abi.T_I64N, outer)
    if outer < 2:
        assert np.all(np.diff(gp) >= 0)
    order = np.lexsort((gb, gp))
    order_e = np.lexsort((eb, ep))
    assert np.array_equal(gp[order], ep[order_e]) and np.array_equal(gb[order], eb[order_e])


@pytest.mark.parametrize("outer", [0, 1, 2])
@pytest.mark.parametrize("sizes", [(0, 5000), (5000, 0), (0, 0)])
def test_hash_join_empty_side(gpu, sizes, outer):
    """An empty build or probe side: an inner join matches nothing, a left join keeps every probe
    row (build index -1), a full join also every build row (probe index -1)."""
    nb, np_ = sizes
    H.say(f"case join empty side nb={nb} np={np_} outer={outer}")
    rng = np.random.default_rng(19)
    bk = rng.integers(0, 100, nb).astype(np.int64)
    pk = rng.integers(0, 100, np_).astype(np.int64)
    # a one-element device buffer stands behind an empty side (its pointer is valid, never read)
    db = ws.DevBuf.from_array(bk if nb else np.zeros(1, np.int64))
    dp = ws.DevBuf.from_array(pk if np_ else np.zeros(1, np.int64))
    gb, gp = run_join(gpu, colview(db, abi.T_I64), nb, colview(dp, abi.T_I64), np_, outer)
    eb, ep = oracle_join(bk, abi.T_I64, pk, abi.T_I64, outer)
    expect = 0 if outer == 0 else np_ if nb == 0 else nb if outer == 2 else 0
    assert len(gb) == len(eb) == expect
    order = np.lexsort((gb, gp))
    order_e = np.lexsort((eb, ep))
    assert np.array_equal(gp[order], ep[order_e]) and np.array_equal(gb[order], eb[order_e])


@pytest.mark.parametrize("unique", [False, True])
def test_hash_join_parts_offsets(gpu, unique):
    H.say(f"case join parts unique={unique}")
    rng = np.random.default_rng(17)
    P = 4
    bsz = [3000, 0, 2500, 4000]
    psz = [50_000, 20_000, 0, 33_333]
    draw = (lambda s: rng.permutation(5000)[:s]) if unique else (lambda s: rng.integers(0, 2000, s))
    bk = np.concatenate([draw(s) + 10_000 * i for i, s in enumerate(bsz)]).astype(np.int32)
    pk = np.concatenate([rng.integers(0, 5500 if unique else 2500, s) + 10_000 * i for i, s in enumerate(psz)]).astype(np.int32)
    boff = np.concatenate([[0], np.cumsum(bsz)]).astype(np.int64)
    poff = np.concatenate([[0], np.cumsum(psz)]).astype(np.int64)
    db, dp = ws.DevBuf.from_array(bk), ws.DevBuf.from_array(pk)
    ob, op, m = C.c_void_p(), C.c_void_p(), C.c_int64()
    off = np.zeros(P + 1, np.int64)
    ws.check(gpu.ws_hash_join_parts(colview(db, abi.T_I32), ws.ptr(boff), P, colview(dp, abi.T_I32), ws.ptr(poff), P,
                                    C.byref(ob), C.byref(op), ws.ptr(off), C.byref(m)))
    gb = ws.DevBuf(m.value * 4, tmp=True, adopt=ob.value).to_array(np.int32, m.value)
    gp = ws.DevBuf(m.value * 4, tmp=True, adopt=op.value).to_array(np.int32, m.value)
    eb, ep = oracle_join(bk, abi.T_I32, pk, abi.T_I32)  # keys are partition-disjoint: same pair set
    assert m.value == len(eb)
    order = np.lexsort((gb, gp))
    order_e = np.lexsort((eb, ep))
    assert np.array_equal(gp[order], ep[order_e]) and np.array_equal(gb[order], eb[order_e])
    for p in range(P + 1):  # out_off[p] = first output pair of probe partition p
        assert off[p] == np.searchsorted(gp, poff[p]) if p < P else off[p] == m.value


@pytest.mark.parametrize("m", [1, 1000, 300_001])
def test_sort_perm(gpu, m):
    H.say(f"case sort m={m}")
    rng = np.random.default_rng(m)
    k0 = rng.integers(0, 50, m).astype(np.float64)
    k0[rng.random(m) < 0.05] = np.nan
    k1 = rng.standard_normal(m)
    k1[rng.random(m) < 0.05] = -0.0
    k1[rng.random(m) < 0.05] = 0.0
    k2 = np.floor(rng.random(m) * 3)
    keys = [k0, k1, k2]
    desc = np.array([1, 0, 1], np.uint8)
    dks = [ws.DevBuf.from_array(k) for k in keys]
    kp = (C.c_void_p * 3)(*[d.p for d in dks])
    perm = C.c_void_p()
    ws.check(gpu.ws_sort_perm(kp, ws.ptr(desc), 3, m, C.byref(perm)))
    got = ws.DevBuf(m * 4, tmp=True, adopt=perm.value).to_array(np.int32, m)
    exp = np.zeros(m, np.int32)
    hp = (C.c_void_p * 3)(*[k.ctypes.data for k in keys])
    H.oracle().or_sort_perm(hp, ws.ptr(desc), 3, m, ws.ptr(exp))
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("m,k,nk", [(476_610, 10, 1), (2_000_000, 16, 1), (100_000, 1, 3), (50_000, 256, 2),
                                    (300_001, 257, 1), (70_000, 100, 4), (70_000, 12, 5), (9_000, 1_000, 1),
                                    (5_000, 7, 2), (3, 10, 1), (1_000_003, 64, 1), (100_000, 1, 1),
                                    (40_000, 13, 1), (10_000_000, 10, 1)])
def test_topk_perm_is_sort_prefix(gpu, m, k, nk):
    """ORDER BY ... LIMIT k: the first k of the stable multi-key sort, ties (few distinct values),
    NULLs (NaN), -0.0 and descending keys included; selection (up to 4 keys, k <= 256) and the
    full-sort fallbacks (k > 256, 5 keys, k not much below m)."""
    H.say(f"case topk m={m} k={k} nk={nk}")
    rng = np.random.default_rng(m + k + nk)
    keys = []
    for j in range(nk):
        if j == 0:
            v = rng.integers(0, 40, m).astype(np.float64)  # heavy ties on the first key
        elif j % 2:
            v = rng.standard_normal(m)
            v[rng.random(m) < 0.05] = -0.0
            v[rng.random(m) < 0.05] = 0.0
        else:
            v = np.floor(rng.random(m) * 3)
        v[rng.random(m) < 0.03] = np.nan
        keys.append(v)
    desc = np.array([(1 + j) % 2 for j in range(nk)], np.uint8)
    dks = [ws.DevBuf.from_array(v) for v in keys]
    kp = (C.c_void_p * nk)(*[d.p for d in dks])
    perm = C.c_void_p()
    ws.check(gpu.ws_topk_perm(kp, ws.ptr(desc), nk, m, k, C.byref(perm)))
    kk = min(k, m)
    got = ws.DevBuf(kk * 4, tmp=True, adopt=perm.value).to_array(np.int32, kk)
    exp = np.zeros(m, np.int32)
    hp = (C.c_void_p * nk)(*[v.ctypes.data for v in keys])
    H.oracle().or_sort_perm(hp, ws.ptr(desc), nk, m, ws.ptr(exp))
    assert np.array_equal(got, exp[:kk])


@pytest.mark.parametrize("order,k,nk", [("rising", 10, 1), ("rising", 200, 1), ("rising", 33, 2), ("falling", 10, 1),
                                         ("sawtooth", 64, 1), ("constant", 10, 1)])
def test_topk_orders_that_defeat_the_threshold(gpu, order, k, nk):
    """ORDER BY ... DESC LIMIT k over inputs where every step brings many new candidates (keys
    rising in row order: each row beats the kept k; a sawtooth of short rising runs; all keys equal,
    so the row index decides), so the selection sorts its steps' new entries instead of ranking them
    one by one; and keys falling, where steps bring none. Against the oracle's stable sort."""
    H.say(f"case topk order={order} k={k} nk={nk}")
    m = 300_000
    rng = np.random.default_rng(k + nk)
    i = np.arange(m, dtype=np.float64)
    first = {"rising": i, "falling": -i, "sawtooth": i % 700, "constant": np.zeros(m)}[order]
    keys = [first] + [np.floor(rng.random(m) * 5) for _ in range(nk - 1)]
    desc = np.ones(nk, np.uint8)
    dks = [ws.DevBuf.from_array(v) for v in keys]
    kp = (C.c_void_p * nk)(*[d.p for d in dks])
    perm = C.c_void_p()
    ws.check(gpu.ws_topk_perm(kp, ws.ptr(desc), nk, m, k, C.byref(perm)))
    got = ws.DevBuf(k * 4, tmp=True, adopt=perm.value).to_array(np.int32, k)
    exp = np.zeros(m, np.int32)
    hp = (C.c_void_p * nk)(*[v.ctypes.data for v in keys])
    H.oracle().or_sort_perm(hp, ws.ptr(desc), nk, m, ws.ptr(exp))
    assert np.array_equal(got, exp[:k])


@pytest.mark.parametrize("ptype,dtype", [(abi.P_I32, np.int32), (abi.P_I64, np.int64)])
def test_gather_multi_raw(gpu, ptype, dtype):
    """Gather across raw batches (the LDS-staged kernel): ragged batches, sorted and unsorted
    indexes, a count that does not fill the last step."""
    H.say(f"case gather_multi raw {dtype.__name__}")
    rng = np.random.default_rng(23)
    sizes = [70_001, 3, 0, 130_000, 99_999]
    n = sum(sizes)
    col = rng.integers(-2 ** 30, 2 ** 30, n).astype(dtype)
    t = H.Table({"c": col}, sizes)
    bound = H.Bound(t, [("c", abi.T_I32 if dtype == np.int32 else abi.T_I64, 0)], True)
    boff = np.array(t.offs[:-1], np.int64)
    for idx in (np.sort(rng.choice(n, 123_457, replace=False)).astype(np.int32), rng.integers(0, n, 77_777).astype(np.int32)):
        di = ws.DevBuf.from_array(idx)
        out = ws.DevBuf(len(idx) * col.itemsize)
        ws.check(gpu.ws_gather_multi(C.cast(bound.views, C.c_void_p), ws.ptr(boff), len(sizes), ptype, col.itemsize, di.p,
                                     len(idx), out.p))
        assert np.array_equal(out.to_array(dtype), col[idx])


def test_gathers_and_codes(gpu):
    H.say("case gathers")
    rng = np.random.default_rng(19)
    n = 100_000
    src = rng.integers(-2 ** 40, 2 ** 40, n).astype(np.int64)
    idx = rng.integers(0, n, 50_000).astype(np.int32)
    ds, di = ws.DevBuf.from_array(src), ws.DevBuf.from_array(idx)
    out = ws.DevBuf(len(idx) * 8)
    ws.check(gpu.ws_gather(ds.p, abi.P_I64, di.p, len(idx), out.p))
    assert np.array_equal(out.to_array(np.int64), src[idx])
    idxn = idx.copy()
    idxn[::3] = -1
    din = ws.DevBuf.from_array(idxn)
    ws.check(gpu.ws_gather_null(ds.p, abi.P_I64, din.p, len(idx), out.p))
    exp = np.where(idxn < 0, abi.NULL_I64, src[np.maximum(idxn, 0)])
    assert np.array_equal(out.to_array(np.int64), exp)
    # encoded gather (u16 frame of reference) and gather across batches
    small = rng.integers(1000, 1000 + 60_000, n).astype(np.int32)
    t = H.Table({"s": small}, [40_000, 60_000])
    t.encode("s")
    bound = H.Bound(t, [("s", abi.T_I32, 0)], True)
    boff = np.array([0, 40_000], np.int64)
    out32 = ws.DevBuf(len(idx) * 4)
    ws.check(gpu.ws_gather_multi(C.cast(bound.views, C.c_void_p), ws.ptr(boff), 2, abi.P_I32, 4, di.p, len(idx), out32.p))
    assert np.array_equal(out32.to_array(np.int32), small[idx])
    v0 = bound.views[0]
    idx0 = idx[idx < 40_000]
    di0 = ws.DevBuf.from_array(idx0)
    ws.check(gpu.ws_gather_enc(C.byref(v0), abi.P_I32, di0.p, len(idx0), out32.p))
    assert np.array_equal(out32.to_array(np.int32, len(idx0)), small[idx0])
    # dictionary code remap
    codes = rng.integers(-3, 120, n).astype(np.int32)
    codes[::11] = abi.NULL_I32
    cmap = rng.integers(0, 1000, 100).astype(np.int32)
    dc, dm, do = ws.DevBuf.from_array(codes), ws.DevBuf.from_array(cmap), ws.DevBuf(n * 4)
    ws.check(gpu.ws_remap_codes(dc.p, n, dm.p, len(cmap), do.p))
    exp = np.zeros(n, np.int32)
    H.oracle().or_remap_codes(ws.ptr(codes), n, ws.ptr(cmap), len(cmap), ws.ptr(exp))
    assert np.array_equal(do.to_array(np.int32), exp)
    # join-filter key bitmap
    keys = rng.integers(0, 5000, 3000).astype(np.int32)
    dk = ws.DevBuf.from_array(keys)
    nbits = 4000
    bits = ws.DevBuf(((nbits + 31) // 32) * 4)
    ws.check(gpu.ws_key_bitmap(colview(dk, abi.T_I32), len(keys), 100, nbits, bits.p))
    w = bits.to_array(np.uint32)
    got = np.nonzero(np.unpackbits(w.view(np.uint8), bitorder="little")[:nbits])[0]
    kk = keys.astype(np.int64) - 100
    assert np.array_equal(got, np.unique(kk[(kk >= 0) & (kk < nbits)]))
    # int64 bits -> double
    bv = rng.integers(-2 ** 50, 2 ** 50, 1000).astype(np.int64)
    db2, dd = ws.DevBuf.from_array(bv), ws.DevBuf(8000)
    ws.check(gpu.ws_bits_to_double(db2.p, 1000, dd.p))
    assert np.array_equal(dd.to_array(np.float64), bv.astype(np.float64))


def test_compact_column_major(gpu):
    """GpuAggregator::compact leaves the groups in VRAM, column-major."""
    H.say("case compact column-major")
    k, v, p = H.gen_c1(50_000)
    t = H.Table({"k": k, "v": v, "p": p}, 1 << 20)
    b = abi.EnvBuilder()
    sk, sv, sp = b.slot(abi.T_I32), b.slot(abi.T_I64), b.slot(abi.T_F64)
    plan = H.plan_c1(b, sk, sv, sp)
    slots = [("k", abi.T_I32, 0), ("v", abi.T_I64, 0), ("p", abi.T_F64, 0)]
    dev = H.Bound(t, slots, True)
    h, _ = H.gpu_aggregate(plan, b.env, dev, cap=1 << 12, fetch=False)
    try:
        kp, ap, g = C.c_void_p(), C.c_void_p(), C.c_int64()
        ws.check(gpu.ws_agg_compact(h, C.byref(kp), C.byref(ap), C.byref(g)))
        keys = ws.DevBuf(g.value * 8, tmp=True, adopt=kp.value).to_array(np.float64, g.value)
        accs = ws.DevBuf(g.value * 48, tmp=True, adopt=ap.value).to_array(np.float64, g.value * 6).reshape(6, g.value).T
    finally:
        gpu.ws_agg_destroy(h)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, H.Bound(t, slots, False))
    H.assert_groups_equal(plan, keys.reshape(-1, 1), accs.copy(), ok, oa)


@pytest.mark.parametrize("keys", ["unique", "duplicate"])
def test_hash_join_sized_equals_plain(gpu, keys):
    """The join with a predicted output size (buffers sized ahead, written without reading the
    count back) gives the plain join's pairs in the same probe order, whether the prediction is
    exact, too small (a second write pass) or too large; duplicate build keys take the two-pass
    form."""
    H.say(f"case sized join {keys}")
    rng = np.random.default_rng(15)
    if keys == "unique":
        bk = rng.permutation(40_000)[:30_000].astype(np.int64)
    else:
        bk = rng.integers(0, 9_000, 30_000).astype(np.int64)
    bk[::89] = abi.NULL_I64
    pk = rng.integers(-100, 42_000, 400_000).astype(np.int64)
    pk[::103] = abi.NULL_I64
    db, dp = ws.DevBuf.from_array(bk), ws.DevBuf.from_array(pk)
    bv, pv = colview(db, abi.T_I64N), colview(dp, abi.T_I64N)
    gb, gp = run_join(gpu, bv, len(bk), pv, len(pk))
    m0 = len(gb)
    eb, ep = oracle_join(bk, abi.T_I64N, pk, abi.T_I64N)
    assert sorted(zip(gp.tolist(), gb.tolist())) == sorted(zip(ep.tolist(), eb.tolist()))
    for expect in (m0, max(1, m0 // 3), 3 * m0 + 7):
        ob, op, m = C.c_void_p(), C.c_void_p(), C.c_int64()
        ws.check(gpu.ws_hash_join_sized(bv, len(bk), pv, len(pk), expect, C.byref(ob), C.byref(op), C.byref(m)))
        assert m.value == m0, (expect, m.value, m0)
        b = ws.DevBuf(m.value * 4, tmp=True, adopt=ob.value).to_array(np.int32, m.value)
        p = ws.DevBuf(m.value * 4, tmp=True, adopt=op.value).to_array(np.int32, m.value)
        # probe order is fixed; the build rows of one probe row follow the table's chain order,
        # which depends on the order the build inserted them (so compare those as sets)
        assert np.array_equal(p, gp), expect
        if keys == "unique":
            assert np.array_equal(b, gb), expect
        else:
            assert np.array_equal(b[np.lexsort((b, p))], gb[np.lexsort((gb, gp))]), expect
    ob, op, m = C.c_void_p(), C.c_void_p(), C.c_int64()
    ws.check(gpu.ws_hash_join_sized(bv, 0, pv, len(pk), 100, C.byref(ob), C.byref(op), C.byref(m)))
    assert m.value == 0 and not ob.value and not op.value


@pytest.mark.parametrize("width", ["i64n_f64", "encoded"])
def test_select_ballot_pass_wide_and_encoded_columns(gpu, width):
    """The ballot pass's four-rows-per-lane loads on tiles without deleted rows, for 8-byte columns
    (nullable BIGINT with NULLs, DOUBLE with NaN: 16-byte pair loads) and for frame-of-reference
    encoded columns (u8 / u16 codes, word loads); ragged batches put batch bases off 16-byte
    alignment so some batches take the row-by-row fallback. Against the oracle and the chunked
    single-pass kernel."""
    H.say(f"case ballot pass {width}")
    n = 2_500_000
    rng = np.random.default_rng(73)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    v[rng.random(n) < 0.05] = abi.NULL_I64
    x = rng.standard_normal(n)
    x[rng.random(n) < 0.05] = np.nan
    r = rng.integers(0, 200, n).astype(np.int32)  # u8 frame-of-reference codes
    u = rng.integers(0, 40_000, n).astype(np.int32)  # u16 codes
    sizes = [1_048_576, 3, 700_001, 751_420]  # 3 and 700_001 leave later batch bases unaligned
    t = H.Table({"v": v, "x": x, "r": r, "u": u}, sizes)
    b = abi.EnvBuilder()
    if width == "i64n_f64":
        slots = [("v", abi.T_I64N, 0), ("x", abi.T_F64, 0)]
        s0, s1 = b.slot(abi.T_I64N), b.slot(abi.T_F64)
        # WHERE v > 100 OR x < -0.5 (NULL and NaN rows take three-valued logic)
        where = b.prog([(abi.OP_COL, s0), b.ilit(100), (abi.OP_GT,), (abi.OP_COL, s1), b.dlit(-0.5), (abi.OP_LT,), (abi.OP_OR,)])
    else:
        e0 = t.encode("r")
        e1 = t.encode("u")
        assert (e0, e1) == (1, 2)
        slots = [("r", abi.T_I32, 0), ("u", abi.T_I32, 0)]
        s0, s1 = b.slot(abi.T_I32, enc=e0), b.slot(abi.T_I32, enc=e1)
        where = b.prog([(abi.OP_COL, s0), b.ilit(37), (abi.OP_LT,), (abi.OP_COL, s1), b.ilit(30_000), (abi.OP_GT,), (abi.OP_AND,)])
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    nb = len(t.sizes)
    bn = np.array(t.sizes, np.int64)
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    fp = abi.AggPlan()
    fp.has_where, fp.where = 1, where
    jit = C.c_void_p()
    ws.check(gpu.ws_jit_get(C.byref(b.env), 2, C.byref(fp), abi.JF_FLAGM, b"filter", C.byref(jit)))
    assert jit.value

    def run():
        out = C.c_void_p()
        per = np.zeros(nb, np.int64)
        cnt = C.c_int64()
        ws.check(gpu.ws_select_multi(C.byref(b.env), 1, where, nb, ws.ptr(bn), None, C.cast(dev.views, C.c_void_p), 2,
                                     C.byref(out), ws.ptr(per), jit, C.byref(cnt)))
        return ws.DevBuf(cnt.value * 4, tmp=True, adopt=out.value).to_array(np.int32, cnt.value), per

    v_old = C.c_int64()
    ws.check(gpu.ws_tune_get(b"select_mask_min_rows", C.byref(v_old)))
    try:
        ws.check(gpu.ws_tune_set(b"select_mask_min_rows", 1 << 40))
        plain, plain_per = run()
        ws.check(gpu.ws_tune_set(b"select_mask_min_rows", 0))
        got, per = run()
    finally:
        ws.check(gpu.ws_tune_set(b"select_mask_min_rows", v_old.value))
    exp = np.zeros(n, np.int32)
    eper = np.zeros(nb, np.int64)
    m = H.oracle().or_select(C.byref(b.env), 1, where, nb, ws.ptr(bn), None, C.cast(host.views, C.c_void_p), 2, ws.ptr(exp),
                             ws.ptr(eper))
    assert m > 10_000
    assert len(got) == m and np.array_equal(got, exp[:m]) and np.array_equal(per, eper)
    assert np.array_equal(plain, exp[:m]) and np.array_equal(plain_per, eper)


@pytest.mark.parametrize("deletes", [False, True])
def test_key_bitmap_where_matches_materialized_bitmap(gpu, deletes):
    """ws_key_bitmap_where (the filtered build side's key bitmap in one scan) against numpy: ragged
    batches, deleted rows, keys below and above the bitmap range (ignored), an expression key, and
    the count of passing rows."""
    H.say(f"case key bitmap where deletes={deletes}")
    rng = np.random.default_rng(17)
    n = 400_000
    key = rng.integers(-1000, 300_000, n).astype(np.int32)
    attr = rng.integers(0, 1000, n).astype(np.int32)
    deleted = rng.random(n) < 0.1 if deletes else None
    t = H.Table({"k": key, "a": attr}, [131_072, 7, 0, 200_000, 68_921], deleted)
    b = abi.EnvBuilder()
    sk, sa = b.slot(abi.T_I32), b.slot(abi.T_I32)
    where = b.prog([(abi.OP_COL, sa), b.ilit(70), (abi.OP_LT,)])
    kprog = b.prog([(abi.OP_COL, sk), b.ilit(5), (abi.OP_ADD,)])
    slots = [("k", abi.T_I32, 0), ("a", abi.T_I32, 0)]
    dev = H.Bound(t, slots, True)
    lo, nbits = 100, 250_000
    bits = ws.DevBuf(((nbits + 31) // 32) * 4)
    cnt = ws.DevBuf(8)
    nb = len(t.sizes)
    live = np.ones(n, bool) if deleted is None else ~deleted
    passing = live & (attr < 70)
    k = key[passing].astype(np.int64) + 5 - lo
    k = k[(k >= 0) & (k < nbits)]
    want = np.zeros((nbits + 31) // 32 * 32, bool)
    want[k] = True
    kp = abi.AggPlan()
    kp.has_where, kp.where, kp.nkeys, kp.keys[0], kp.key_int[0] = 1, where, 1, kprog, 1
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    jit = C.c_void_p()
    ws.check(gpu.ws_jit_get(C.byref(b.env), 2, C.byref(kp), abi.JF_FLAGM, b"key bitmap", C.byref(jit)))
    assert jit.value
    sizes = np.array(t.sizes, np.int64)
    for plan in (None, jit):  # the interpreter, then the compiled plan
        ws.check(gpu.ws_key_bitmap_where(C.byref(b.env), 1, where, kprog, nb, ws.ptr(sizes),
                                         C.cast(dev.dels, C.c_void_p), C.cast(dev.views, C.c_void_p), 2, lo, nbits, bits.p,
                                         cnt.p, plan))
        got = bits.to_array(np.uint32, (nbits + 31) // 32)
        assert np.array_equal(got, np.packbits(want, bitorder="little").view(np.uint32))
        assert int(cnt.to_array(np.uint64, 1)[0]) == int(passing.sum())
    # a floating-point key is refused
    sp = b.slot(abi.T_F64)
    fprog = b.prog([(abi.OP_COL, sp)])
    rc = gpu.ws_key_bitmap_where(C.byref(b.env), 1, where, fprog, nb, ws.ptr(sizes),
                                 C.cast(dev.dels, C.c_void_p), C.cast(dev.views, C.c_void_p), 2, lo, nbits, bits.p, None,
                                 None)
    assert rc == ws.WS_ERR_INTERNAL
