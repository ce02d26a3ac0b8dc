"""Pins the CPU oracle (test infrastructure) before it is trusted as the parity reference:
  - against golden vectors recorded from the reference's own expression semantics (its
    devtypes.h compiled on the host, tests/golden/make_golden.py), and against that harness
    directly when it is built here;
  - against brute force / closed forms (numpy group-bys, nested-loop joins, stable sorts,
    Hinnant's civil-date identities, 128-bit arithmetic);
  - against the known answers of the reference's smoke test (tests/smoke.py:20-35)."""
import ctypes as C
import json
import os
import random
import struct

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi
from golden.make_golden import Case, encoded_cases, gen_cases, NROWS

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "vm_golden.json")


def fbits(x: float) -> str:
    return struct.pack("<d", x).hex()


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_vm_matches_reference_golden(golden):
    O = H.oracle()
    bad = []
    for i, case in enumerate(golden["vm"]):
        code = [tuple(c) for c in case["code"]]
        c = Case(case["cols"], code, case["sets"], case["maps"])
        for r in range(NROWS):
            got = fbits(O.or_eval_prog(C.byref(c.env), None, c.prog, r))
            if got != case["expect"][r]:
                bad.append((i, r, got, case["expect"][r], code))
    assert not bad, f"{len(bad)} mismatches, first: {bad[:3]}"


def test_typed_evaluation_matches_golden_where_exact(golden):
    """The compiled-plan (typed) restatement equals the reference VM's golden value on every row
    where no int64 node changed the value; every row where one did has a cause the double VM
    cannot represent (operands or results beyond 2^53, wrapped overflow, a raw NULL sentinel read
    as an integer, YEAR of a fractional or out-of-range double through the datetime path). A
    mistyped or mis-dispatched operator shows up as an unexplained difference or as a mismatch."""
    O = H.oracle()
    O.or_eval_typed.restype = C.c_double
    O.or_eval_typed.argtypes = [C.POINTER(abi.ExprEnv), C.c_void_p, abi.Prog, C.c_int64, C.POINTER(C.c_int)]
    bad, explained, exact = [], 0, 0
    for i, case in enumerate(golden["vm"]):
        code = [tuple(c) for c in case["code"]]
        c = Case(case["cols"], code, case["sets"], case["maps"])
        for r in range(NROWS):
            fl = C.c_int()
            got = O.or_eval_typed(C.byref(c.env), None, c.prog, r, C.byref(fl))
            want = struct.unpack("<d", bytes.fromhex(case["expect"][r]))[0]
            if fl.value & 2:
                bad.append(("unexplained", i, r, fbits(got), case["expect"][r], code))
            elif fl.value & 1:
                explained += 1
            elif not ((np.isnan(got) and np.isnan(want)) or fbits(got) == case["expect"][r]):
                bad.append(("mismatch", i, r, fbits(got), case["expect"][r], code))
            else:
                exact += 1
    assert not bad, f"{len(bad)} bad rows, first: {bad[:3]}"
    assert exact > 0.9 * len(golden["vm"]) * NROWS, (exact, explained)


def test_typed_evaluation_closed_forms():
    """The typed rules on their own: int64 + - * wrap (2^62 * 4 -> 0), NEG of INT64_MIN wraps,
    comparisons beyond 2^53 are exact, a raw DECIMAL sentinel read as an integer is the sentinel,
    YEAR of an integer day is the civil year."""
    O = H.oracle()
    O.or_eval_typed.restype = C.c_double
    O.or_eval_typed.argtypes = [C.POINTER(abi.ExprEnv), C.c_void_p, abi.Prog, C.c_int64, C.POINTER(C.c_int)]
    big = np.array([2 ** 62, -2 ** 63, 2 ** 53 + 1, 19723], np.int64)
    dec = np.array([abi.NULL_I64, 5, 7, 9], np.int64)
    b = abi.EnvBuilder()
    sb, sd = b.slot(abi.T_I64), b.slot(abi.T_DECIMAL, 2)
    b.env.cols[sb].p, b.env.cols[sd].p = big.ctypes.data, dec.ctypes.data

    def ev(prog, row):
        fl = C.c_int()
        return O.or_eval_typed(C.byref(b.env), None, prog, row, C.byref(fl)), fl.value

    assert ev(b.prog([(abi.OP_COL, sb), b.ilit(4), (abi.OP_MUL,)]), 0) == (0.0, 1)  # wraps; the VM gives 2^64
    assert ev(b.prog([(abi.OP_COL, sb), (abi.OP_NEG,)]), 1) == (float(-2 ** 63), 1)
    assert ev(b.prog([(abi.OP_COL, sb), b.ilit(2 ** 53), (abi.OP_GT,)]), 2) == (1.0, 1)  # doubles say equal
    assert ev(b.prog([(abi.OP_COLRAW, sd, 1)]), 0) == (float(abi.NULL_I64), 1)
    assert ev(b.prog([(abi.OP_COLRAW, sd, 0)]), 0)[0] != ev(b.prog([(abi.OP_COLRAW, sd, 0)]), 0)[0]  # NULL
    assert ev(b.prog([(abi.OP_COL, sb), (abi.OP_YEAR,)]), 3) == (2024.0, 0)
    assert ev(b.prog([(abi.OP_COL, sb), b.ilit(3), (abi.OP_ADD,)]), 3) == (19726.0, 0)


def test_vm_matches_reference_harness_live():
    R = H.ref_lib()
    if R is None:
        pytest.skip("reference harness not built here (no reference tree)")
    O = H.oracle()
    for case in gen_cases(seed=99, n=120):
        c = Case(case["cols"], case["code"], case["sets"], case["maps"])
        for r in range(NROWS):
            a = fbits(O.or_eval_prog(C.byref(c.env), None, c.prog, r))
            b = fbits(R.ref_eval_prog(C.addressof(c.env), None, 0, c.prog.len, r))
            assert a == b, (case["code"], r, a, b)


def test_decoders_match_reference(golden):
    O = H.oracle()
    for d in golden["decode"]:
        blob = np.frombuffer(bytes.fromhex(d["blob"]), np.uint8).copy()
        v = abi.ColView()
        v.p, v.t, v.scale, v.enc, v.off = blob.ctypes.data, d["t"], d["scale"], d["enc"], d["off"]
        for r in range(d["rows"]):
            assert fbits(O.or_value(C.byref(v), r)) == d["load"][r], (d["enc"], d["t"], r)
            if d["enc"] < 4:
                assert O.or_stored_int(C.byref(v), r) == d["colint"][r]


def test_dates_checked_and_functions(golden):
    O = H.oracle()
    for z, y, m, d, back in golden["dates"]:
        yy, mm, dd = C.c_int64(), C.c_uint(), C.c_uint()
        O.or_civil_from_days(z, C.byref(yy), C.byref(mm), C.byref(dd))
        assert (yy.value, mm.value, dd.value) == (y, m, d)
        assert O.or_days_from_civil(y, m, d) == back == z
    for op, a, b, r, f in golden["checked"]:
        flag = C.c_int(0)
        assert O.or_checked(op, a, b, C.byref(flag)) == r and flag.value == f
    for fn, x, want in golden["call1"]:
        xv = struct.unpack("<d", bytes.fromhex(x))[0]
        assert fbits(O.or_call1(fn, xv)) == want, (fn, xv)
    for fn, a, b, want in golden["call2"]:
        av, bv = (struct.unpack("<d", bytes.fromhex(h))[0] for h in (a, b))
        assert fbits(O.or_call2(fn, av, bv)) == want, (fn, av, bv)


def test_civil_dates_closed_forms():
    O = H.oracle()
    assert O.or_days_from_civil(1970, 1, 1) == 0
    assert O.or_days_from_civil(2000, 3, 1) == 11017
    assert O.or_days_from_civil(2024, 1, 1) == 19723
    # consecutive days: +1 each, month lengths of a leap year
    d0 = O.or_days_from_civil(2024, 1, 1)
    lengths = [31, 29, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    acc = d0
    for m, ln in enumerate(lengths, start=1):
        assert O.or_days_from_civil(2024, m, 1) == acc
        acc += ln
    assert O.or_days_from_civil(2025, 1, 1) == acc


def test_checked_arithmetic_closed_form():
    O = H.oracle()
    rng = random.Random(3)
    for _ in range(2000):
        a = rng.randint(-2 ** 63, 2 ** 63 - 1)
        b = rng.choice([rng.randint(-2 ** 63, 2 ** 63 - 1), rng.randint(-1000, 1000)])
        for op, exact in ((0, a + b), (1, a - b), (2, a * b)):
            flag = C.c_int(0)
            r = O.or_checked(op, a, b, C.byref(flag))
            wrapped = (exact + 2 ** 63) % 2 ** 64 - 2 ** 63
            assert r == wrapped and flag.value == (exact != wrapped)


def _agg_case(rng, n, nulls=True):
    key = rng.integers(0, 37, n).astype(np.int32)
    v = rng.integers(-2 ** 40, 2 ** 40, n).astype(np.int64)
    if nulls:
        v[rng.random(n) < 0.1] = abi.NULL_I64
    x = rng.standard_normal(n)
    x[rng.random(n) < 0.05] = np.nan
    deleted = rng.random(n) < 0.15
    return key, v, x, deleted


def test_aggregate_matches_numpy_bruteforce():
    rng = np.random.default_rng(21)
    n = 20_000
    key, v, x, deleted = _agg_case(rng, n)
    t = H.Table({"k": key, "v": v, "x": x}, [7000, 0, 13000], deleted)
    b = abi.EnvBuilder()
    sk, sv, sx = b.slot(abi.T_I32), b.slot(abi.T_I64N), b.slot(abi.T_F64)
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sx), b.dlit(1.5), (abi.OP_LT,)])  # NULL x fails the filter
    plan.nkeys, plan.keys[0], plan.key_int[0] = 1, b.col(sk), 1
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX, abi.AGG_SUM, abi.AGG_MAX]
    plan.naggs = len(ops)
    for a, o in enumerate(ops):
        plan.ops[a] = o
    plan.agg_int[0] = 1
    for a in (1, 2, 3, 4):
        plan.aggs[a], plan.agg_int[a], plan.agg_bits[a], plan.agg_null[a] = b.col(sv), 1, 1, abi.T_I64N
    plan.agg_part[1], plan.agg_part[2] = 1, 2
    plan.aggs[5] = plan.aggs[6] = b.col(sx)
    host = H.Bound(t, [("k", abi.T_I32, 0), ("v", abi.T_I64N, 0), ("x", abi.T_F64, 0)], False)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    live = ~deleted & (x < 1.5)
    keys = np.unique(key[live])
    assert np.array_equal(ok[:, 0].view(np.int64), keys)
    for i, kk in enumerate(keys):
        m = live & (key == kk)
        vv = v[m][v[m] != abi.NULL_I64]
        assert oa[i, 0] == m.sum()
        lo = sum(int(q) & 0xffffffff for q in vv)
        hi = sum(int(q) >> 32 for q in vv)
        assert oa[i, 1:2].view(np.int64)[0] == lo and oa[i, 2:3].view(np.int64)[0] == hi
        assert hi * 2 ** 32 + lo == sum(int(q) for q in vv)  # the split recombines exactly
        if len(vv):
            assert oa[i, 3:4].view(np.int64)[0] == vv.min() and oa[i, 4:5].view(np.int64)[0] == vv.max()
        np.testing.assert_allclose(oa[i, 5], x[m].sum(), rtol=1e-12)
        assert oa[i, 6] == x[m].max()


def test_nokeys_128bit_sum_closed_form():
    n = 5000
    v = np.full(n, 2 ** 62 + 12345, np.int64)
    t = H.Table({"v": v}, n)
    b = abi.EnvBuilder()
    s = b.slot(abi.T_I64)
    plan = abi.AggPlan()
    plan.naggs = 2
    plan.ops[0], plan.ops[1] = abi.AGG_COUNT, abi.AGG_SUM
    plan.agg_int[0] = plan.agg_int[1] = 1
    plan.aggs[1] = b.col(s)
    _, oa, hi = H.oracle_aggregate(plan, b.env, H.Bound(t, [("v", abi.T_I64, 0)], False))
    total = (int(hi[1]) << 64) + int(oa[0, 1:2].view(np.uint64)[0])
    assert total == n * (2 ** 62 + 12345)


def test_smoke_known_answers():
    """The reference's smoke test table (tests/smoke.py:16-35), names dictionary-encoded."""
    names = ["apple", "banana", "cherry"]
    name = np.array([0, 1, 2, 0], np.int32)
    price = np.array([1.5, 0.25, 3.0, 2.5])
    O = H.oracle()
    day = lambda y, m, d: O.or_days_from_civil(y, m, d)  # noqa: E731
    d = np.array([day(2024, 1, 1), day(2024, 2, 1), day(2024, 3, 1), day(2024, 4, 1)], np.int32)
    t = H.Table({"name": name, "price": price, "d": d}, 4)
    host = H.Bound(t, [("name", abi.T_I32, 0), ("price", abi.T_F64, 0), ("d", abi.T_I32, 0)], False)
    # GROUP BY name -> COUNT(*), SUM(price), MIN(d)
    b = abi.EnvBuilder()
    sn, sp, sd = b.slot(abi.T_I32), b.slot(abi.T_F64), b.slot(abi.T_I32)
    plan = abi.AggPlan()
    plan.nkeys, plan.keys[0], plan.key_int[0] = 1, b.col(sn), 1
    plan.naggs = 3
    plan.ops[0], plan.ops[1], plan.ops[2] = abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_MIN
    plan.agg_int[0] = 1
    plan.aggs[1] = b.col(sp)
    plan.aggs[2], plan.agg_int[2] = b.col(sd), 1
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host)
    got = [(names[int(k)], int(a[0]), a[1], int(a[2])) for k, a in zip(ok[:, 0].view(np.int64), oa)]
    assert got == [("apple", 2, 4.0, day(2024, 1, 1)), ("banana", 1, 0.25, day(2024, 2, 1)),
                   ("cherry", 1, 3.0, day(2024, 3, 1))]
    # WHERE name LIKE 'b%' OR price > 2 -> COUNT(*), SUM(price) = (3, 5.75); LIKE is a dictionary bitmap
    like = np.array([0b010], np.uint32)
    b = abi.EnvBuilder()
    sn, sp = b.slot(abi.T_I32), b.slot(abi.T_F64)
    si = b.add_set(like.ctypes.data, 3)
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sn), (abi.OP_INSET, si), (abi.OP_COL, sp), b.dlit(2.0), (abi.OP_GT,), (abi.OP_OR,)])
    plan.naggs = 2
    plan.ops[0], plan.ops[1] = abi.AGG_COUNT, abi.AGG_SUM
    plan.agg_int[0] = 1
    plan.aggs[1] = b.col(sp)
    _, oa, _ = H.oracle_aggregate(plan, b.env, H.Bound(t, [("name", abi.T_I32, 0), ("price", abi.T_F64, 0)], False))
    assert (int(oa[0, 0]), oa[0, 1]) == (3, 5.75)
    # GROUP BY CASE WHEN price > 10 THEN 1 ELSE 0 END after the +10 update of apples -> (0, 2), (1, 2)
    price2 = np.array([11.5, 0.25, 3.0, 12.5])
    t2 = H.Table({"price": price2}, 4)
    b = abi.EnvBuilder()
    sp = b.slot(abi.T_F64)
    plan = abi.AggPlan()
    plan.nkeys = 1
    plan.keys[0] = b.prog([(abi.OP_COL, sp), b.dlit(10.0), (abi.OP_GT,), b.ilit(1), b.ilit(0), (abi.OP_SELECT,)])
    plan.naggs, plan.ops[0], plan.agg_int[0] = 1, abi.AGG_COUNT, 1
    ok, oa, _ = H.oracle_aggregate(plan, b.env, H.Bound(t2, [("price", abi.T_F64, 0)], False))
    assert [(k, int(a[0])) for k, a in zip(ok[:, 0], oa)] == [(0.0, 2), (1.0, 2)]
    # YEAR(d) GROUP BY -> (2024, 4)
    b = abi.EnvBuilder()
    sd = b.slot(abi.T_I32)
    plan = abi.AggPlan()
    plan.nkeys, plan.keys[0] = 1, b.prog([(abi.OP_COL, sd), (abi.OP_YEAR,)])
    plan.naggs, plan.ops[0], plan.agg_int[0] = 1, abi.AGG_COUNT, 1
    ok, oa, _ = H.oracle_aggregate(plan, b.env, H.Bound(t, [("d", abi.T_I32, 0)], False))
    assert [(k, int(a[0])) for k, a in zip(ok[:, 0], oa)] == [(2024.0, 4)]


def test_select_join_sort_bruteforce():
    O = H.oracle()
    rng = np.random.default_rng(5)
    # joins: nested loops on small inputs
    bk = rng.integers(0, 30, 60).astype(np.int32)
    bk[::13] = abi.NULL_I32
    pk = rng.integers(-5, 40, 200).astype(np.int32)
    pk[::17] = abi.NULL_I32
    for outer in (0, 1, 2):
        cap = 100000
        ob, op = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        bv, pv = abi.ColView(), abi.ColView()
        bv.p, bv.t, pv.p, pv.t = bk.ctypes.data, abi.T_I32N, pk.ctypes.data, abi.T_I32N
        m = O.or_join(C.byref(bv), len(bk), C.byref(pv), len(pk), outer, ob.ctypes.data, op.ctypes.data, cap)
        exp = []
        matched = set()
        for j, k in enumerate(pk):
            hits = [i for i, kb in enumerate(bk) if k != abi.NULL_I32 and kb == k]
            matched.update(hits)
            exp += [(j, i) for i in hits] or ([(j, -1)] if outer else [])
        if outer == 2:
            exp += [(-1, i) for i in range(len(bk)) if i not in matched]
        assert sorted(zip(op[:m], ob[:m])) == sorted(exp)
    # stable sort: numpy lexsort with the NULL / -0 / descending conventions
    m = 5000
    k0 = rng.integers(0, 9, m).astype(np.float64)
    k0[rng.random(m) < 0.1] = np.nan
    k1 = np.round(rng.standard_normal(m), 1)
    k1[k1 == 0] = -0.0
    desc = np.array([1, 0], np.uint8)
    perm = np.zeros(m, np.int32)
    O.or_sort_perm((C.c_void_p * 2)(k0.ctypes.data, k1.ctypes.data), desc.ctypes.data, 2, m, perm.ctypes.data)
    n0 = np.where(np.isnan(k0), -np.inf, k0)
    n1 = np.where(k1 == 0, 0.0, k1)
    exp = np.lexsort((np.arange(m), n1, -n0))
    assert np.array_equal(perm, exp)


def _and3(a, b):
    """SQL AND over {1, 0, NaN = NULL}: false wins, then NULL."""
    return np.where((a == 0) | (b == 0), 0.0, np.where(np.isnan(a) | np.isnan(b), np.nan, 1.0))


def _or3(a, b):
    return np.where((a == 1) | (b == 1), 1.0, np.where(np.isnan(a) | np.isnan(b), np.nan, 0.0))


def _cmp3(mask, *nulls):
    out = mask.astype(np.float64)
    for n in nulls:
        out[n] = np.nan
    return out


def test_select_matches_numpy_mask():
    """or_select against a numpy three-valued mask and np.nonzero: ragged batches with an empty
    one, deleted rows, NULL sentinels (I64N, I32N) and NaN doubles, AND / OR / NOT over NULLs."""
    O = H.oracle()
    rng = np.random.default_rng(31)
    n = 30_000
    k = rng.integers(0, 10, n).astype(np.int32)
    k[rng.random(n) < 0.07] = abi.NULL_I32
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    v[rng.random(n) < 0.1] = abi.NULL_I64
    x = rng.standard_normal(n)
    x[rng.random(n) < 0.1] = np.nan
    deleted = rng.random(n) < 0.2
    sizes = [9000, 0, 12345, 8655]
    t = H.Table({"k": k, "v": v, "x": x}, sizes, deleted)
    host = H.Bound(t, [("k", abi.T_I32N, 0), ("v", abi.T_I64N, 0), ("x", abi.T_F64, 0)], False)
    b = abi.EnvBuilder()
    sk, sv, sx = b.slot(abi.T_I32N), b.slot(abi.T_I64N), b.slot(abi.T_F64)
    # WHERE (x < 0.5 AND v > -200) OR NOT (k <> 3)
    where = b.prog([(abi.OP_COL, sx), b.dlit(0.5), (abi.OP_LT,), (abi.OP_COL, sv), b.ilit(-200), (abi.OP_GT,), (abi.OP_AND,),
                    (abi.OP_COL, sk), b.ilit(3), (abi.OP_NE,), (abi.OP_NOT,), (abi.OP_OR,)])
    xl = _cmp3(x < 0.5, np.isnan(x))
    vg = _cmp3(v > -200, v == abi.NULL_I64)
    kne = _cmp3(k != 3, k == abi.NULL_I32)
    knot = np.where(np.isnan(kne), np.nan, 1.0 - kne)
    res = _or3(_and3(xl, vg), knot)
    for has_where, sel in ((1, (res == 1) & ~deleted), (0, ~deleted)):
        out = np.zeros(n, np.int32)
        per = np.zeros(len(sizes), np.int64)
        sz = np.array(sizes, np.int64)
        m = O.or_select(C.byref(b.env), has_where, where, len(sizes), ws.ptr(sz),
                        C.cast(host.dels, C.c_void_p), C.cast(host.views, C.c_void_p), host.nslots, ws.ptr(out), ws.ptr(per))
        exp = np.nonzero(sel)[0]
        assert np.array_equal(out[:m], exp)
        offs = np.concatenate([[0], np.cumsum(sizes)])
        assert per.tolist() == [int(sel[offs[i]:offs[i + 1]].sum()) for i in range(len(sizes))]
    # a NULL WHERE (NaN) never selects; a non-zero non-NaN value always does
    b = abi.EnvBuilder()
    sx = b.slot(abi.T_F64)
    t2 = H.Table({"x": np.array([np.nan, 0.0, -0.0, 2.5, -1e-300])}, 5)
    h2 = H.Bound(t2, [("x", abi.T_F64, 0)], False)
    out = np.zeros(5, np.int32)
    five = np.array([5], np.int64)
    m = O.or_select(C.byref(b.env), 1, b.col(sx), 1, ws.ptr(five), C.cast(h2.dels, C.c_void_p),
                    C.cast(h2.views, C.c_void_p), 1, ws.ptr(out), None)
    assert out[:m].tolist() == [3, 4]


def test_remap_codes_matches_numpy():
    """or_remap_codes against np.where: NULL stays NULL, codes outside [0, len(map)) -> -2."""
    O = H.oracle()
    rng = np.random.default_rng(32)
    cmap = rng.integers(-2, 500, 300).astype(np.int32)
    codes = rng.integers(-5, 320, 10_000).astype(np.int32)
    codes[::11] = abi.NULL_I32
    out = np.zeros_like(codes)
    O.or_remap_codes(ws.ptr(codes), len(codes), ws.ptr(cmap), len(cmap), ws.ptr(out))
    inside = (codes >= 0) & (codes < len(cmap))
    exp = np.where(codes == abi.NULL_I32, abi.NULL_I32, np.where(inside, cmap[np.clip(codes, 0, len(cmap) - 1)], -2))
    assert np.array_equal(out, exp)
    O.or_remap_codes(ws.ptr(codes), len(codes), ws.ptr(cmap), 0, ws.ptr(out))  # empty map: everything absent
    assert np.array_equal(out, np.where(codes == abi.NULL_I32, abi.NULL_I32, -2))


def test_aggregate_dense_matches_bruteforce():
    """or_aggregate(dense=1) against Python-integer brute force: COUNT; integer SUM without
    agg_bits as the double of the exact int64 sum (values near 2^56, so sums pass 2^53 and the
    rounding matters); split BIGINT SUM halves; integer MIN / MAX as doubles (no agg_bits) and as
    exact bits (agg_bits); double SUM / MIN / MAX skipping NaN; deleted rows; two keys."""
    rng = np.random.default_rng(33)
    n = 12_000
    k0 = rng.integers(3, 9, n).astype(np.int32)
    k1 = rng.integers(-2, 2, n).astype(np.int64)
    v = rng.integers(2 ** 55, 2 ** 56, n).astype(np.int64)
    v[rng.random(n) < 0.5] *= -1
    w = rng.integers(-2 ** 40, 2 ** 40, n).astype(np.int64)
    w[rng.random(n) < 0.1] = abi.NULL_I64
    x = rng.standard_normal(n)
    x[rng.random(n) < 0.1] = np.nan
    deleted = rng.random(n) < 0.1
    t = H.Table({"k0": k0, "k1": k1, "v": v, "w": w, "x": x}, [5000, 7000], deleted)
    host = H.Bound(t, [("k0", abi.T_I32, 0), ("k1", abi.T_I64, 0), ("v", abi.T_I64, 0), ("w", abi.T_I64N, 0),
                       ("x", abi.T_F64, 0)], False)
    b = abi.EnvBuilder()
    s0, s1, sv, sw, sx = (b.slot(tc) for tc in (abi.T_I32, abi.T_I64, abi.T_I64, abi.T_I64N, abi.T_F64))
    plan = abi.AggPlan()
    plan.nkeys = 2
    plan.keys[0], plan.keys[1] = b.col(s0), b.col(s1)
    plan.key_int[0] = plan.key_int[1] = 1
    # 0 COUNT(*), 1 SUM(v) no bits, 2 MIN(v) no bits, 3 MAX(v) no bits, 4/5 SUM(w) lo/hi, 6 MIN(w) bits,
    # 7 MAX(w) bits, 8 SUM(x), 9 MIN(x), 10 MAX(x)
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX,
           abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX]
    plan.naggs = len(ops)
    for a, o in enumerate(ops):
        plan.ops[a] = o
    plan.agg_int[0] = 1
    for a in (1, 2, 3):
        plan.aggs[a], plan.agg_int[a] = b.col(sv), 1
    for a in (4, 5, 6, 7):
        plan.aggs[a], plan.agg_int[a], plan.agg_bits[a], plan.agg_null[a] = b.col(sw), 1, 1, abi.T_I64N
    plan.agg_part[4], plan.agg_part[5] = 1, 2
    for a in (8, 9, 10):
        plan.aggs[a] = b.col(sx)
    ok, oa, _ = H.oracle_aggregate(plan, b.env, host, dense=True)
    live = ~deleted
    keys = sorted({(int(a), int(c)) for a, c in zip(k0[live], k1[live])}, key=lambda kk: (kk[0], kk[1]))
    assert [(int(a), int(c)) for a, c in ok.view(np.int64)] == keys
    beyond = 0
    for i, (a0, a1) in enumerate(keys):
        m = live & (k0 == a0) & (k1 == a1)
        vv = [int(q) for q in v[m]]
        ww = [int(q) for q in w[m] if q != abi.NULL_I64]
        xx = x[m][~np.isnan(x[m])]
        row = oa[i]
        assert row[0] == m.sum()
        exact = sum(vv)
        beyond += abs(exact) > 2 ** 53
        assert row[1] == float(exact)  # the double nearest the exact sum (correct rounding)
        assert row[2] == float(min(vv)) and row[3] == float(max(vv))
        bits = row.view(np.int64)
        assert bits[4] == sum(q & 0xffffffff for q in ww) and bits[5] == sum(q >> 32 for q in ww)
        assert bits[6] == (min(ww) if ww else 2 ** 63 - 1) and bits[7] == (max(ww) if ww else -2 ** 63)
        np.testing.assert_allclose(row[8], xx.sum(), rtol=1e-12)
        assert row[9] == (xx.min() if len(xx) else np.inf) and row[10] == (xx.max() if len(xx) else -np.inf)
    assert beyond > 0  # the no-bits sums really pass 2^53


def test_c2_query_oracle_vs_numpy():
    user, region, date, url, dur = H.gen_c2(200_000, n_users=5000, n_urls=3000)
    full = H.oracle_c2(user, region, date, url, dur)
    m = (date < 19723 + 300) & (region < 224)
    u, inv = np.unique(user[m], return_inverse=True)
    cnt = np.bincount(inv)
    s = np.bincount(inv, weights=dur[m].astype(np.float64)).astype(np.int64)
    pairs = np.unique(np.stack([user[m], url[m]], 1), axis=0)
    dcount = dict(zip(*np.unique(pairs[:, 0], return_counts=True)))
    exp = sorted(((int(a), int(b), int(c), int(dcount[a])) for a, b, c in zip(u, cnt, s)), key=lambda r: (-r[1], r[0]))
    assert [tuple(r) for r in full] == exp


def test_c3_query_oracle_vs_numpy():
    fk, qty, price, flag, dkey, dattr = H.gen_c3(300_000, 20_000)
    got = H.oracle_c3(fk, qty, price, flag, dkey, dattr)
    sel = set(dkey[dattr < 50].tolist())
    m = np.isin(fk, list(sel))
    exp = []
    for f in np.unique(flag[m]):
        mm = m & (flag == f)
        exp.append([f, mm.sum(), int(price[mm].sum()), int(qty[mm].sum())])
    assert np.array_equal(got, np.array(exp, np.int64))


def test_encoded_decode_roundtrip_via_generator():
    """Frame-of-reference encoding decodes back to the original values (NULL sentinel kept)."""
    rng = np.random.default_rng(9)
    O = H.oracle()
    for dtype, lo, hi, tcode in ((np.int32, 0, 250, abi.T_I32N), (np.int32, -1000, 60000, abi.T_I32N),
                                 (np.int64, 10 ** 12, 10 ** 12 + 3 * 10 ** 9, abi.T_I64N)):
        x = rng.integers(lo, hi, 3000).astype(dtype)
        x[::29] = abi.NULL_I32 if dtype == np.int32 else abi.NULL_I64
        out, e = H.encode_column(x)
        assert e in (1, 2, 3)
        v = abi.ColView()
        v.p, v.t, v.enc = out.ctypes.data, tcode, e
        back = np.array([O.or_stored_int(C.byref(v), r) for r in range(len(x))])
        assert np.array_equal(back, x.astype(np.int64))


# ------------------------------------------------------------------ column statistics (sizing)

def _stats(arr, tcode, deleted=None):
    words = None
    if deleted is not None:
        words = np.zeros((len(arr) + 31) // 32, np.uint32)
        idx = np.nonzero(deleted)[0]
        np.bitwise_or.at(words, idx >> 5, (np.uint32(1) << (idx & 31).astype(np.uint32)))
    return H.oracle_column_stats(H.host_view(arr, tcode), len(arr), words)


def test_column_stats_hll_error_bound():
    """HyperLogLog with 2048 registers: relative standard error 1.04/sqrt(2048) = 2.3 %; every
    estimate here must be within 4 sigma of the true distinct count (the estimator's published
    bound), from 1 to 10^6 distinct values, with duplicates not changing the estimate."""
    L = ws.lib()
    rng = np.random.default_rng(12)
    for d in (1, 7, 100, 3000, 50_000, 1_000_000):
        vals = rng.choice(np.arange(d * 7, dtype=np.int64) * 1_000_003, d, replace=False)
        st = _stats(vals, abi.T_I64)
        est = L.ws_hll_estimate(st.hll)
        assert abs(est - d) <= max(1.0, 4 * 0.023 * d), (d, est)
        dup = np.concatenate([vals, vals[::-1], vals[: d // 2]])
        st2 = _stats(dup, abi.T_I64)
        assert bytes(st2.hll) == bytes(st.hll)  # idempotent under duplicates
        assert st2.rows == len(dup) and st.imin == vals.min() and st.imax == vals.max()


def test_column_stats_merge_nulls_signed_zero():
    """Stats of two halves merged == stats of the whole; NULLs counted apart; -0.0 == 0.0; NaN
    and INT64_MIN sentinels are NULL; deleted rows are not seen."""
    L = ws.lib()
    rng = np.random.default_rng(5)
    x = rng.integers(-1000, 1000, 20_000).astype(np.int64)
    x[rng.random(len(x)) < 0.1] = abi.NULL_I64
    deleted = rng.random(len(x)) < 0.2
    whole = _stats(x, abi.T_I64N, deleted)
    a = _stats(x[:7000], abi.T_I64N, deleted[:7000])
    b = _stats(x[7000:], abi.T_I64N, deleted[7000:])
    L.ws_col_stats_merge(C.byref(a), C.byref(b))
    assert bytes(a.hll) == bytes(whole.hll)
    assert (a.rows, a.nulls, a.imin, a.imax) == (whole.rows, whole.nulls, whole.imin, whole.imax)
    live = x[~deleted]
    assert whole.rows == len(live) and whole.nulls == int((live == abi.NULL_I64).sum())
    nn = live[live != abi.NULL_I64]
    assert whole.imin == nn.min() and whole.imax == nn.max() and whole.dmin == float(nn.min())
    d = np.array([0.0, -0.0, np.nan, 1.5, -2.0, 1.5], np.float64)
    sd = _stats(d, abi.T_F64)
    assert sd.nulls == 1 and sd.rows == 6 and sd.dmin == -2.0 and sd.dmax == 1.5
    sz = _stats(np.array([0.0]), abi.T_F64)
    szn = _stats(np.array([-0.0]), abi.T_F64)
    assert bytes(sz.hll) == bytes(szn.hll)
    assert round(L.ws_hll_estimate(sd.hll)) == 3  # {0, 1.5, -2}


# ------------------------------------------------------------------ sizing model (MLP) oracle

def _mlp():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mlp_oracle", os.path.join(H.ROOT, "oracle", "mlp_oracle.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_mlp_oracle_gradient_matches_finite_differences():
    """The oracle's analytic gradient equals central finite differences of its loss (float64),
    for every one of the 1633 parameters; a dropped ReLU mask, a transposed weight or a missing
    1/B would fail it."""
    M = _mlp()
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, M.N_PARAMS)
    x = rng.random((24, M.N_IN))
    t = rng.random(24)
    g = M.grad(p, x, t)
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(M.N_PARAMS):
        q = p.copy()
        q[i] += eps
        lp = M.loss(q, x, t)
        q[i] -= 2 * eps
        lm = M.loss(q, x, t)
        fd[i] = (lp - lm) / (2 * eps)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-8), np.abs(g - fd).max()
    assert M.N_PARAMS == 1633


def test_mlp_oracle_adam_closed_form():
    """First Adam step with bias correction moves each parameter by lr * sign(g) (|g| >> eps)."""
    M = _mlp()
    rng = np.random.default_rng(4)
    p = rng.normal(0, 1, 50)
    g = rng.normal(0, 1, 50)
    q, m, v = M.adam(p, g, np.zeros(50), np.zeros(50), 1, 0.01)
    assert np.allclose(p - q, 0.01 * np.sign(g), atol=1e-6)


def test_sizing_features_basic():
    """Feature builder: selectivity, key count and distinct-count features from known stats."""
    L = ws.lib()
    vals = np.arange(5000, dtype=np.int64) * 3 + 100
    st = H.oracle_column_stats(H.host_view(vals, abi.T_I64), len(vals))
    f = np.zeros(16, np.float32)
    ws.check(L.ws_sizing_features(C.byref(st), 1, None, 1 << 20, 0.25, ws.ptr(f)))
    assert abs(f[0] - np.log2(2 ** 20 + 1) / 32) < 1e-6 and abs(f[1] - 0.25) < 1e-7
    assert abs(f[3] * 32 - np.log2(5001)) < 0.1  # HyperLogLog within a few %
    assert abs(f[4] * 32 - np.log2(3 * 4999 + 2)) < 1e-4  # integer span
    assert f[11] == 0.25 and f[5] == 0 and f[6] == 0
    assert ws.lib().ws_sizing_features(C.byref(st), 5, None, 10, 1.0, ws.ptr(f)) != 0
