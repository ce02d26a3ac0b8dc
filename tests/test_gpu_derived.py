"""GPU parity of the operators around the scan that the engine calls for derived tables, IN /
EXISTS subqueries, ORDER BY on strings and materialisation (reference src/gpu.h:142-201):
integer evaluation, distinct values, temp-table columns, value bitmaps, wide-row gathers, cross
joins, host-column sort permutations, rank maps, deleted-row bits, and gathers over encoded batches
of every segment encoding. Expected values come from numpy restatements of each entry point's
contract (gpu.h) and from the CPU oracle; all comparisons are bit-exact."""
import ctypes as C
import math

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu

NULL32, NULL64 = abi.NULL_I32, abi.NULL_I64


def llround(x: float) -> int:
    """C's llround: nearest integer, halves away from zero (exact for every double)."""
    f = math.floor(abs(x))
    return int(math.copysign(f + (1 if abs(x) - f >= 0.5 else 0), x))


def llround_vec(y: np.ndarray) -> np.ndarray:
    a = np.abs(y)
    f = np.floor(a)
    return (np.sign(y) * (f + (a - f >= 0.5))).astype(np.int64)


def i64_of(d: float) -> int:
    return int(np.array([d]).view(np.int64)[0])


def eval_env(n, rng):
    """Columns: a I32 (with INT32_MIN), b I64N (NULLs), c I32N (NULLs), d DECIMAL(2), x F64 (NaN)."""
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    a[::97] = NULL32
    bcol = rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64)
    bcol[::13] = NULL64
    c = rng.integers(-50, 50, n).astype(np.int32)
    c[::7] = NULL32
    d = rng.integers(-10 ** 6, 10 ** 6, n).astype(np.int64)
    x = rng.normal(0, 1e6, n)
    x[::11] = np.nan
    b = abi.EnvBuilder()
    keep = []
    for arr, t, sc in ((a, abi.T_I32, 0), (bcol, abi.T_I64N, 0), (c, abi.T_I32N, 0), (d, abi.T_DECIMAL, 2), (x, abi.T_F64, 0)):
        s = b.slot(t, scale=sc)
        dv = ws.DevBuf.from_array(arr)
        keep.append(dv)
        b.env.cols[s].p = dv.p
    return b, keep, (a, bcol, c, d, x)


def test_eval_at_int_forms(gpu):
    """One column (a non-nullable INT's INT32_MIN reads as NULL), exact int64 programs (NULLs
    propagate, INT32_MIN of a non-nullable INT is a value), and the rounded double form
    (reference gpu.cu:826-883)."""
    H.say("case eval_at_int")
    n = 20_000
    rng = np.random.default_rng(31)
    b, keep, (a, bb, c, d, x) = eval_env(n, rng)
    idx = np.sort(rng.choice(n, 7_001, replace=False)).astype(np.int32)
    di = ws.DevBuf.from_array(idx)
    out = ws.DevBuf(n * 8)
    A = a.astype(object)
    B = [None if v == NULL64 else int(v) for v in bb]
    Cn = [None if v == NULL32 else int(v) for v in c]

    def run(prog, use_idx):
        m = len(idx) if use_idx else n
        ws.check(gpu.ws_eval_at_int(C.byref(b.env), prog, di.p if use_idx else None, m, out.p))
        return out.to_array(np.int64, m)

    # one column
    got = run(b.col(0), False)
    assert np.array_equal(got, np.where(a == NULL32, NULL64, a.astype(np.int64)))
    assert np.array_equal(run(b.col(3), True), d[idx])  # DECIMAL: the stored scaled integer
    # exact: a + b * 3 - c, then -(b)
    p = b.prog([(abi.OP_COL, 0), (abi.OP_COL, 1), b.ilit(3), (abi.OP_MUL,), (abi.OP_ADD,), (abi.OP_COL, 2), (abi.OP_SUB,)])
    exp = [NULL64 if (B[i] is None or Cn[i] is None) else int(A[i]) + B[i] * 3 - Cn[i] for i in idx]
    assert np.array_equal(run(p, True), np.array(exp, np.int64))
    p = b.prog([(abi.OP_COL, 1), (abi.OP_NEG,)])
    assert np.array_equal(run(p, False), np.array([NULL64 if v is None else -v for v in B], np.int64))
    # double form: x * 2.5 rounded half away from zero, NaN -> NULL
    p = b.prog([(abi.OP_COL, 4), b.dlit(2.5), (abi.OP_MUL,)])
    y = x * 2.5
    exp = np.where(np.isnan(y), NULL64, llround_vec(np.nan_to_num(y)))
    assert np.array_equal(run(p, False), exp)


def test_distinct_ints(gpu):
    """Sorted distinct values of an exact integer expression within [base, base + span) (NULLs and
    out-of-span values skipped), read back through the host path (small span) and compacted on the
    device (large span); a truncated output still reports the full count; non-exact programs (a
    DECIMAL column) have none (reference gpu.cu:890-907)."""
    H.say("case distinct_ints")
    n = 300_000
    rng = np.random.default_rng(37)
    b, keep, (a, bb, c, d, x) = eval_env(n, rng)
    idx = rng.integers(0, n, 100_000).astype(np.int32)
    di = ws.DevBuf.from_array(idx)
    cap = 1 << 21
    out = np.zeros(cap, np.int64)
    cnt = C.c_int64()
    p = b.prog([(abi.OP_COL, 2), b.ilit(7), (abi.OP_MUL,)])  # c * 7 (NULL skipped)
    cv = c[idx].astype(np.int64)
    vals = np.unique(cv[c[idx] != NULL32] * 7)
    for base, span in ((-200, 500), (-3_000_000, 5_000_000)):
        ws.check(gpu.ws_distinct_ints(C.byref(b.env), p, di.p, len(idx), base, span, ws.ptr(out), cap, C.byref(cnt)))
        exp = vals[(vals >= base) & (vals < base + span)]
        assert cnt.value == len(exp) and np.array_equal(out[:cnt.value], exp), (base, span)
    # a BIGINT expression over all rows in a 40M-value window: device compaction, then a
    # truncated output
    p2 = b.prog([(abi.OP_COL, 1), b.ilit(10 ** 6), (abi.OP_SUB,)])
    bv = bb[bb != NULL64] - 10 ** 6
    lo = int(np.median(bv))
    span = 40_000_000
    exp = np.unique(bv[(bv >= lo) & (bv < lo + span)])
    small = 5
    ws.check(gpu.ws_distinct_ints(C.byref(b.env), p2, None, n, lo, span, ws.ptr(out), small, C.byref(cnt)))
    assert cnt.value == len(exp) and np.array_equal(out[:min(small, len(exp))], exp[:small])
    ws.check(gpu.ws_distinct_ints(C.byref(b.env), p2, None, n, lo, span, ws.ptr(out), cap, C.byref(cnt)))
    assert cnt.value == len(exp) and np.array_equal(out[:cnt.value], exp)
    # DECIMAL single column: not an exact integer expression, no values
    ws.check(gpu.ws_distinct_ints(C.byref(b.env), b.col(3), None, n, -10 ** 7, 2 * 10 ** 7, ws.ptr(out), cap, C.byref(cnt)))
    assert cnt.value == 0


def temp_expect(spec, keys, accs, g):
    """The TempColSpec contract (reference gpu.h:155-166) restated per group."""
    vals, lo, hi = [], None, None
    for r in range(g):
        A = lambda j: accs[j][r]  # noqa: E731
        nul = spec.src == 1 and spec.cnt >= 0 and A(spec.cnt) == 0
        is_int, iv, dv = False, 0, 0.0
        if spec.src == 0:
            k = keys[spec.idx][r]
            if spec.key_int:
                iv, is_int = i64_of(k), True
                nul = iv in (NULL64, NULL32)
            else:
                dv, nul = k, np.isnan(k)
        elif not nul:
            if spec.kind == 1:
                iv, is_int = i64_of(A(spec.idx)), True
            elif spec.kind == 2:
                iv = (i64_of(A(spec.hi)) * 2 ** 32 + i64_of(A(spec.idx)) + 2 ** 63) % 2 ** 64 - 2 ** 63
                is_int = True
            elif spec.kind == 3:
                iv, is_int = int(A(spec.idx)), True
            elif spec.kind == 4:
                if spec.sum_bits:
                    s = float(i64_of(A(spec.hi))) * 4294967296.0 + float(i64_of(A(spec.idx))) if spec.hi >= 0 else float(
                        i64_of(A(spec.idx)))
                else:
                    s = A(spec.idx)
                dv = s / spec.div / A(spec.cnt)
            else:
                dv = A(spec.idx) / spec.div
                nul = np.isnan(dv)
        if spec.mul != 1 and not nul:
            dv, is_int = (float(iv) if is_int else dv) * spec.mul, False
        if spec.out == 2:
            vals.append(np.nan if nul else (float(iv) if is_int else dv))
            continue
        v = 0 if nul else (iv if is_int else llround(dv))
        if not nul:
            lo = v if lo is None else min(lo, v)
            hi = v if hi is None else max(hi, v)
        vals.append((NULL64 if spec.out == 1 else NULL32) if nul else v)
    return vals, lo, hi


@pytest.mark.parametrize("form", ["key_int", "key_dbl", "bits", "split", "count", "avg_bits", "avg", "value", "mul"])
def test_temp_column(gpu, form):
    H.say(f"case temp_column {form}")
    rng = np.random.default_rng(41)
    g = 3001
    k0 = rng.integers(-10 ** 9, 10 ** 9, g).astype(np.int64)
    k0[::17] = NULL64
    k1 = rng.normal(0, 100, g)
    k1[::19] = np.nan
    keys = [k0.view(np.float64), k1]
    cnt = rng.integers(0, 5, g).astype(np.float64)  # COUNT slot (0: the group's argument was all NULL)
    lo = rng.integers(0, 2 ** 32, g).astype(np.int64)
    hi = rng.integers(-2 ** 20, 2 ** 20, g).astype(np.int64)
    dsum = rng.normal(0, 1e4, g)
    accs = [cnt, lo.view(np.float64), hi.view(np.float64), dsum]
    s = abi.TempColSpec()
    s.hi, s.cnt, s.div, s.mul, s.out = -1, -1, 1.0, 1.0, 1
    if form == "key_int":
        s.src, s.idx, s.key_int, s.out = 0, 0, 1, 1
    elif form == "key_dbl":
        s.src, s.idx, s.out = 0, 1, 2
    elif form == "bits":
        s.src, s.idx, s.kind, s.cnt, s.out = 1, 1, 1, 0, 1
    elif form == "split":
        s.src, s.idx, s.hi, s.kind, s.cnt, s.out = 1, 1, 2, 2, 0, 1
    elif form == "count":
        s.src, s.idx, s.kind, s.out = 1, 0, 3, 0
    elif form == "avg_bits":
        s.src, s.idx, s.hi, s.kind, s.cnt, s.sum_bits, s.div, s.out = 1, 1, 2, 4, 0, 1, 100.0, 2
    elif form == "avg":
        s.src, s.idx, s.kind, s.cnt, s.out = 1, 3, 4, 0, 2
    elif form == "value":
        s.src, s.idx, s.kind, s.cnt, s.div, s.out = 1, 3, 0, 0, 10.0, 1
    else:
        s.src, s.idx, s.kind, s.cnt, s.mul, s.out = 1, 1, 1, 0, 0.5, 0
    dk = ws.DevBuf.from_array(np.concatenate(keys))
    da = ws.DevBuf.from_array(np.concatenate(accs))
    mn, mx, out = C.c_double(), C.c_double(), C.c_void_p()
    ws.check(gpu.ws_temp_column(dk.p, 2, da.p, 4, g, C.byref(s), C.byref(mn), C.byref(mx), C.byref(out)))
    w = 4 if s.out == 0 else 8
    got = ws.DevBuf(g * w, adopt=out.value)  # caller-owned device memory: freed with ws_dev_free
    arr = got.to_array(np.int32 if s.out == 0 else np.int64 if s.out == 1 else np.float64, g)
    got.free()
    vals, lo_v, hi_v = temp_expect(s, keys, accs, g)
    if s.out == 2:
        assert np.array_equal(arr, np.array(vals), equal_nan=True)
        assert mn.value == -np.inf and mx.value == np.inf
    else:
        assert np.array_equal(arr.astype(np.int64), np.array(vals, np.int64).astype(np.int32 if s.out == 0 else np.int64))
        assert (mn.value, mx.value) == ((float(lo_v), float(hi_v)) if lo_v is not None else (-np.inf, np.inf))


@pytest.mark.parametrize("case", ["int32_lds", "int32_global", "int64_global", "empty"])
def test_bits_from_column(gpu, case):
    """Bitmap of the non-NULL values of an INT / BIGINT device column offset by base (values outside
    [base, base + nbits) skipped), and whether a NULL was seen (reference gpu.cu:966-995); small
    bitmaps over many rows take the LDS form."""
    H.say(f"case bits_from_column {case}")
    rng = np.random.default_rng(43)
    if case == "int32_lds":
        n, nbits, base = 3_000_000, 4096, -100
        x = rng.integers(-200, 5000, n).astype(np.int32)
        x[5] = NULL32
    elif case == "int32_global":
        n, nbits, base = 200_000, 3_000_000, 1000
        x = rng.integers(0, 4_000_000, n).astype(np.int32)
    elif case == "int64_global":
        n, nbits, base = 200_000, 1_000_003, 2 ** 40
        x = (rng.integers(-1000, 1_200_000, n) + 2 ** 40).astype(np.int64)
        x[::1000] = NULL64
    else:
        n, nbits, base = 0, 100, 0
        x = np.zeros(1, np.int32)
    dx = ws.DevBuf.from_array(x)
    saw, bits = C.c_int(), C.c_void_p()
    ws.check(gpu.ws_bits_from_column(dx.p, 1 if x.dtype == np.int64 else 0, n, base, nbits, C.byref(saw), C.byref(bits)))
    words = max(1, (nbits + 31) // 32)
    db = ws.DevBuf(words * 4, adopt=bits.value)  # caller-owned device memory: freed with ws_dev_free
    out = db.to_array(np.uint32)
    db.free()
    xs = x[:n].astype(np.int64)
    nul = NULL64 if x.dtype == np.int64 else NULL32
    k = xs[xs != nul] - base
    k = k[(k >= 0) & (k < nbits)]
    exp = np.zeros(words * 32, np.uint8)
    exp[k] = 1
    assert np.array_equal(np.unpackbits(out.view(np.uint8), bitorder="little"), exp)
    assert saw.value == int(bool(n) and bool((x[:n] == nul).any()))


@pytest.mark.parametrize("words", [1, 3, 64, 130])
def test_gather_bytes(gpu, words):
    """Rows of `words` 32-bit words gathered by index; -1 gives a NULL row (NaN words) or zeros
    (reference gpu.cu:1041-1049)."""
    H.say(f"case gather_bytes {words}")
    rng = np.random.default_rng(47 + words)
    n, m = 5000, 7777
    src = rng.integers(0, 2 ** 32, (n, words), dtype=np.uint64).astype(np.uint32)
    idx = rng.integers(-1, n, m).astype(np.int32)
    ds, di = ws.DevBuf.from_array(src), ws.DevBuf.from_array(idx)
    out = ws.DevBuf(m * words * 4)
    for nulls, fill in ((1, 0x7fc00000), (0, 0)):
        ws.check(gpu.ws_gather_bytes(ds.p, words * 4, di.p, m, out.p, nulls))
        got = out.to_array(np.uint32).reshape(m, words)
        exp = np.where((idx < 0)[:, None], np.uint32(fill), src[np.maximum(idx, 0)])
        assert np.array_equal(got, exp)


def test_cross_join_and_host_sort(gpu):
    H.say("case cross_join, sort_perm_host, map_rank, set_bits")
    a, bcount = C.c_void_p(), C.c_void_p()
    m = C.c_int64()
    ws.check(gpu.ws_cross_join(301, 7, C.byref(a), C.byref(bcount), C.byref(m)))
    assert m.value == 301 * 7
    ga = ws.DevBuf(m.value * 4, tmp=True, adopt=a.value).to_array(np.int32)
    gb = ws.DevBuf(m.value * 4, tmp=True, adopt=bcount.value).to_array(np.int32)
    assert np.array_equal(ga, np.repeat(np.arange(301), 7)) and np.array_equal(gb, np.tile(np.arange(7), 301))
    assert gpu.ws_cross_join(20000, 20000, C.byref(a), C.byref(bcount), C.byref(m)) == ws.WS_ERR_INTERNAL
    # sort permutation of host columns: stable, NULL-free ints and doubles (NaN first, -0 == 0)
    rng = np.random.default_rng(53)
    for arr, pt in ((rng.integers(-50, 50, 100_000).astype(np.int32), abi.P_I32),
                    (rng.integers(-2 ** 40, 2 ** 40, 70_000).astype(np.int64), abi.P_I64)):
        out = np.zeros(len(arr), np.int32)
        ws.check(gpu.ws_sort_perm_host(ws.ptr(arr), pt, len(arr), ws.ptr(out)))
        assert np.array_equal(out, np.argsort(arr, kind="stable"))
    x = rng.choice([1.5, -2.0, 0.0, -0.0, np.nan, 7.25], 50_000)
    out = np.zeros(len(x), np.int32)
    ws.check(gpu.ws_sort_perm_host(ws.ptr(x), abi.P_F64, len(x), ws.ptr(out)))
    keyed = np.where(np.isnan(x), -np.inf, x + 0.0)
    assert np.array_equal(out, np.argsort(keyed, kind="stable"))
    # dictionary code -> rank (NULL -> -1)
    codes = rng.integers(0, 1000, 40_000).astype(np.float64)
    codes[::9] = np.nan
    rank = rng.permutation(1000).astype(np.int32)
    dk, dr = ws.DevBuf.from_array(codes), ws.DevBuf.from_array(rank)
    ws.check(gpu.ws_map_rank(dk.p, dr.p, len(codes)))
    exp = np.where(np.isnan(codes), -1.0, rank[np.nan_to_num(codes).astype(np.int64)].astype(np.float64))
    assert np.array_equal(dk.to_array(np.float64), exp)
    # deleted-row bits
    bits = np.zeros(4000, np.uint32)
    bits[::5] = 0x80000001
    db = ws.DevBuf.from_array(bits)
    rows = rng.integers(0, 4000 * 32, 30_000).astype(np.int32)
    ws.check(gpu.ws_set_bits(db.p, ws.DevBuf.from_array(rows).p, len(rows)))
    exp = bits.copy()
    np.bitwise_or.at(exp, rows >> 5, (np.uint32(1) << (rows & 31).astype(np.uint32)))
    assert np.array_equal(db.to_array(np.uint32), exp)


@pytest.mark.parametrize("kind", ["i32_u8", "i32_u16", "i64_u32", "f64_dict", "f64_dec16", "f64_dec32", "mixed_raw"])
def test_gather_multi_encoded(gpu, kind):
    """Gathers across ragged batches stored in every segment encoding (ws_encode_column), sorted and
    random indexes, an index array that is not 16-byte aligned; against the decoded values."""
    H.say(f"case gather_multi encoded {kind}")
    rng = np.random.default_rng(59)
    sizes = [70_001, 3, 0, 130_000, 99_999, 1]
    n = sum(sizes)
    if kind == "i32_u8":
        col, pt, t, want = rng.integers(-100, 150, n).astype(np.int32), abi.P_I32, abi.T_I32N, 1
        col[::31] = NULL32
    elif kind == "i32_u16":
        col, pt, t, want = rng.integers(1000, 61000, n).astype(np.int32), abi.P_I32, abi.T_I32N, 2
    elif kind == "i64_u32":
        col, pt, t, want = (rng.integers(0, 3 * 10 ** 9, n) + 10 ** 13).astype(np.int64), abi.P_I64, abi.T_I64N, 3
        col[::29] = NULL64
    elif kind == "f64_dict":
        col, pt, t, want = rng.choice(rng.normal(0, 1, 200), n), abi.P_F64, abi.T_F64, 4
        col[::37] = np.nan
    elif kind == "f64_dec16":
        col, pt, t, want = rng.integers(0, 60000, n) / 100.0, abi.P_F64, abi.T_F64, 5
        col[::41] = np.nan
    elif kind == "f64_dec32":
        col, pt, t, want = rng.integers(-10 ** 8, 10 ** 8, n) / 1000.0, abi.P_F64, abi.T_F64, 6
    else:
        col, pt, t, want = rng.integers(-2 ** 30, 2 ** 30, n).astype(np.int32), abi.P_I32, abi.T_I32N, 0
    tbl = H.Table({"c": col}, sizes)
    if want:
        tbl.encode("c")
        assert tbl.enc["c"][3][1] == want, kind  # the largest batch; tiny ones may pick another encoding
    bound = H.Bound(tbl, [("c", t, 0)], True)
    boff = np.array(tbl.offs[:-1], np.int64)
    for idx in (np.sort(rng.choice(n, 123_457, replace=False)).astype(np.int32), rng.integers(0, n, 77_777).astype(np.int32)):
        di = ws.DevBuf.from_array(np.concatenate([np.zeros(1, np.int32), idx]))  # idx at byte 4: unaligned
        out = ws.DevBuf(len(idx) * col.itemsize)
        ws.check(gpu.ws_gather_multi(C.cast(bound.views, C.c_void_p), ws.ptr(boff), len(sizes), pt, col.itemsize, di.p + 4,
                                     len(idx), out.p))
        got = out.to_array(col.dtype)
        exp = col[idx]
        if col.dtype == np.float64:
            assert np.array_equal(got.view(np.int64), exp.view(np.int64)) or np.array_equal(got, exp, equal_nan=True)
        else:
            assert np.array_equal(got, exp)
    # the plain index gather with an unaligned index pointer (scalar index loads)
    if col.dtype != np.float64:
        src = ws.DevBuf.from_array(col)
        idx = rng.integers(0, n, 10_001).astype(np.int32)
        di = ws.DevBuf.from_array(np.concatenate([np.zeros(1, np.int32), idx]))
        out = ws.DevBuf(len(idx) * col.itemsize)
        ws.check(gpu.ws_gather(src.p, pt, di.p + 4, len(idx), out.p))
        assert np.array_equal(out.to_array(col.dtype), col[idx])
