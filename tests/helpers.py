"""Shared test helpers: tables split into batches, plans built with the engine's slot conventions,
GPU runs through the C ABI and the CPU oracle on the same host data."""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import warpstore_amd as ws  # noqa: E402
from warpstore_amd import abi  # noqa: E402

P, I, I64, D = C.c_void_p, C.c_int, C.c_int64, C.c_double

_oracle = None
_ref = None


def oracle():
    """The CPU oracle library (oracle/liboracle.so)."""
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import build as ob  # noqa: E402
            ob.build_oracle()
        L = C.CDLL(path)
        L.or_eval_prog.restype = D
        L.or_eval_prog.argtypes = [C.POINTER(abi.ExprEnv), P, abi.Prog, I64]
        for f in ("or_stored_int",):
            getattr(L, f).restype = I64
            getattr(L, f).argtypes = [C.POINTER(abi.ColView), I64]
        for f in ("or_stored_dbl", "or_value"):
            getattr(L, f).restype = D
            getattr(L, f).argtypes = [C.POINTER(abi.ColView), I64]
        L.or_call1.restype = D
        L.or_call1.argtypes = [I, D]
        L.or_call2.restype = D
        L.or_call2.argtypes = [I, D, D]
        L.or_days_from_civil.restype = I64
        L.or_days_from_civil.argtypes = [I64, C.c_uint, C.c_uint]
        L.or_civil_from_days.argtypes = [I64, C.POINTER(I64), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
        L.or_checked.restype = I64
        L.or_checked.argtypes = [I, I64, I64, C.POINTER(I)]
        L.or_aggregate.restype = I64
        L.or_aggregate.argtypes = [C.POINTER(abi.AggPlan), I, C.POINTER(abi.ExprEnv), I, P, P, P, I, P, P, P, I64]
        L.or_select.restype = I64
        L.or_select.argtypes = [C.POINTER(abi.ExprEnv), I, abi.Prog, I, P, P, P, I, P, P]
        L.or_join.restype = I64
        L.or_join.argtypes = [C.POINTER(abi.ColView), I64, C.POINTER(abi.ColView), I64, I, P, P, I64]
        L.or_sort_perm.argtypes = [P, P, I, I64, P]
        L.or_remap_codes.argtypes = [P, I64, P, I64, P]
        _oracle = L
    return _oracle


def ref_lib():
    """The reference's own devtypes.h compiled on the host (oracle/_ref), or None."""
    global _ref
    if _ref is None:
        path = os.path.join(ROOT, "oracle", "_ref", "libref_devtypes.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_eval_prog.restype = D
        L.ref_eval_prog.argtypes = [P, P, C.c_int16, C.c_int16, I64]
        L.ref_load_col.restype = D
        L.ref_load_col.argtypes = [P, I64]
        L.ref_col_int.restype = I64
        L.ref_col_int.argtypes = [P, I64]
        L.ref_col_dbl.restype = D
        L.ref_col_dbl.argtypes = [P, I64]
        L.ref_call1.restype = D
        L.ref_call1.argtypes = [I, D]
        L.ref_call2.restype = D
        L.ref_call2.argtypes = [I, D, D]
        L.ref_days_from_civil.restype = I64
        L.ref_days_from_civil.argtypes = [I64, C.c_uint, C.c_uint]
        L.ref_civil_from_days.argtypes = [I64, C.POINTER(I64), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
        L.ref_checked.restype = I64
        L.ref_checked.argtypes = [I, I64, I64, C.POINTER(I)]
        L.ref_sizeof.restype = I64
        L.ref_offsetof.restype = I64
        _ref = L
    return _ref


def encode_column(x: np.ndarray):
    """(encoded bytes, encoding) of one segment column through the library's encoder."""
    x = np.ascontiguousarray(x)
    ptype = {np.dtype(np.int32): abi.P_I32, np.dtype(np.int64): abi.P_I64, np.dtype(np.float64): abi.P_F64}[x.dtype]
    L = ws.lib()
    out = np.zeros(L.ws_encoded_bytes_max(len(x)), np.uint8)
    e, nbytes = C.c_int(), C.c_size_t()
    ws.check(L.ws_encode_column(ws.ptr(x), ptype, len(x), ws.ptr(out), len(out), C.byref(e), C.byref(nbytes)))
    return out[:nbytes.value].copy(), e.value


def say(*a) -> None:
    """Names the case before a GPU step (flushed, so a lost box still shows where it was)."""
    print(*a, flush=True)


# ------------------------------------------------------------------ tables and batches

class Table:
    """Columns (numpy) cut into batches like segments. `enc` maps a column name to an encoded
    copy (frame of reference) used instead of the raw column."""

    def __init__(self, cols: dict, batch_rows, deleted: np.ndarray | None = None):
        self.cols = {k: np.ascontiguousarray(v) for k, v in cols.items()}
        self.n = len(next(iter(self.cols.values())))
        if isinstance(batch_rows, int):
            sizes, r = [], self.n
            while r > 0:
                sizes.append(min(batch_rows, r))
                r -= sizes[-1]
        else:
            sizes = list(batch_rows)
        assert sum(sizes) == self.n
        self.sizes = sizes
        self.offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.deleted = deleted  # bool per row, or None
        self.enc = {}

    def encode(self, name: str) -> int:
        """Stores column `name` per batch in its segment encoding (ws_encode_column: frame of
        reference for integers, dictionary or decimal codes for doubles); returns the last
        batch's encoding."""
        x = self.cols[name]
        pieces = []
        e = None
        for b in range(len(self.sizes)):
            part = np.ascontiguousarray(x[self.offs[b]:self.offs[b + 1]])
            out, e = encode_column(part)
            assert e, "column does not fit an encoding"
            pieces.append((out, e))
        self.enc[name] = pieces
        return e

    def del_bitmaps(self):
        if self.deleted is None:
            return [None] * len(self.sizes)
        out = []
        for b in range(len(self.sizes)):
            d = self.deleted[self.offs[b]:self.offs[b + 1]]
            words = np.zeros((len(d) + 31) // 32, np.uint32)
            idx = np.nonzero(d)[0]
            np.bitwise_or.at(words, idx >> 5, (np.uint32(1) << (idx & 31).astype(np.uint32)))
            out.append(words)
        return out


class Bound:
    """A table bound to slot views (one ColView per (batch, slot)) on the host or the device."""

    def __init__(self, t: Table, slots: list, device: bool):
        # slots: list of (column name, type code, scale)
        self.t = t
        self.keep = []
        nb = len(t.sizes)
        self.views = (abi.ColView * (nb * len(slots)))()
        dels = t.del_bitmaps()
        self.dels = (C.c_void_p * nb)()
        for b in range(nb):
            for s, (name, tcode, scale) in enumerate(slots):
                enc = 0
                if name in t.enc:
                    arr, enc = t.enc[name][b]
                else:
                    arr = np.ascontiguousarray(t.cols[name][t.offs[b]:t.offs[b + 1]])
                p = self._place(arr, device)
                v = self.views[b * len(slots) + s]
                v.p, v.t, v.scale, v.enc, v.off = p, tcode, scale, enc, 0
            if dels[b] is not None:
                self.dels[b] = self._place(dels[b], device)
        self.nslots = len(slots)

    def _place(self, arr: np.ndarray, device: bool) -> int:
        if device:
            b = ws.DevBuf.from_array(arr)
            self.keep.append(b)
            return b.p
        a = np.ascontiguousarray(arr)
        self.keep.append(a)
        return a.ctypes.data

    def batch_views(self, b: int):
        return self.views[b * self.nslots:(b + 1) * self.nslots]


# ------------------------------------------------------------------ aggregation runs

def env_with_batch(env: abi.ExprEnv, bound: Bound, b: int) -> abi.ExprEnv:
    e = abi.ExprEnv.from_buffer_copy(env)
    for s in range(bound.nslots):
        e.cols[s] = bound.views[b * bound.nslots + s]
    return e


LAST_RUN = {"partitioned": False, "retries": 0}


def gpu_aggregate(plan, env, bound: Bound, cap: int = 1 << 16, dense=None, jit=None, fetch=True, windows=False,
                  strategy=None, expected=0, features=None):
    """Runs GpuAggregator through the C ABI, retrying bigger on overflow like the engine
    (engine_select.cpp:1514-1522). Returns (keys g x nk, accs g x na, hi, cap used).
    strategy: None (library default) or ws_agg_set_strategy's value, with `expected` groups.
    LAST_RUN records whether the final run was partitioned and how many retries it took."""
    L = ws.lib()
    LAST_RUN["retries"] = 0
    while True:
        h = C.c_void_p()
        ws.check(L.ws_agg_create(C.byref(plan), cap, C.byref(dense) if dense is not None else None, C.byref(h)))
        if strategy is not None:
            ws.check(L.ws_agg_set_strategy(h, strategy, expected))
        if features is not None:
            ws.check(L.ws_agg_set_features(h, ws.ptr(features)))
        try:
            for b in range(len(bound.t.sizes)):
                e = env_with_batch(env, bound, b)
                ws.check(L.ws_agg_add_batch(h, C.byref(e), bound.nslots, bound.t.sizes[b], bound.dels[b]))
                if windows:
                    ws.check(L.ws_agg_run_window(h, C.byref(env), jit))
            ws.check(L.ws_agg_run(h, C.byref(env), jit))
            ov = C.c_int()
            ws.check(L.ws_agg_overflowed(h, C.byref(ov)))
            LAST_RUN["partitioned"] = bool(L.ws_agg_partitioned(h))
            if ov.value:
                gc = C.c_int64()
                ws.check(L.ws_agg_group_count(h, C.byref(gc)))
                want = max(cap * 4, gc.value * 4)
                cap = 1
                while cap < want:
                    cap <<= 1
                LAST_RUN["retries"] += 1
                if not fetch:
                    L.ws_agg_destroy(h)
                continue
            if not fetch:
                return h, cap
            g = C.c_int64()
            ws.check(L.ws_agg_fetch_begin(h, C.byref(g)))
            nk, na = plan.nkeys, plan.naggs
            keys = np.zeros(max(g.value * nk, 1))
            accs = np.zeros(max(g.value * na, 1))
            ws.check(L.ws_agg_fetch_copy(h, ws.ptr(keys), ws.ptr(accs)))
            hi = np.zeros(na, np.int64)
            L.ws_agg_sum_high(h, ws.ptr(hi))
            return keys[:g.value * nk].reshape(g.value, nk), accs[:g.value * na].reshape(g.value, na), hi, cap
        finally:
            if fetch:
                L.ws_agg_destroy(h)


def oracle_aggregate(plan, env, bound: Bound, dense: bool = False, max_groups: int = 1 << 22):
    O = oracle()
    nb = len(bound.t.sizes)
    n = np.array(bound.t.sizes, np.int64)
    nk, na = plan.nkeys, plan.naggs
    keys = np.zeros(max(max_groups * nk, 1))
    accs = np.zeros(max(max_groups * na, 1))
    hi = np.zeros(na, np.int64)
    g = O.or_aggregate(C.byref(plan), 1 if dense else 0, C.byref(env), nb, ws.ptr(n), C.cast(bound.dels, C.c_void_p),
                       C.cast(bound.views, C.c_void_p), bound.nslots, ws.ptr(keys), ws.ptr(accs), ws.ptr(hi), max_groups)
    assert g >= 0, "oracle group buffer too small"
    return keys[:g * nk].reshape(g, nk), accs[:g * na].reshape(g, na), hi


def sort_groups(keys: np.ndarray, accs: np.ndarray):
    """Orders groups by their key bit patterns (as int64, lexicographic), like the oracle."""
    if len(keys) == 0:
        return keys, accs
    kb = keys.view(np.int64)
    order = np.lexsort(tuple(kb[:, k] for k in reversed(range(kb.shape[1]))))
    return keys[order], accs[order]


def assert_groups_equal(plan, gk, ga, ok, oa, fp_rel=1e-12, dense=False):
    """Integer, count, MIN / MAX and key results bit-exact; double sums within fp_rel."""
    gk, ga = sort_groups(gk, ga)
    assert gk.shape == ok.shape, (gk.shape, ok.shape)
    assert np.array_equal(gk.view(np.int64), ok.view(np.int64)), "group keys differ"
    for a in range(plan.naggs):
        g, o = ga[:, a], oa[:, a]
        fp_sum = plan.ops[a] == abi.AGG_SUM and not plan.agg_int[a]
        if fp_sum:
            np.testing.assert_allclose(g, o, rtol=fp_rel, atol=0, err_msg=f"slot {a} (double sum)")
        else:
            bad = np.nonzero(g.view(np.int64) != o.view(np.int64))[0]
            assert len(bad) == 0, f"slot {a}: {len(bad)} groups differ, e.g. gpu {g[bad[:3]]} oracle {o[bad[:3]]}"


# ------------------------------------------------------------------ query plans (engine slot layout)

def plan_c1(b: abi.EnvBuilder, sk: int, sv: int, sp: int, k_lt: int = 500, v_gt: int = 1000):
    """SELECT k, COUNT(*), SUM(v), MIN(v), MAX(v), AVG(p) FROM t WHERE k < 500 AND v > 1000 GROUP BY k.
    Slots (engine_select.cpp:647-724, 1130-1149): 0 COUNT(*); 1/2 SUM(v) low / high 32-bit halves
    (BIGINT SUM with GROUP BY); 3 MIN(v); 4 MAX(v) (exact int64 bits); 5 SUM(p) for AVG."""
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sk), b.ilit(k_lt), (abi.OP_LT,), (abi.OP_COL, sv), b.ilit(v_gt), (abi.OP_GT,),
                         (abi.OP_AND,)])
    plan.nkeys = 1
    plan.keys[0] = b.col(sk)
    plan.key_int[0] = 1
    ops = [abi.AGG_COUNT, abi.AGG_SUM, abi.AGG_SUM, abi.AGG_MIN, abi.AGG_MAX, abi.AGG_SUM]
    plan.naggs = len(ops)
    for a, o in enumerate(ops):
        plan.ops[a] = o
    plan.agg_int[0] = 1
    for a in (1, 2, 3, 4):
        plan.aggs[a] = b.col(sv)
        plan.agg_int[a] = 1
        plan.agg_bits[a] = 1
    plan.agg_part[1], plan.agg_part[2] = 1, 2
    plan.aggs[5] = b.col(sp)
    return plan


def dense_spec(plan, kmin: list, rng: list) -> abi.DenseSpec:
    ds = abi.DenseSpec()
    G = 1
    for k in range(plan.nkeys):
        ds.kmin[k], ds.range[k], ds.stride[k] = kmin[k], rng[k], G
        G *= rng[k]
    ds.G = G
    off = 0
    for a in range(plan.naggs):
        ds.kind[a] = abi.dense_kind(plan.ops[a], bool(plan.agg_int[a]))
        ds.slot_off[a] = off
        off += (G * (4 if ds.kind[a] == abi.DK_COUNT else 8) + 7) // 8 * 8
    ds.copy_bytes = off
    ds.copies = max(1, min(8, (48 * 1024) // max(off, 1))) if off <= 48 * 1024 else 0
    ds.priv = 1 if G * plan.naggs <= 88 else 0
    return ds


def gen_c1(n: int, seed: int = 1):
    g = ws.datagen()
    k = np.empty(n, np.int32)
    v = np.empty(n, np.int64)
    p = np.empty(n, np.float64)
    g.wsg_c1(seed, n, ws.ptr(k), ws.ptr(v), ws.ptr(p))
    return k, v, p


def gen_c2(n: int, seed: int = 2, n_users: int = 1000000, n_urls: int = 100000):
    g = ws.datagen()
    user = np.empty(n, np.int64)
    region = np.empty(n, np.int32)
    date = np.empty(n, np.int32)
    url = np.empty(n, np.int32)
    dur = np.empty(n, np.int32)
    g.wsg_c2(seed, n, n_users, n_urls, ws.ptr(user), ws.ptr(region), ws.ptr(date), ws.ptr(url), ws.ptr(dur))
    return user, region, date, url, dur


# ------------------------------------------------------------------ query-level oracles

def oracle_c2(user, region, date, url, dur, date_lt=19723 + 300, region_lt=224):
    O = oracle()
    O.or_c2_query.restype = I64
    O.or_c2_query.argtypes = [P, P, P, P, P, I64, C.c_int32, C.c_int32, P, I64]
    cap = len(np.unique(user)) + 1
    out = np.zeros((cap, 4), np.int64)
    g = O.or_c2_query(ws.ptr(user), ws.ptr(region), ws.ptr(date), ws.ptr(url), ws.ptr(dur), len(user), date_lt, region_lt,
                      ws.ptr(out), cap)
    assert g >= 0
    return out[:g]


def oracle_c3(fk, qty, price, flag, dkey, dattr, attr_lt=50):
    O = oracle()
    O.or_c3_query.restype = I64
    O.or_c3_query.argtypes = [P, P, P, P, I64, P, P, I64, C.c_int32, P, I64]
    out = np.zeros((256, 4), np.int64)
    g = O.or_c3_query(ws.ptr(fk), ws.ptr(qty), ws.ptr(price), ws.ptr(flag), len(fk), ws.ptr(dkey), ws.ptr(dattr), len(dkey),
                      attr_lt, ws.ptr(out), 256)
    assert g >= 0
    return out[:g]


def check_topn(got: np.ndarray, full: np.ndarray) -> None:
    """ORDER BY count DESC LIMIT n parity with ties: the counts equal the oracle's top counts,
    and every returned row is exactly the oracle's row for that key (tie order is unspecified)."""
    n = len(got)
    assert n == min(len(full), n)
    assert np.array_equal(got[:, 1], full[:n, 1]), (got[:, 1], full[:n, 1])
    by_key = {int(r[0]): r for r in full}
    for r in got:
        assert np.array_equal(r, by_key[int(r[0])]), (r, by_key.get(int(r[0])))


def gen_c3(n_fact: int, n_dim: int, seed: int = 3):
    g = ws.datagen()
    dkey = np.empty(n_dim, np.int32)
    dattr = np.empty(n_dim, np.int32)
    g.wsg_c3_dim(seed, n_dim, ws.ptr(dkey), ws.ptr(dattr))
    fk = np.empty(n_fact, np.int32)
    qty = np.empty(n_fact, np.int64)
    price = np.empty(n_fact, np.int64)
    flag = np.empty(n_fact, np.int32)
    g.wsg_c3_fact(seed, n_fact, n_dim, ws.ptr(fk), ws.ptr(qty), ws.ptr(price), ws.ptr(flag))
    return fk, qty, price, flag, dkey, dattr


# ------------------------------------------------------------------ column statistics

def oracle_column_stats(view: abi.ColView, n: int, del_words=None) -> abi.ColStats:
    """The oracle's column statistics of one host batch, as a ws_col_stats structure."""
    O = oracle()
    O.or_column_stats.restype = None
    O.or_column_stats.argtypes = [C.POINTER(abi.ColView), I64, P, P, P, P]
    counts = np.zeros(4, np.int64)
    mm = np.zeros(2)
    regs = np.zeros(2048, np.uint8)
    O.or_column_stats(C.byref(view), n, ws.ptr(del_words) if del_words is not None else None, ws.ptr(counts), ws.ptr(mm),
                      ws.ptr(regs))
    st = abi.ColStats()
    st.rows, st.nulls, st.imin, st.imax = (int(x) for x in counts)
    st.dmin, st.dmax = float(mm[0]), float(mm[1])
    C.memmove(st.hll, regs.ctypes.data, 2048)
    return st


def host_view(arr: np.ndarray, tcode: int, scale: int = 0) -> abi.ColView:
    v = abi.ColView()
    v.p, v.t, v.scale = arr.ctypes.data, tcode, scale
    return v
