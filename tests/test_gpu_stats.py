"""GPU column statistics (sizing inputs) against the CPU oracle: live / NULL counts, exact min / max
and every HyperLogLog register bit-exact, per batch and merged, for raw, nullable, encoded,
decimal and double columns with deleted rows."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu


def _same(a: abi.ColStats, b: abi.ColStats):
    assert (a.rows, a.nulls, a.imin, a.imax) == (b.rows, b.nulls, b.imin, b.imax)
    for x, y in ((a.dmin, b.dmin), (a.dmax, b.dmax)):
        assert (np.isnan(x) and np.isnan(y)) or x == y
    assert bytes(a.hll) == bytes(b.hll)


def test_column_stats_match_oracle(gpu):
    H.say("case column stats")
    n = 300_000
    rng = np.random.default_rng(31)
    cols = {
        "i32": rng.integers(-5000, 5000, n).astype(np.int32),
        "i64n": rng.integers(-(1 << 50), 1 << 50, n).astype(np.int64),
        "f64": np.round(rng.normal(0, 100, n), 1),
        "enc": rng.integers(0, 60_000, n).astype(np.int32),
        "dec": rng.integers(100, 10_000_000, n).astype(np.int64),
    }
    cols["i64n"][rng.random(n) < 0.07] = abi.NULL_I64
    cols["f64"][rng.random(n) < 0.03] = np.nan
    cols["f64"][rng.random(n) < 0.01] = -0.0
    cols["dec"][rng.random(n) < 0.02] = abi.NULL_I64
    deleted = rng.random(n) < 0.1
    t = H.Table(cols, [100_000, 0, 131_072, 68_928], deleted)
    t.encode("enc")
    slots = [("i32", abi.T_I32, 0), ("i64n", abi.T_I64N, 0), ("f64", abi.T_F64, 0), ("enc", abi.T_I32, 0),
             ("dec", abi.T_DECIMAL, 2)]
    dev, host = H.Bound(t, slots, True), H.Bound(t, slots, False)
    hdels = t.del_bitmaps()
    L = gpu
    for s in range(len(slots)):
        merged_g = merged_o = None
        for b in range(len(t.sizes)):
            nb = t.sizes[b]
            g = abi.ColStats()
            ws.check(L.ws_column_stats(C.byref(dev.views[b * len(slots) + s]), nb, dev.dels[b], C.byref(g)))
            o = H.oracle_column_stats(host.views[b * len(slots) + s], nb, hdels[b])
            if nb:
                _same(g, o)
            if merged_g is None:
                merged_g, merged_o = g, o
            else:
                L.ws_col_stats_merge(C.byref(merged_g), C.byref(g))
                L.ws_col_stats_merge(C.byref(merged_o), C.byref(o))
        _same(merged_g, merged_o)
        live = ~deleted
        assert merged_g.rows == int(live.sum())
        name = slots[s][0]
        x = cols[name][live]
        if name in ("i64n", "dec"):
            x = x[x != abi.NULL_I64]
        elif name == "f64":
            x = x[~np.isnan(x)]
        true_d = len(np.unique(x + 0.0 if name == "f64" else x))
        est = L.ws_hll_estimate(merged_g.hll)
        assert abs(est - true_d) <= 4 * 0.023 * true_d + 1, (name, est, true_d)
