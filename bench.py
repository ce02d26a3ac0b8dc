"""Benchmark of the analytic hot path on one MI355X.

Workload (one step = one query over GPU-resident segments, 2^20 rows each):
  web-analytics table, 10M rows, seed 2: user_id int64 Zipf(1.1) over 1e6 ids, region int16 in
  [0,256) stored frame-of-reference encoded (uint16 codes), event_date int32 over 365 days, url
  int32 dictionary code Zipf(1.1) over 1e5, dur int32 lognormal(ln 30000, 1);
  SELECT user_id, COUNT(*) c, SUM(dur), COUNT(DISTINCT url) FROM h
  WHERE event_date < 19723+300 AND region < 224 GROUP BY user_id ORDER BY c DESC LIMIT 10
run through the engine-side driver exactly as the planner issues it (compiled phase-1 plan,
two-phase COUNT DISTINCT, GPU top-N), with results checked against the CPU oracle.

value = GB/s scanned: 22 bytes of referenced columns per row (8 + 2 + 4 + 4 + 4) x 10M rows per
query, over the wall time of the timed steps. Inputs are resident in HBM before timing starts;
steps rotate over copies of the table that together exceed the 256 MB Infinity Cache.
roofline: the dominant kernel (most GPU time over the warm-up queries, among the scan kernels of
the compiled phase-1 aggregation) timed live with HIP events on its launch stream during the timed
steps; achieved = its algorithmic bytes per launch (the referenced columns of the 10M rows it scans:
22 B/row, or 18 B/row for the partition count pass, which does not read `dur`) / mean duration.
cpu_baseline: the CPU oracle's restatement of the same query on this host, one thread.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GB/s scanned (% of HBM peak) for filter+agg/join; queries/s at p99, 1 GPU"
HBM_PEAK_GBS = 8000.0
ROWS = 10_000_000
BYTES_PER_ROW = 22


def gen_c2(n, seed=2):
    import warpstore_amd as ws
    g = ws.datagen()
    user = np.empty(n, np.int64)
    region = np.empty(n, np.int32)
    date = np.empty(n, np.int32)
    url = np.empty(n, np.int32)
    dur = np.empty(n, np.int32)
    g.wsg_c2(seed, n, 1_000_000, 100_000, ws.ptr(user), ws.ptr(region), ws.ptr(date), ws.ptr(url), ws.ptr(dur))
    return user, region, date, url, dur


def oracle_lib():
    import ctypes as C
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    O = C.CDLL(path)
    P, I64 = C.c_void_p, C.c_int64
    O.or_c2_query.restype = I64
    O.or_c2_query.argtypes = [P, P, P, P, P, I64, C.c_int32, C.c_int32, P, I64]
    return O


def oracle_c2(O, cols, n):
    user, region, date, url, dur = (c[:n] for c in cols)
    cap = n + 1
    out = np.zeros((cap, 4), np.int64)
    g = O.or_c2_query(user.ctypes.data, region.ctypes.data, date.ctypes.data, url.ctypes.data, dur.ctypes.data, n,
                      19723 + 300, 224, out.ctypes.data, cap)
    return out[:g]


def check_top(top, full):
    if not np.array_equal(top[:, 1], full[:len(top), 1]):
        return False
    by_key = {int(r[0]): r for r in full[:max(len(top) * 50, 1000)]}
    return all(int(r[0]) in by_key and np.array_equal(r, by_key[int(r[0])]) for r in top)


def cpu_baseline(cols, target_s: float):
    """Times the oracle on a bounded prefix sample of the same table (about target_s seconds)."""
    O = oracle_lib()
    n = 1_000_000
    t0 = time.perf_counter()
    oracle_c2(O, cols, n)
    dt = time.perf_counter() - t0
    n2 = int(min(ROWS, max(n, n * target_s / max(dt, 1e-3))))
    if n2 > n:
        t0 = time.perf_counter()
        oracle_c2(O, cols, n2)
        dt = time.perf_counter() - t0
        n = n2
    return {"value": round(n * BYTES_PER_ROW / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"first {n} rows of the same table, same query, single-thread CPU oracle ({dt:.2f} s)"}


# Scan kernels of the phase-1 aggregation that can dominate, with their algorithmic bytes per launch
# and the kernel symbol they launch.
PHASE1 = "[aggregate (distinct pairs)]"  # the plan-cache kind of the phase-1 plan: its launches' profile tag
ALGO = {
    "agg_hash_jit" + PHASE1: (22 * ROWS, "ws_q_grouped", "compiled filter + shared-table hash GROUP BY (user_id, url)"),
    "agg_part_scatter_jit" + PHASE1: (22 * ROWS, "ws_q_grouped_pscatter",
                                      "compiled filter + partition scatter of (user_id, url) partial groups"),
    "agg_part_count_jit" + PHASE1: (18 * ROWS, "ws_q_grouped_pcount", "compiled filter + partition count"),
}


def pmc_traffic(symbol: str, name: str = "pmc_c2_web_analytics.json"):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter runs (or None)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            d = json.load(f)
        return d["kernels"][symbol]["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--copies", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a tuning knob of the physical plans (ws_tune_set) before running")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 1 GiB device copy kernel (known bytes, for counter calibration)")
    ap.add_argument("--workload", choices=["web", "batch", "serve", "join"], default="web",
                    help="web: the headline 10M-row web-analytics query (default); batch: 256 small GROUP BY "
                         "queries, one batched launch vs per-query; serve: the open-loop mixed stream, "
                         "continuous batching vs per-query serving")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path's multi-GPU sharding is a design note, no collective runs here)", file=sys.stderr)
        return 2
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "join":
        return bench_join(args)
    if args.workload == "serve":
        return bench_serve(args)

    import warpstore_amd as ws
    from warpstore_amd import abi, queries as Q
    L = ws.lib()
    ws.check(L.ws_init())
    L.ws_jit_set_mode(abi.JIT_SYNC)
    tuned = {}
    for kv in args.tune:
        key, val = kv.split("=", 1)
        ws.check(L.ws_tune_set(key.encode(), int(val)))
        tuned[key] = int(val)
    if args.calibrate:
        src, dst = ws.DevBuf(1 << 30), ws.DevBuf(1 << 30)
        ws.check(L.ws_copy_kernel(dst.p, src.p, 1 << 30))
        L.ws_sync()
        src.free()
        dst.free()

    cols = gen_c2(ROWS)
    types = [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32]
    tables = [Q.DeviceTable(list(cols), types, [0, 1, 0, 0, 0]) for _ in range(args.copies)]

    top = None
    for w in range(max(args.warmup, 1)):
        top, stats = Q.c2(tables[w % args.copies], 10)
    L.ws_sync()
    # the dominant kernel, from one more (untimed) query after the warm-ups
    L.ws_prof_reset()
    L.ws_prof_only(b"")
    L.ws_prof_enable(1)
    top, stats = Q.c2(tables[0], 10)
    L.ws_sync()
    warm = ws.prof_rows()
    L.ws_prof_enable(0)
    dom = max((n for n in ALGO if n in warm), key=lambda n: warm[n][1], default="agg_hash_jit" + PHASE1)
    algo_bytes, symbol, what = ALGO[dom]

    L.ws_prof_reset()
    L.ws_prof_only(dom.encode())
    L.ws_prof_enable(1)
    lat = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        top, stats = Q.c2(tables[s % args.copies], 10)
        lat.append(time.perf_counter() - ts)
    L.ws_sync()
    t1 = time.perf_counter()
    prof = ws.prof_rows()
    L.ws_prof_enable(0)

    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    gbs = ROWS * BYTES_PER_ROW * args.steps / elapsed / 1e9
    launches, total_ms = prof.get(dom, (0, 0.0))
    kern_ms = total_ms / max(launches, 1)
    achieved = algo_bytes / (kern_ms / 1e3) / 1e9 if kern_ms > 0 else 0.0

    full = oracle_c2(oracle_lib(), cols, ROWS)
    parity = bool(check_top(top, full))

    result = {
        "metric": METRIC,
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded generator, seed 2)",
        "config": {
            "workload": "web-analytics GROUP BY user_id + COUNT DISTINCT url + top 10, 10M rows",
            "rows": ROWS, "bytes_per_row": BYTES_PER_ROW, "segments_rows": 1 << 20, "table_copies": args.copies,
            "distinct_pairs": int(stats[0]), "groups": int(stats[1]), "phase1_capacity": int(stats[2]),
            "overflow_retries": int(stats[3]),
            **({"tuning": tuned} if tuned else {}),
        },
        "pct_of_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 3),
        "queries_per_s": round(args.steps / elapsed, 3),
        "p50_ms": round(float(np.percentile(np.array(lat) * 1e3, 50)), 4),
        "p99_ms": round(float(np.percentile(np.array(lat) * 1e3, 99)), 4),
        "parity_vs_oracle": parity,
        "roofline": {
            "kernel": f"{symbol} ({what})",
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": pmc_traffic(symbol),
            "algorithmic_bytes": algo_bytes, "mean_launch_ms": round(kern_ms, 4), "launches": int(launches),
        },
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cols, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0 if parity else 1


def bench_batch(args) -> int:
    """256 filter + GROUP BY queries (seed 4): rows lognormal in [4096, 4M], keys with 10 to 1e5
    distinct values, WHERE val > theta_q; 12 B/row scanned. One step = all 256 queries, run as one
    batched launch + one batched fetch, and (separately timed) one by one through the per-query path."""
    import warpstore_amd as ws
    from warpstore_amd import abi, queries as Q
    L = ws.lib()
    ws.check(L.ws_init())
    L.ws_jit_set_mode(abi.JIT_SYNC)
    for kv in args.tune:
        key, val = kv.split("=", 1)
        ws.check(L.ws_tune_set(key.encode(), int(val)))
    s = Q.C4Set(seed=4, nq=256)
    res = {}
    for mode in ("batched", "per_query"):
        for _ in range(max(args.warmup, 1)):
            d, g = s.run(batched=mode == "batched")
        L.ws_sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            d, g = s.run(batched=mode == "batched")
        L.ws_sync()
        dt = (time.perf_counter() - t0) / args.steps
        res[mode] = (dt, d, g)
    same = bool(np.array_equal(res["batched"][1], res["per_query"][1]))
    bt, pt = res["batched"][0], res["per_query"][0]
    out = {
        "metric": "GB/s scanned by 256 batched filter+GROUP BY queries", "value": round(s.rows * 12 / bt / 1e9, 3),
        "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(bt * 1e3, 4),
        "higher_is_better": True, "dtype": "int64", "data": "synthetic (seeded generator, seed 4)",
        "config": {"workload": "256 filter + GROUP BY queries, one batched launch", "rows": s.rows, "bytes_per_row": 12},
        "queries_per_s": round(256 / bt, 1), "per_query_ms_per_step": round(pt * 1e3, 4),
        "per_query_queries_per_s": round(256 / pt, 1), "speedup_vs_per_query": round(pt / bt, 3),
        "results_equal_per_query": same, "groups": int(res["batched"][2].sum()),
    }
    s.free()
    print(json.dumps(out), flush=True)
    return 0 if same else 1


JOIN_FACT, JOIN_DIM = 60_000_000, 1_500_000
JOIN_BYTES = 24  # fact columns the query references: fk 4 + qty 8 + price 8 + flag 4


def oracle_c3(O, fact, dim):
    import ctypes as C
    fk, qty, price, flag = fact
    dkey, dattr = dim
    P, I64 = C.c_void_p, C.c_int64
    O.or_c3_query.restype = I64
    O.or_c3_query.argtypes = [P, P, P, P, I64, P, P, I64, C.c_int32, P, I64]
    out = np.zeros((256, 4), np.int64)
    g = O.or_c3_query(fk.ctypes.data, qty.ctypes.data, price.ctypes.data, flag.ctypes.data, len(fk), dkey.ctypes.data,
                      dattr.ctypes.data, len(dkey), 50, out.ctypes.data, 256)
    return out[:g]


def bench_join(args) -> int:
    """Star join (seed 3): fact 60M rows (fk int32 into the dimension keys, qty int64, price
    DECIMAL(15,2) as int64, flag < 25), dimension 1.5M rows (shuffled unique int32 key, attr
    < 1000). SELECT flag, COUNT(*), SUM(price), SUM(qty) FROM fact JOIN dim ON fk = key WHERE
    attr < 50 GROUP BY flag: dimension filter, exact key bitmap, fused fact filter, hash join
    (build on the dimension), gathers, dense GROUP BY. GB/s = 24 B of referenced fact columns per
    fact row / query time."""
    import warpstore_amd as ws
    from warpstore_amd import abi, queries as Q
    L = ws.lib()
    ws.check(L.ws_init())
    L.ws_jit_set_mode(abi.JIT_SYNC)
    tuned = {}
    for kv in args.tune:
        key, val = kv.split("=", 1)
        ws.check(L.ws_tune_set(key.encode(), int(val)))
        tuned[key] = int(val)
    if args.calibrate:
        src, dst = ws.DevBuf(1 << 30), ws.DevBuf(1 << 30)
        ws.check(L.ws_copy_kernel(dst.p, src.p, 1 << 30))
        L.ws_sync()
        src.free()
        dst.free()
    g = ws.datagen()
    dkey, dattr = np.empty(JOIN_DIM, np.int32), np.empty(JOIN_DIM, np.int32)
    g.wsg_c3_dim(3, JOIN_DIM, ws.ptr(dkey), ws.ptr(dattr))
    fk, flag = np.empty(JOIN_FACT, np.int32), np.empty(JOIN_FACT, np.int32)
    qty, price = np.empty(JOIN_FACT, np.int64), np.empty(JOIN_FACT, np.int64)
    g.wsg_c3_fact(3, JOIN_FACT, JOIN_DIM, ws.ptr(fk), ws.ptr(qty), ws.ptr(price), ws.ptr(flag))
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    for _ in range(max(args.warmup, 1)):
        rows, st = Q.c3(fact, dim)
    L.ws_sync()
    # stage times from one more (untimed) query: a warm-up's first launches also load code objects,
    # which would land between a stage's events
    L.ws_prof_reset()
    L.ws_prof_only(b"")
    L.ws_prof_enable(1)
    rows, st = Q.c3(fact, dim)
    L.ws_sync()
    warm = ws.prof_rows()
    # the fact filter's launch: the persistent (sets staged in LDS) or the chunked selection kernel
    fact_filter = next((k for k in warm if k.endswith("[filter (fact by join keys)]")), "select_jit[filter (fact by join keys)]")
    L.ws_prof_enable(0)
    L.ws_prof_reset()
    L.ws_prof_only(fact_filter.encode())
    L.ws_prof_enable(1)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rows, st = Q.c3(fact, dim)
        lat.append(time.perf_counter() - ts)
    L.ws_sync()
    elapsed = time.perf_counter() - t0
    prof = ws.prof_rows()
    L.ws_prof_enable(0)
    launches, total_ms = prof.get(fact_filter, (0, 0.0))
    kern_ms = total_ms / max(launches, 1)
    algo = 4 * JOIN_FACT  # the fused fact filter reads fk (4 B) per fact row
    achieved = algo / (kern_ms / 1e3) / 1e9 if kern_ms > 0 else 0.0
    O = oracle_lib()
    t1 = time.perf_counter()
    exp = oracle_c3(O, (fk, qty, price, flag), (dkey, dattr))
    cpu_s = time.perf_counter() - t1
    parity = bool(np.array_equal(np.asarray(rows)[np.argsort(np.asarray(rows)[:, 0])], exp[np.argsort(exp[:, 0])]))
    gbs = JOIN_FACT * JOIN_BYTES * args.steps / elapsed / 1e9
    out = {
        "metric": "GB/s scanned (% of HBM peak) for filter+agg/join; queries/s at p99, 1 GPU", "value": round(gbs, 3),
        "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic (seeded generator, seed 3)",
        "config": {"workload": "star join: fact 60M x dim 1.5M, 5 % dimension filter, GROUP BY flag",
                   "fact_rows": JOIN_FACT, "dim_rows": JOIN_DIM, "bytes_per_fact_row": JOIN_BYTES,
                   "dim_selected": int(st[0]), "fact_selected": int(st[1]), "pairs": int(st[2]),
                   **({"tuning": tuned} if tuned else {})},
        "pct_of_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 3),
        "queries_per_s": round(args.steps / elapsed, 3),
        "p50_ms": round(float(np.percentile(np.array(lat) * 1e3, 50)), 4),
        "p99_ms": round(float(np.percentile(np.array(lat) * 1e3, 99)), 4),
        "parity_vs_oracle": parity,
        "stage_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in sorted(warm.items(), key=lambda kv: -kv[1][1])[:8]},
        "roofline": {"kernel": fact_filter + " (compiled fact filter: fk - lo IN dimension-key bitmap)", "bound": "hbm",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": pmc_traffic("ws_q_select_staged (fact filter)", "pmc_c3_star_join.json"),
                     "algorithmic_bytes": algo,
                     "mean_launch_ms": round(kern_ms, 4), "launches": int(launches)},
        "cpu_baseline": {"value": round(JOIN_FACT * JOIN_BYTES / cpu_s / 1e9, 4), "unit": "GB/s", "cores": 1,
                         "kind": "port", "sample": f"the whole query, single-thread CPU oracle ({cpu_s:.2f} s)"},
    }
    fact.free()
    dim.free()
    print(json.dumps(out), flush=True)
    return 0 if parity else 1


def bench_serve(args) -> int:
    """The open-loop mixed stream (seed 5): Poisson arrivals, 70 % small aggregates on 64k-row
    segments, 25 % web-analytics queries on the 10M-row table, 5 % star joins (fact 60M, dim 1.5M).
    Per offered rate, continuous batching (one thread batches the small aggregates, 2 threads run
    the others concurrently) against per-query serving (4 concurrent queries, each alone); reports
    queries/s and p50 / p99 latency."""
    import warpstore_amd as ws
    from warpstore_amd import abi, queries as Q
    L = ws.lib()
    ws.check(L.ws_init())
    L.ws_jit_set_mode(abi.JIT_SYNC)
    pool = Q.C4Set(seed=5, nq=256, rows_override=65536)
    cols = gen_c2(ROWS)
    web = Q.DeviceTable(list(cols), [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32], [0, 1, 0, 0, 0])
    g = ws.datagen()
    nf, nd = 60_000_000, 1_500_000
    dkey, dattr = np.empty(nd, np.int32), np.empty(nd, np.int32)
    g.wsg_c3_dim(3, nd, ws.ptr(dkey), ws.ptr(dattr))
    fk, qty, price, flag = np.empty(nf, np.int32), np.empty(nf, np.int64), np.empty(nf, np.int64), np.empty(nf, np.int32)
    g.wsg_c3_fact(3, nf, nd, ws.ptr(fk), ws.ptr(qty), ws.ptr(price), ws.ptr(flag))
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    del fk, qty, price, flag
    Q.serve(pool, web, fact, dim, rate=200.0, count=60, mode=1, seed=1)  # warm-up: compile plans
    Q.serve(pool, web, fact, dim, rate=200.0, count=60, mode=0, seed=1)
    rows = []
    for rate in (400.0, 1600.0, 3200.0, 6400.0):
        count = int(min(4000, max(400, rate * 1.5)))
        for mode, name in ((1, "continuous_batching"), (0, "per_query")):
            st, _, _ = Q.serve(pool, web, fact, dim, rate=rate, count=count, mode=mode, workers=2 if mode else 4,
                               max_batch=256, seed=5)
            st.update({"offered_rate": rate, "mode": name})
            rows.append(st)
            print(json.dumps({k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}), flush=True)
    for t in (web, fact, dim):
        t.free()
    pool.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
