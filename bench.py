"""Benchmark of the analytic hot path on one MI355X.

Default workload (one step = one query over GPU-resident segments of 2^20 rows): the star join,
fact 60M rows (fk int32 into the dimension keys, qty int64 [1,50], price DECIMAL(15,2) as scaled
int64, flag [0,25)), dimension 1.5M rows (shuffled unique int32 key, attr int32 [0,1000)), seed 3:
  SELECT f.flag, COUNT(*), SUM(f.price), SUM(f.qty) FROM fact f JOIN dim d ON f.fk = d.key
  WHERE d.attr < 50 GROUP BY f.flag
run through the engine-side driver as the planner issues it, with results checked against the CPU
oracle. It is the largest single-GPU configuration of the metric (1.27 GB of referenced columns).

value = GB/s read: the bytes the query has to read per step (fk of every fact row, key and attr of
every dimension row, and qty, price and flag of the fact rows that survive the join filter) over
the wall time of the timed steps; inputs are resident in HBM before timing starts.
effective_gbs = 24 B of referenced fact columns x 60M rows over the same time (the figure a
column engine would quote for a full scan; most of qty / price / flag is never read).
roofline: the dominant kernel (most GPU time in a profiled untimed query), timed live with HIP
events on its launch stream during the timed steps; achieved = its algorithmic bytes per launch /
mean launch duration.
cpu_baseline: the multithreaded CPU restatement of the same query (oracle/cpu_baseline.cpp) on
this host's cores, plus its single-thread figure.

Other workloads (--workload): web (10M-row web-analytics GROUP BY user_id + COUNT DISTINCT + top
10), filter (1M-row filter + GROUP BY, rotating over 320 MB of table copies), batch (256 batched
GROUP BY queries), serve (the open-loop mixed stream).
"""
from __future__ import annotations

import argparse
import ctypes as C
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


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_lib():
    """The CPU baseline library (test infrastructure: timed beside the GPU path, never the product)."""
    B = C.CDLL(os.path.join(ROOT, "oracle", "libcpu_baseline.so"))
    P, I64, I = C.c_void_p, C.c_int64, C.c_int
    B.cb_hardware_threads.restype = I
    B.cb_c1.restype = I64
    B.cb_c1.argtypes = [P, P, P, I64, C.c_int32, I64, I, P, P, I64]
    B.cb_c2.restype = I64
    B.cb_c2.argtypes = [P, P, P, P, P, I64, C.c_int32, C.c_int32, I, P, I64]
    B.cb_c3.restype = I64
    B.cb_c3.argtypes = [P, P, P, P, I64, P, P, I64, C.c_int32, I, P, I64]
    return B


def time_cpu(run, target_s: float):
    """Runs run() repeatedly for about target_s seconds (at least once); returns seconds per run."""
    t0 = time.perf_counter()
    run()
    dt = time.perf_counter() - t0
    reps = int(max(1, min(20, target_s / max(dt, 1e-6) - 1)))
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    return (time.perf_counter() - t0) / reps


def cpu_baseline_entry(bytes_per_query: float, run_threads, target_s: float, sample: str) -> dict:
    """Times the CPU restatement with all hardware threads and with one thread."""
    B = cpu_lib()
    threads = int(B.cb_hardware_threads())
    mt = time_cpu(lambda: run_threads(threads), target_s * 0.6)
    st = time_cpu(lambda: run_threads(1), target_s * 0.4)
    return {"value": round(bytes_per_query / mt / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": sample, "ms_per_query": round(mt * 1e3, 3), "single_thread_value": round(bytes_per_query / st / 1e9, 4),
            "single_thread_ms_per_query": round(st * 1e3, 3), "cpu_model": cpu_model()}


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
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    O = C.CDLL(path)
    P, I64 = C.c_void_p, C.c_int64
    O.or_c2_query.restype = I64
    O.or_c2_query.argtypes = [P, P, P, P, P, I64, C.c_int32, C.c_int32, P, I64]
    O.or_c3_query.restype = I64
    O.or_c3_query.argtypes = [P, P, P, P, I64, P, P, I64, C.c_int32, P, I64]
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


def pmc_traffic(symbol: str, name: str):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter runs (or None)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            d = json.load(f)
        return d["kernels"][symbol]["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def setup(args):
    import warpstore_amd as ws
    from warpstore_amd import abi
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
    return ws, L, tuned


def timed_steps(ws, L, step, steps: int, prof_kernel: str):
    """Times exactly `steps` calls of step() between two device synchronizations, with the given
    kernel (profile name) bracketed by HIP events on its launch stream; returns
    (elapsed s, per-step latencies s, launches, kernel total ms, last result)."""
    L.ws_prof_reset()
    L.ws_prof_only(prof_kernel.encode())
    L.ws_prof_enable(1)
    L.ws_sync()
    lat = []
    out = None
    t0 = time.perf_counter()
    for s in range(steps):
        ts = time.perf_counter()
        out = step(s)
        lat.append(time.perf_counter() - ts)
    L.ws_sync()
    elapsed = time.perf_counter() - t0
    prof = ws.prof_rows()
    L.ws_prof_enable(0)
    launches, total_ms = prof.get(prof_kernel, (0, 0.0))
    return elapsed, lat, launches, total_ms, out


def profiled_query(ws, L, step):
    """One untimed query with every kernel bracketed by events: {profile name: (launches, ms)}."""
    L.ws_prof_reset()
    L.ws_prof_only(b"")
    L.ws_prof_enable(1)
    out = step(0)
    L.ws_sync()
    warm = ws.prof_rows()
    L.ws_prof_enable(0)
    return warm, out


def base_line(args, value, elapsed, lat, data, config, dtype="int64"):
    return {
        "metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": data, "config": config,
        "pct_of_hbm_peak": round(100 * value / HBM_PEAK_GBS, 3),
        "queries_per_s": round(args.steps / elapsed, 3),
        "p50_ms": round(float(np.percentile(np.array(lat) * 1e3, 50)), 4),
        "p99_ms": round(float(np.percentile(np.array(lat) * 1e3, 99)), 4),
    }


def roofline(kernel_desc, algo_bytes, launches, total_ms, traffic):
    kern_ms = total_ms / max(launches, 1)
    achieved = algo_bytes / (kern_ms / 1e3) / 1e9 if kern_ms > 0 else 0.0
    return {"kernel": kernel_desc, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "algorithmic_bytes": int(algo_bytes),
            "mean_launch_ms": round(kern_ms, 4), "launches": int(launches)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--copies", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a tuning knob of the physical plans (ws_tune_set) before running")
    ap.add_argument("--no-autotune", action="store_true", help="batch workload: skip the autotuner, use the knobs as set")
    ap.add_argument("--no-learned", action="store_true", help="join / web workloads: skip the learned-sizing comparison")
    ap.add_argument("--no-concurrent", action="store_true",
                    help="join workload: skip the concurrent-sessions figure (profiles of the single-session query)")
    ap.add_argument("--heavy-workers", type=int, default=3,
                    help="serve: threads running web / join queries beside the batching thread (continuous-batching form)")
    ap.add_argument("--budget-mb", type=int, default=1536,
                    help="serve workload: HBM budget for the requests' estimated temporaries (0: none)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 1 GiB device copy kernel (known bytes, for counter calibration)")
    ap.add_argument("--workload", choices=["join", "web", "filter", "batch", "serve"], default="join",
                    help="join: the star join, fact 60M x dimension 1.5M (default); web: the 10M-row "
                         "web-analytics query; filter: the 1M-row filter + GROUP BY; batch: 256 small GROUP BY "
                         "queries, one batched launch vs per-query; serve: the open-loop mixed stream")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path's multi-GPU sharding is a design note, no collective runs here)", file=sys.stderr)
        return 2
    return {"join": bench_join, "web": bench_web, "filter": bench_filter, "batch": bench_batch,
            "serve": bench_serve}[args.workload](args)


# Scan kernels of the web query's phase-1 aggregation that can dominate, with their algorithmic
# bytes per launch and the kernel symbol they launch.
PHASE1 = "[aggregate (distinct pairs)]"  # the plan-cache kind of the phase-1 plan: its launches' profile tag
WEB_ALGO = {
    "agg_hash_jit" + PHASE1: (22 * ROWS, "ws_q_grouped", "compiled filter + shared-table hash GROUP BY (user_id, url)"),
    "agg_part_scatter_jit" + PHASE1: (22 * ROWS, "ws_q_grouped_pscatter",
                                      "compiled filter + partition scatter of (user_id, url) partial groups"),
    "agg_part_count_jit" + PHASE1: (18 * ROWS, "ws_q_grouped_pcount", "compiled filter + partition count"),
}


def bench_web(args) -> int:
    """The web-analytics query (seed 2): 10M rows, user_id int64 Zipf(1.1) over 1e6 ids, region in
    [0,256) stored frame-of-reference encoded (uint16 codes), event_date over 365 days, url code
    Zipf(1.1) over 1e5, dur lognormal; SELECT user_id, COUNT(*) c, SUM(dur), COUNT(DISTINCT url) FROM
    h WHERE event_date < 19723+300 AND region < 224 GROUP BY user_id ORDER BY c DESC LIMIT 10.
    value = 22 B of referenced columns per row x 10M rows per query / wall time; steps rotate over
    copies of the table that together exceed the 256 MB Infinity Cache."""
    from warpstore_amd import abi, queries as Q
    ws, L, tuned = setup(args)
    cols = gen_c2(ROWS)
    types = [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32]
    tables = [Q.DeviceTable(list(cols), types, [0, 1, 0, 0, 0]) for _ in range(args.copies)]
    for w in range(max(args.warmup, 1)):
        Q.c2(tables[w % args.copies], 10)
    L.ws_sync()
    warm, _ = profiled_query(ws, L, lambda s: Q.c2(tables[0], 10))
    dom = max((n for n in WEB_ALGO if n in warm), key=lambda n: warm[n][1], default="agg_hash_jit" + PHASE1)
    algo_bytes, symbol, what = WEB_ALGO[dom]
    elapsed, lat, launches, total_ms, (top, stats) = timed_steps(
        ws, L, lambda s: Q.c2(tables[s % args.copies], 10), args.steps, dom)
    gbs = ROWS * BYTES_PER_ROW * args.steps / elapsed / 1e9
    full = oracle_c2(oracle_lib(), cols, ROWS)
    parity = bool(check_top(top, full))
    result = base_line(args, gbs, elapsed, lat, "synthetic (seeded generator, seed 2)", {
        "workload": "web-analytics GROUP BY user_id + COUNT DISTINCT url + top 10, 10M rows",
        "rows": ROWS, "bytes_per_row": BYTES_PER_ROW, "segments_rows": 1 << 20, "table_copies": args.copies,
        "distinct_pairs": int(stats[0]), "groups": int(stats[1]), "phase1_capacity": int(stats[2]),
        "overflow_retries": int(stats[3]), **({"tuning": tuned} if tuned else {})})
    result["rows_per_s"] = round(ROWS * args.steps / elapsed, 1)
    result["parity_vs_oracle"] = parity
    result["roofline"] = roofline(f"{symbol} ({what})", algo_bytes, launches, total_ms,
                                  pmc_traffic(symbol, "pmc_c2_web_analytics.json"))
    if not args.no_learned:
        result["learned_sizing"] = web_learned(ws, L, Q, tables, full, args, elapsed / args.steps)
    if not args.no_cpu:
        B = cpu_lib()
        cap = ROWS + 1
        out = np.zeros((cap, 4), np.int64)
        ptrs = [c.ctypes.data for c in cols]
        result["cpu_baseline"] = cpu_baseline_entry(
            ROWS * BYTES_PER_ROW, lambda t: B.cb_c2(*ptrs, ROWS, 19723 + 300, 224, t, out.ctypes.data, cap), args.cpu_seconds,
            "the whole query over the same 10M rows (filter, partition by user, sort, group, order)")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0 if parity else 1


FILTER_ROWS = 1_000_000
FILTER_BYTES = 20  # k 4 + v 8 + p 8


def bench_filter(args) -> int:
    """The plain filter + GROUP BY (seed 1): 1M rows, k int32 [0,1000), v int64 [0,1e6), p double
    [0,1000); SELECT k, COUNT(*), SUM(v), MIN(v), MAX(v), AVG(p) FROM t WHERE k < 500 AND v > 1000
    GROUP BY k. One 20 MB table fits the 256 MB Infinity Cache, so steps rotate over 16 copies
    (320 MB) and every step reads a copy the previous 15 steps did not touch."""
    from warpstore_amd import abi, queries as Q
    ws, L, tuned = setup(args)
    g = ws.datagen()
    k, v, p = np.empty(FILTER_ROWS, np.int32), np.empty(FILTER_ROWS, np.int64), np.empty(FILTER_ROWS, np.float64)
    g.wsg_c1(1, FILTER_ROWS, ws.ptr(k), ws.ptr(v), ws.ptr(p))
    ncopies = 16
    tables = [Q.DeviceTable([k, v, p], [abi.T_I32, abi.T_I64, abi.T_F64]) for _ in range(ncopies)]
    for w in range(max(args.warmup, 1)):
        Q.c1(tables[w % ncopies], True)
    L.ws_sync()
    warm, _ = profiled_query(ws, L, lambda s: Q.c1(tables[0], True))
    scan = [n for n in warm if n.startswith("agg_dense") or n.startswith("agg_hash") or n.startswith("agg_nokeys")]
    dom = max(scan, key=lambda n: warm[n][1]) if scan else max(warm, key=lambda n: warm[n][1])
    elapsed, lat, launches, total_ms, (rows, avg) = timed_steps(
        ws, L, lambda s: Q.c1(tables[(s + args.warmup) % ncopies], True), args.steps, dom)
    gbs = FILTER_ROWS * FILTER_BYTES * args.steps / elapsed / 1e9
    # parity against the CPU restatement
    B = cpu_lib()
    out = np.zeros((1000, 5), np.int64)
    psum = np.zeros(1000)
    ng = B.cb_c1(k.ctypes.data, v.ctypes.data, p.ctypes.data, FILTER_ROWS, 500, 1000, 1, out.ctypes.data, psum.ctypes.data, 1000)
    order = np.argsort(rows[:, 0])
    parity = bool(len(rows) == ng and np.array_equal(rows[order], out[:ng]) and
                  np.allclose(avg[order], psum[:ng] / out[:ng, 1], rtol=1e-12, atol=0))
    result = base_line(args, gbs, elapsed, lat, "synthetic (seeded generator, seed 1)", {
        "workload": "filter + GROUP BY k, 1M rows (k < 500 AND v > 1000; COUNT, SUM, MIN, MAX v, AVG p)",
        "rows": FILTER_ROWS, "bytes_per_row": FILTER_BYTES, "table_copies": ncopies, "groups": int(len(rows)),
        **({"tuning": tuned} if tuned else {})})
    result["rows_per_s"] = round(FILTER_ROWS * args.steps / elapsed, 1)
    result["parity_vs_oracle"] = parity
    result["roofline"] = roofline(dom, FILTER_ROWS * FILTER_BYTES, launches, total_ms, None)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline_entry(
            FILTER_ROWS * FILTER_BYTES,
            lambda t: B.cb_c1(k.ctypes.data, v.ctypes.data, p.ctypes.data, FILTER_ROWS, 500, 1000, t, out.ctypes.data,
                              psum.ctypes.data, 1000), args.cpu_seconds, "the whole query over the same 1M rows")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0 if parity else 1


def bench_batch(args) -> int:
    """256 filter + GROUP BY queries (seed 4): rows lognormal in [4096, 4M], keys with 10 to 1e5
    distinct values, WHERE val > theta_q; 12 B/row scanned. One step = all 256 queries, their plans
    built, run as one batched launch and every group fetched into page-locked memory; `value` times
    steps pipelined three deep (a step's build and result copy overlap later steps' kernels), and
    the same steps one at a time are reported beside it. Also timed: the queries one by one
    through the per-query path, and as a batch prepared once whose launches replay from a HIP graph
    (the repeated-dashboard form), one replay at a time and pipelined over three prepared copies
    (each batch's groups copied to the host while the next two batches' kernels run), and pipelined
    three deep with every batch built from scratch (plans, aggregators, descriptors and graph). Every timed form fetches every group to the host; the digests
    that compare the forms are computed after the timing. Unless --no-autotune, the autotuner first
    picks the batched form's tile, LDS table and grid knobs on this workload."""
    import warpstore_amd as ws
    from warpstore_amd import abi, queries as Q
    L = ws.lib()
    ws.check(L.ws_init())
    L.ws_jit_set_mode(abi.JIT_SYNC)
    for kv in args.tune:
        key, val = kv.split("=", 1)
        ws.check(L.ws_tune_set(key.encode(), int(val)))
    s = Q.C4Set(seed=4, nq=256)
    tuned = None if args.no_autotune else s.autotune(reps=3)
    res = {}
    for mode in ("batched", "per_query"):
        batched = mode == "batched"
        for _ in range(max(args.warmup, 1)):
            s.run(batched=batched, digests=False)
        L.ws_sync()
        if batched:
            elapsed, lat, launches, total_ms, _ = timed_steps(ws, L, lambda i: s.run(batched=True, digests=False), args.steps,
                                                              "agg_batch_dense_jit")
        else:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                s.run(batched=False, digests=False)
            L.ws_sync()
            elapsed = time.perf_counter() - t0
        res[mode] = elapsed / args.steps
    # the same batch prepared once (its launches recorded in a HIP graph) and replayed
    pr = s.prepare()
    for _ in range(max(args.warmup, 1)):
        pr.run(digests=False)
    L.ws_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pr.run(digests=False)
    L.ws_sync()
    res["replay"] = (time.perf_counter() - t0) / args.steps
    dr, gr = pr.run()
    # three prepared copies in a pipeline: each batch's groups are copied to the host while the
    # next two batches' kernels run (every batch still fetches all its groups)
    more = [s.prepare(), s.prepare()]
    pr.pipeline(more, max(args.warmup, 3), digests=False)
    L.ws_sync()
    t0 = time.perf_counter()
    pr.pipeline(more, args.steps, digests=False)
    L.ws_sync()
    res["pipelined"] = (time.perf_counter() - t0) / args.steps
    dl, gl = pr.pipeline(more, 4)
    for p in more:
        p.free()
    pr.free()
    # every batch built from scratch (plans, aggregators, descriptors, graph), pipelined 3 deep
    s.fresh_pipeline(max(args.warmup, 3), 3, digests=False)
    L.ws_sync()
    t0 = time.perf_counter()
    s.fresh_pipeline(args.steps, 3, digests=False)
    L.ws_sync()
    res["fresh_pipelined"] = (time.perf_counter() - t0) / args.steps
    df, gf = s.fresh_pipeline(4, 3)
    db, gb = s.run(batched=True)
    dp, gp = s.run(batched=False)
    same = bool(np.array_equal(db, dp) and np.array_equal(gb, gp) and np.array_equal(dr, dp) and np.array_equal(gr, gp)
                and np.array_equal(dl, dp) and np.array_equal(gl, gp) and np.array_equal(df, dp) and np.array_equal(gf, gp))
    bt, pt, ft = res["batched"], res["per_query"], res["fresh_pipelined"]
    scan = s.rows * 12
    # value: the whole job's throughput, batches built from scratch and pipelined (the same work per
    # step as one batch at a time, steps overlapped); the roofline is the batched kernel measured
    # in the one-at-a-time steps, where no other batch's kernel shares the GPU
    out = {
        "metric": "GB/s scanned by 256 batched filter+GROUP BY queries", "value": round(scan / ft / 1e9, 3),
        "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ft * 1e3, 4),
        "higher_is_better": True, "dtype": "int64", "data": "synthetic (seeded generator, seed 4)",
        "config": {"workload": "256 filter + GROUP BY queries per step, each step's plans built from scratch, one batched "
                               "launch per step, steps pipelined 3 deep", "rows": s.rows, "bytes_per_row": 12},
        "queries_per_s": round(256 / ft, 1), "per_query_ms_per_step": round(pt * 1e3, 4),
        "per_query_queries_per_s": round(256 / pt, 1), "speedup_vs_per_query": round(pt / ft, 3),
        "one_at_a_time_ms_per_step": round(bt * 1e3, 4), "one_at_a_time_value": round(scan / bt / 1e9, 3),
        "results_equal_per_query": same, "groups": int(gb.sum()), "autotuner": tuned,
        "graph_replay_ms_per_step": round(res["replay"] * 1e3, 4), "graph_replay_value": round(scan / res["replay"] / 1e9, 3),
        "pipelined_ms_per_step": round(res["pipelined"] * 1e3, 4), "pipelined_value": round(scan / res["pipelined"] / 1e9, 3),
        "fresh_pipelined_ms_per_step": round(res["fresh_pipelined"] * 1e3, 4),
        "fresh_pipelined_value": round(scan / res["fresh_pipelined"] / 1e9, 3),
        "roofline": roofline("agg_batch_dense_jit (all 256 queries' filter + dense GROUP BY in one launch)", scan, launches,
                             total_ms, None),
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


def c3_data(ws):
    g = ws.datagen()
    dkey, dattr = np.empty(JOIN_DIM, np.int32), np.empty(JOIN_DIM, np.int32)
    g.wsg_c3_dim(3, JOIN_DIM, ws.ptr(dkey), ws.ptr(dattr))
    fk, flag = np.empty(JOIN_FACT, np.int32), np.empty(JOIN_FACT, np.int32)
    qty, price = np.empty(JOIN_FACT, np.int64), np.empty(JOIN_FACT, np.int64)
    g.wsg_c3_fact(3, JOIN_FACT, JOIN_DIM, ws.ptr(fk), ws.ptr(qty), ws.ptr(price), ws.ptr(flag))
    return (fk, qty, price, flag), (dkey, dattr)


def c3_bytes_read(fact_selected: int) -> int:
    """Bytes the star join has to read: fk of every fact row, key and attr of every dimension row,
    and qty (8 B), price (8 B) and flag (4 B) of each fact row that survives the join filter."""
    return 4 * JOIN_FACT + 8 * JOIN_DIM + 20 * int(fact_selected)


def bench_join(args) -> int:
    """Star join (seed 3): fact 60M rows (fk int32 into the dimension keys, qty int64, price
    DECIMAL(15,2) as int64, flag < 25), dimension 1.5M rows (shuffled unique int32 key, attr
    < 1000). SELECT flag, COUNT(*), SUM(price), SUM(qty) FROM fact JOIN dim ON fk = key WHERE
    attr < 50 GROUP BY flag: dimension filter, exact key bitmap, fused fact filter, hash join
    (build on the dimension), gathers, dense GROUP BY."""
    from warpstore_amd import abi, queries as Q
    ws, L, tuned = setup(args)
    (fk, qty, price, flag), (dkey, dattr) = c3_data(ws)
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    for _ in range(max(args.warmup, 1)):
        Q.c3(fact, dim)
    L.ws_sync()
    # stage times from one more (untimed) query: a warm-up's first launches also load code objects
    warm, (_, wst) = profiled_query(ws, L, lambda s: Q.c3(fact, dim))
    # the roofline kernel: the one with the most GPU time in that query (the fused fact scan when
    # the dimension keys are distinct, else the fact filter of the hash-join plan)
    dom = max(warm, key=lambda k: warm[k][1])
    elapsed, lat, launches, total_ms, (rows, st) = timed_steps(ws, L, lambda s: Q.c3(fact, dim), args.steps, dom)
    O = oracle_lib()
    exp = oracle_c3(O, (fk, qty, price, flag), (dkey, dattr))
    rows = np.asarray(rows)
    parity = bool(np.array_equal(rows[np.argsort(rows[:, 0])], exp[np.argsort(exp[:, 0])]))
    bytes_read = c3_bytes_read(int(st[1]))
    gbs = bytes_read * args.steps / elapsed / 1e9
    out = base_line(args, gbs, elapsed, lat, "synthetic (seeded generator, seed 3)", {
        "workload": "star join: fact 60M x dim 1.5M, 5 % dimension filter, GROUP BY flag",
        "fact_rows": JOIN_FACT, "dim_rows": JOIN_DIM, "bytes_read_per_query": bytes_read,
        "dim_selected": int(st[0]), "fact_selected": int(st[1]), "pairs": int(st[2]),
        **({"tuning": tuned} if tuned else {})})
    out["effective_gbs"] = round(JOIN_FACT * JOIN_BYTES * args.steps / elapsed / 1e9, 3)
    out["fact_rows_per_s"] = round(JOIN_FACT * args.steps / elapsed, 1)
    out["parity_vs_oracle"] = parity
    out["stage_ms"] = {k: round(v[1] / max(v[0], 1), 4) for k, v in sorted(warm.items(), key=lambda kv: -kv[1][1])[:8]}
    if dom.startswith("agg_dense_staged"):
        # fk of every fact row, flag (4 B), price and qty (8 B each) of every passing row
        desc = " (fused fact scan: fk - lo IN dimension-key bitmap staged in LDS, GROUP BY flag of passing rows)"
        alg, sym = 4 * JOIN_FACT + 20 * int(st[1]), "ws_q_dense_staged (fused fact scan)"
    else:
        desc, alg, sym = " (compiled fact filter: fk - lo IN dimension-key bitmap)", 4 * JOIN_FACT, "ws_q_select_staged (fact filter)"
    out["roofline"] = roofline(dom + desc, alg, launches, total_ms, pmc_traffic(sym, "pmc_c3_star_join.json"))
    out["plan"] = "join as bitmap filter fused into the fact aggregation (distinct dimension keys)" if dom.startswith(
        "agg_dense_staged") else "hash join (filtered fact rows probe the filtered dimension)"
    out["concurrent_sessions"] = None if args.no_concurrent else join_concurrent(L, Q, fact, dim, rows, bytes_read, args)
    if not args.no_learned:
        out["learned_sizing"] = join_learned(ws, L, Q, fact, dim, exp, args)
    if not args.no_cpu:
        B = cpu_lib()
        res = np.zeros((256, 4), np.int64)
        ptrs = [a.ctypes.data for a in (fk, qty, price, flag)]
        out["cpu_baseline"] = cpu_baseline_entry(
            bytes_read, lambda t: B.cb_c3(*ptrs, JOIN_FACT, dkey.ctypes.data, dattr.ctypes.data, JOIN_DIM, 50, t,
                                          res.ctypes.data, 256), args.cpu_seconds,
            "the whole query over the same 60M x 1.5M tables (filtered build with multiplicities, probe, GROUP BY flag)")
    else:
        out["cpu_baseline"] = None
    fact.free()
    dim.free()
    print(json.dumps(out), flush=True)
    return 0 if parity else 1


def join_concurrent(L, Q, fact, dim, want, bytes_read, args, sessions: int = 4) -> dict:
    """The star join from `sessions` host threads at once, as the reference engine runs at most
    four queries concurrently (engine.cpp:1262-1337): each thread issues its launches on its own
    stream, so one query's host-side launch work overlaps another's fact scan. Whole-job figures
    over sessions x steps queries (timed from the release of the warmed-up threads to the end of
    the last query); every thread's last result is checked against the single-session one."""
    import threading
    steps = max(args.steps, 1)
    start, done = threading.Barrier(sessions + 1), threading.Barrier(sessions + 1)
    got, errors = [None] * sessions, []

    def worker(i):
        try:
            for _ in range(max(args.warmup, 1)):
                Q.c3(fact, dim)
            L.ws_sync()
            start.wait()
            for _ in range(steps):
                r, _ = Q.c3(fact, dim)
            got[i] = np.asarray(r)
            done.wait()
        except BaseException as e:  # noqa: BLE001 - reported below, the barriers released
            errors.append(repr(e))
            start.abort()
            done.abort()

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(sessions)]
    for t in threads:
        t.start()
    try:
        start.wait()
        t0 = time.perf_counter()
        done.wait()
        elapsed = time.perf_counter() - t0
    except threading.BrokenBarrierError:
        elapsed = None
    for t in threads:
        t.join()
    if errors or elapsed is None:
        return {"sessions": sessions, "error": errors[:1]}
    key = np.argsort(want[:, 0])
    same = all(g is not None and np.array_equal(g[np.argsort(g[:, 0])], want[key]) for g in got)
    q = sessions * steps
    return {"sessions": sessions, "queries": q, "queries_per_s": round(q / elapsed, 1),
            "value": round(bytes_read * q / elapsed / 1e9, 3), "unit": "GB/s", "ms_per_query": round(elapsed / q * 1e3, 4),
            "results_equal_single_session": bool(same)}


JOIN_TRAIN_FILTERS = (3, 8, 15, 25, 35, 65, 90, 130, 200, 300, 450, 700)  # dimension filters attr < x (not 50)


def join_learned(ws, L, Q, fact, dim, exp, args):
    """The hash-join plan of the star join (the join kept although the dimension keys are
    distinct), plain and with the learned join-size model: trained on the GPU on the exact output
    sizes of the same join under 12 other dimension filters (the measured filter attr < 50 is held
    out), installed in the drivers, then timed like the plain path; results checked again."""
    from warpstore_amd import sizing

    def timed():
        for _ in range(max(args.warmup, 1)):
            Q.c3(fact, dim)
        L.ws_sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = Q.c3(fact, dim)
        L.ws_sync()
        return (time.perf_counter() - t0) / args.steps, out

    Q.set_join_rewrite(False)
    try:
        plain_s, _ = timed()
        sizing.log_clear("joins")
        for a in JOIN_TRAIN_FILTERS:
            Q.c3(fact, dim, attr_lt=a)
        model = sizing.train_from_log("joins", epochs=600)
        Q.set_sizers(joins=model.h)
        try:
            learned_s, (rows, st) = timed()
        finally:
            Q.set_sizers(joins=None)
    finally:
        Q.set_join_rewrite(True)
    rows = np.asarray(rows)
    same = bool(np.array_equal(rows[np.argsort(rows[:, 0])], exp[np.argsort(exp[:, 0])]))
    return {"plan": "hash join (the default plan runs the join as the fused bitmap filter)",
            "model": "MLP 16-32-32-1 on join key statistics, trained on the GPU", "training_joins": model.samples,
            "train_loss": round(model.loss, 6), "pairs": int(st[2]), "predicted_pairs": int(st[3]),
            "plain_ms_per_query": round(plain_s * 1e3, 4), "learned_ms_per_query": round(learned_s * 1e3, 4),
            "speedup": round(plain_s / learned_s, 3), "parity_vs_oracle": same}


WEB_TRAIN_FILTERS = [(19723 + d, r) for d in (30, 90, 150, 220, 365) for r in (16, 64, 128, 192, 256)]


def web_learned(ws, L, Q, tables, full, args, plain_s):
    """The web query with the learned group-count model: trained on the GPU on the exact group
    counts both phases log under 25 other filters (the measured filter is held out), installed,
    then timed over the same table copies; results checked again."""
    from warpstore_amd import sizing
    sizing.log_clear("groups")
    for d, r in WEB_TRAIN_FILTERS:
        Q.c2(tables[0], 10, date_lt=d, region_lt=r)
    model = sizing.train_from_log("groups", epochs=600)
    Q.set_sizers(groups=model.h)
    try:
        for w in range(max(args.warmup, 1)):
            Q.c2(tables[w % len(tables)], 10)
        L.ws_sync()
        t0 = time.perf_counter()
        for s in range(args.steps):
            top, st = Q.c2(tables[s % len(tables)], 10)
        L.ws_sync()
        learned_s = (time.perf_counter() - t0) / args.steps
    finally:
        Q.set_sizers(groups=None)
    return {"model": "MLP 16-32-32-1 on segment statistics, trained on the GPU", "training_aggregations": model.samples,
            "train_loss": round(model.loss, 6), "phase1_capacity": int(st[2]), "overflow_retries": int(st[3]),
            "plain_ms_per_query": round(plain_s * 1e3, 4), "learned_ms_per_query": round(learned_s * 1e3, 4),
            "speedup": round(plain_s / learned_s, 3), "parity_vs_oracle": bool(check_top(top, full))}


def bench_serve(args) -> int:
    """The open-loop mixed stream (seed 5): Poisson arrivals, 70 % small aggregates on 64k-row
    segments, 25 % web-analytics queries on the 10M-row table, 5 % star joins (fact 60M, dim 1.5M).
    Both forms admit requests under the same HBM budget for their estimated temporaries
    (--budget-mb). Per offered rate: the serving form (continuous batching: one thread batches the
    small aggregates, 2 threads run the others concurrently; learned group-count and join-size
    models installed, so estimates are predictions) against the plain per-query form (4 concurrent
    queries, each alone; the range rules' upper-bound estimates). The models are trained first on
    the GPU from the exact counts of the same queries under other filters. Reports queries/s with
    p50 / p99 latency, one line per rate and form, then a summary."""
    import warpstore_amd as ws
    from warpstore_amd import abi, queries as Q, sizing
    L = ws.lib()
    ws.check(L.ws_init())
    L.ws_jit_set_mode(abi.JIT_SYNC)
    pool = Q.C4Set(seed=5, nq=256, rows_override=65536)
    cols = gen_c2(ROWS)
    web = Q.DeviceTable(list(cols), [abi.T_I64, abi.T_I32, abi.T_I32, abi.T_I32, abi.T_I32], [0, 1, 0, 0, 0])
    (fk, qty, price, flag), (dkey, dattr) = c3_data(ws)
    fact = Q.DeviceTable([fk, qty, price, flag], [abi.T_I32, abi.T_I64, abi.T_DECIMAL, abi.T_I32])
    dim = Q.DeviceTable([dkey, dattr], [abi.T_I32, abi.T_I32])
    del fk, qty, price, flag
    sizing.log_clear("groups")
    sizing.log_clear("joins")
    for d, r in WEB_TRAIN_FILTERS:
        Q.c2(web, 10, date_lt=d, region_lt=r)
    Q.set_join_rewrite(False)  # the join sizes come from the hash-join plan
    try:
        for a in JOIN_TRAIN_FILTERS:
            Q.c3(fact, dim, attr_lt=a)
    finally:
        Q.set_join_rewrite(True)
    gm, jm = sizing.train_from_log("groups", epochs=600), sizing.train_from_log("joins", epochs=600)
    budget = args.budget_mb << 20
    # the batched form: one batching thread + heavy_workers heavy-query threads; the per-query
    # form: the reference's at most 4 queries at a time
    forms = (("learned_batched", 1, args.heavy_workers, (gm.h, jm.h)), ("plain_per_query", 0, 4, (None, None)))
    for name, mode, workers, models in forms:  # warm-up: compile plans
        Q.set_sizers(*models)
        Q.serve(pool, web, fact, dim, rate=200.0, count=60, mode=mode, workers=workers, seed=1, budget_bytes=budget)
    rows = []
    # (12800/s saturates both forms: at 6400 the batched form drains only just behind the arrivals,
    # so its q/s there is bounded by the offered rate, not by its capacity)
    for rate in (400.0, 1600.0, 3200.0, 6400.0, 12800.0):
        count = int(min(4000, max(400, rate * 1.5)))
        for name, mode, workers, models in forms:
            Q.set_sizers(*models)
            st, _, _ = Q.serve(pool, web, fact, dim, rate=rate, count=count, mode=mode, workers=workers, max_batch=256, seed=5,
                               budget_bytes=budget)
            st.update({"offered_rate": rate, "form": name, "budget_mb": args.budget_mb,
                       "gpu_threads": workers + 1 if mode == 1 else workers})
            rows.append(st)
            print(json.dumps({k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}), flush=True)
    Q.set_sizers(None, None)
    # capacity at equal p99: the highest sustained q/s of each form among the rates where its p99
    # stays within the plain form's p99 at its own highest sustained rate
    summary = {}
    plain = [r for r in rows if r["form"] == "plain_per_query"]
    p99_cap = max(plain, key=lambda r: r["qps"])["p99_ms"]
    for name in ("learned_batched", "plain_per_query"):
        rs = [r for r in rows if r["form"] == name]
        ok = [r for r in rs if r["p99_ms"] <= p99_cap] or rs
        best = max(ok, key=lambda r: r["qps"])
        summary[name] = {"max_qps": round(best["qps"], 1), "p99_ms_at_max": round(best["p99_ms"], 3),
                         "offered_rate_at_max": best["offered_rate"]}
    summary["p99_bound_ms"] = round(p99_cap, 3)
    summary["qps_ratio"] = round(summary["learned_batched"]["max_qps"] / summary["plain_per_query"]["max_qps"], 3)
    print(json.dumps({"serving_summary": summary, "models": {"groups_trained_on": gm.samples, "joins_trained_on": jm.samples}}),
          flush=True)
    for t in (web, fact, dim):
        t.free()
    pool.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
