"""ctypes mirror of libws_queries.so: the engine-side drivers of the measured query shapes,
which call the operator library only through the drop-in gdb:: surface."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import PKG, check, lib
from .abi import T_F64, T_I32, T_I64, T_DECIMAL  # noqa: F401

_q = None


def qlib():
    global _q
    if _q is None:
        lib()  # load the operator library first (the drivers link against it)
        path = os.path.join(PKG, "libws_queries.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with warpstore_amd/build.py")
        Q = C.CDLL(path)
        P, I, I64 = C.c_void_p, C.c_int, C.c_int64
        Q.wsq_last_error.restype = C.c_char_p
        Q.wsq_table_load.argtypes = [I, P, P, I64, P, C.POINTER(C.c_void_p)]
        Q.wsq_table_load.restype = I
        Q.wsq_table_free.argtypes = [P]
        Q.wsq_c1.argtypes = [P, I, P, P, I64, C.POINTER(I64)]
        Q.wsq_c1.restype = I
        Q.wsq_c2.argtypes = [P, I64, P, C.POINTER(I64), P]
        Q.wsq_c2.restype = I
        Q.wsq_c2_where.argtypes = [P, I64, I64, I64, P, C.POINTER(I64), P]
        Q.wsq_c2_where.restype = I
        Q.wsq_c3.argtypes = [P, P, I64, P, C.POINTER(I64), P]
        Q.wsq_c3.restype = I
        Q.wsq_set_join_sizer.argtypes = [P]
        Q.wsq_set_join_sizer.restype = None
        Q.wsq_set_join_rewrite.argtypes = [I]
        Q.wsq_set_join_rewrite.restype = None
        Q.wsq_c4_load.argtypes = [C.c_uint64, I64, I64, C.POINTER(C.c_void_p), C.POINTER(I64)]
        Q.wsq_c4_load.restype = I
        Q.wsq_c4_free.argtypes = [P]
        Q.wsq_c4_run.argtypes = [P, I, P, P]
        Q.wsq_c4_run.restype = I
        Q.wsq_serve.argtypes = [P, P, P, P, C.c_double, I64, I, I, I, C.c_uint64, I64, P, P, P]
        Q.wsq_serve.restype = I
        Q.wsq_serve_alone.argtypes = [P, P, P, P, C.c_double, I64, C.c_uint64, I64, C.POINTER(C.c_uint64)]
        Q.wsq_serve_alone.restype = I
        Q.wsq_set_sizer.argtypes = [P]
        Q.wsq_set_sizer.restype = None
        Q.wsq_c4_prepare.argtypes = [P, C.POINTER(C.c_void_p)]
        Q.wsq_c4_prepare.restype = I
        Q.wsq_c4_replay.argtypes = [P, P, P]
        Q.wsq_c4_replay.restype = I
        Q.wsq_c4_pipeline.argtypes = [P, I, I, P, P]
        Q.wsq_c4_pipeline.restype = I
        Q.wsq_c4_fresh_pipeline.argtypes = [P, I, I, P, P]
        Q.wsq_c4_fresh_pipeline.restype = I
        Q.wsq_c4_plan_handle.argtypes = [P]
        Q.wsq_c4_plan_handle.restype = C.c_void_p
        Q.wsq_c4_release.argtypes = [P]
        Q.wsq_c4_release.restype = None
        for f in ("wsq_c4_autotune", "wsq_c2_autotune"):
            getattr(Q, f).argtypes = [P, I, P, P, C.POINTER(I)]
            getattr(Q, f).restype = I
        _q = Q
    return _q


def _ck(rc: int) -> None:
    if rc != 0:
        from . import WsError
        raise WsError(qlib().wsq_last_error().decode(errors="replace"), rc)


class DeviceTable:
    """A GPU-resident table cut into 2^20-row segments (optionally encoded columns)."""

    def __init__(self, arrays: list, types: list, encode: list | None = None) -> None:
        arrays = [np.ascontiguousarray(a) for a in arrays]
        n = len(arrays[0])
        t = (C.c_int * len(types))(*types)
        e = (C.c_int * len(types))(*(encode or [0] * len(types)))
        ptrs = (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
        h = C.c_void_p()
        _ck(qlib().wsq_table_load(len(types), t, e, n, ptrs, C.byref(h)))
        self.h = h
        self.n = n

    def free(self) -> None:
        if self.h:
            qlib().wsq_table_free(self.h)
            self.h = None

    def __del__(self) -> None:
        try:
            self.free()
        except Exception:
            pass


def c1(t: DeviceTable, dense: bool = True):
    """Rows (k, count, sum(v), min(v), max(v)) int64 and avg(p) double, in group order."""
    oi = np.zeros((1000, 5), np.int64)
    od = np.zeros(1000)
    g = C.c_int64()
    _ck(qlib().wsq_c1(t.h, 1 if dense else 0, oi.ctypes.data, od.ctypes.data, 1000, C.byref(g)))
    return oi[:g.value], od[:g.value]


def c2(t: DeviceTable, limit: int = 10, date_lt: int = 19723 + 300, region_lt: int = 224):
    """Top rows (user_id, count, sum(dur), count distinct url) and stats (pairs, groups, cap1, retries)."""
    out = np.zeros((limit, 4), np.int64)
    st = np.zeros(4, np.int64)
    n = C.c_int64()
    _ck(qlib().wsq_c2_where(t.h, date_lt, region_lt, limit, out.ctypes.data, C.byref(n), st.ctypes.data))
    return out[:n.value], st


def c3(fact: DeviceTable, dim: DeviceTable, attr_lt: int = 50):
    """Rows (flag, count, raw price sum, qty sum) and stats (dim rows, filtered fact rows, pairs,
    predicted pairs: 0 without a join model)."""
    out = np.zeros((64, 4), np.int64)
    st = np.zeros(4, np.int64)
    g = C.c_int64()
    _ck(qlib().wsq_c3(fact.h, dim.h, attr_lt, out.ctypes.data, C.byref(g), st.ctypes.data))
    return out[:g.value], st


def set_join_rewrite(on: bool) -> None:
    """True (default): the star join with distinct filtered build keys runs as a bitmap filter
    fused into the fact aggregation; False keeps the hash join."""
    qlib().wsq_set_join_rewrite(1 if on else 0)


def set_sizers(groups=None, joins=None) -> None:
    """Installs sizing models (ws_sizer handles, or None for the plain rules) in the drivers."""
    qlib().wsq_set_sizer(groups)
    qlib().wsq_set_join_sizer(joins)


class C4Set:
    """The many-small-aggregations workload: nq generated segments (rows lognormal in [4096, 4M],
    or rows_override each), one filter + GROUP BY query per segment."""

    def __init__(self, seed: int, nq: int, rows_override: int = 0) -> None:
        h = C.c_void_p()
        tot = C.c_int64()
        _ck(qlib().wsq_c4_load(seed, nq, rows_override, C.byref(h), C.byref(tot)))
        self.h, self.nq, self.rows = h, nq, tot.value

    def run(self, batched: bool, digests: bool = True):
        """(digests, group counts) per query; digests=False fetches the results without digesting
        them (what a timing run measures) and returns None for them."""
        d = np.zeros(self.nq, np.uint64)
        g = np.zeros(self.nq, np.int64)
        _ck(qlib().wsq_c4_run(self.h, 1 if batched else 0, d.ctypes.data if digests else None, g.ctypes.data))
        return (d if digests else None), g

    def fresh_pipeline(self, steps: int, depth: int = 3, digests: bool = True):
        """steps batches of the set, each built from scratch (plans, aggregators, descriptors and
        graph) and pipelined depth deep; (digests, group counts) of the last batch."""
        d = np.zeros(self.nq, np.uint64)
        g = np.zeros(self.nq, np.int64)
        _ck(qlib().wsq_c4_fresh_pipeline(self.h, steps, depth, d.ctypes.data if digests else None, g.ctypes.data))
        return (d if digests else None), g

    def prepare(self) -> "PreparedC4":
        """The set's batch prepared once (launches recorded in a HIP graph) for replays."""
        h = C.c_void_p()
        _ck(qlib().wsq_c4_prepare(self.h, C.byref(h)))
        return PreparedC4(h, self.nq)

    def autotune(self, reps: int = 3) -> dict:
        """Runs the autotuner on this set's batched workload; returns its choice and timings."""
        chosen = np.zeros(3, np.int64)
        ms = np.zeros(2)
        trials = C.c_int()
        _ck(qlib().wsq_c4_autotune(self.h, reps, chosen.ctypes.data, ms.ctypes.data, C.byref(trials)))
        return {"shape": "dense GROUP BY batch (1 key, 5 slots)",
                "chosen": dict(zip(["batch_tile_rows", "batch_lds_slots", "batch_blocks_per_cu"], chosen.tolist())),
                "start_ms": round(float(ms[0]), 3), "tuned_ms": round(float(ms[1]), 3), "trials": trials.value}


    def free(self) -> None:
        if self.h:
            qlib().wsq_c4_free(self.h)
            self.h = None


class PreparedC4:
    """A prepared batch of the C4 workload: run() replays it (digests as C4Set.run)."""

    def __init__(self, h, nq: int) -> None:
        self.h, self.nq = h, nq

    def run(self, digests: bool = True):
        d = np.zeros(self.nq, np.uint64)
        g = np.zeros(self.nq, np.int64)
        _ck(qlib().wsq_c4_replay(self.h, d.ctypes.data if digests else None, g.ctypes.data))
        return (d if digests else None), g

    def pipeline(self, others, steps: int, digests: bool = True):
        """steps batches taking this copy and others (further prepare()s of the same set) in turn:
        with D copies, each batch's groups are copied to the host while the next D - 1 batches'
        kernels run. (digests, group counts) of the last batch."""
        if isinstance(others, PreparedC4):
            others = [others]
        hs = (C.c_void_p * (1 + len(others)))(self.h, *[o.h for o in others])
        d = np.zeros(self.nq, np.uint64)
        g = np.zeros(self.nq, np.int64)
        _ck(qlib().wsq_c4_pipeline(hs, len(hs), steps, d.ctypes.data if digests else None, g.ctypes.data))
        return (d if digests else None), g

    def free(self) -> None:
        if self.h:
            qlib().wsq_c4_release(self.h)
            self.h = None


def autotune_web(t: "DeviceTable", reps: int = 3) -> dict:
    """Runs the autotuner on the web query over t; returns its choice and timings."""
    chosen = np.zeros(3, np.int64)
    ms = np.zeros(2)
    trials = C.c_int()
    _ck(qlib().wsq_c2_autotune(t.h, reps, chosen.ctypes.data, ms.ctypes.data, C.byref(trials)))
    return {"shape": "partitioned GROUP BY (web query, both phases)",
            "chosen": dict(zip(["part_chunks_per_cu", "part_scatter_lds_kib", "part_merge_lds_kib"], chosen.tolist())),
            "start_ms": round(float(ms[0]), 3), "tuned_ms": round(float(ms[1]), 3), "trials": trials.value}

def serve(pool: C4Set, web=None, fact=None, dim=None, rate: float = 100.0, count: int = 200, mode: int = 1,
          workers: int = 2, max_batch: int = 256, seed: int = 5, budget_bytes: int = 0):
    """Runs the open-loop mixed stream; returns (stats dict, digests, types). budget_bytes > 0
    admits requests only while their estimated temporaries fit that much device memory."""
    st = np.zeros(11)
    d = np.zeros(count, np.uint64)
    ty = np.zeros(count, np.int32)
    _ck(qlib().wsq_serve(pool.h, web.h if web else None, fact.h if fact else None, dim.h if dim else None, rate, count, mode,
                         workers, max_batch, seed, budget_bytes, st.ctypes.data, d.ctypes.data, ty.ctypes.data))
    keys = ["completed", "wall_s", "qps", "p50_ms", "p99_ms", "small_p50_ms", "small_p99_ms", "heavy_p50_ms", "heavy_p99_ms",
            "memory_waits", "peak_reserved_bytes"]
    return dict(zip(keys, st.tolist())), d, ty


def serve_alone(pool: C4Set, web=None, fact=None, dim=None, rate: float = 100.0, count: int = 200, seed: int = 5,
                index: int = 0) -> int:
    out = C.c_uint64()
    _ck(qlib().wsq_serve_alone(pool.h, web.h if web else None, fact.h if fact else None, dim.h if dim else None, rate, count,
                               seed, index, C.byref(out)))
    return out.value
