"""The CPU oracle, the CPU baseline and the workload generators built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only) and exercised in a child process: aggregation,
selection, joins, sort, the query restatements, typed evaluation over the golden programs, and
every generator. Any sanitizer report fails the test."""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import ctypes as C, json, os, sys
sys.path.insert(0, os.path.join(sys.argv[1], "tests")); sys.path.insert(0, sys.argv[1])
import numpy as np
import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi
from golden.make_golden import Case, NROWS
H._oracle = None
os.environ["WS_SANITIZED"] = "1"
O = C.CDLL(sys.argv[2]); G = C.CDLL(sys.argv[3]); B = C.CDLL(sys.argv[4])
# reuse the binding code on the sanitized libraries
H._oracle = None
orig = H.oracle
import ctypes
def bound_oracle():
    L = O
    P, I, I64, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
    L.or_aggregate.restype = I64
    L.or_aggregate.argtypes = [C.POINTER(abi.AggPlan), I, C.POINTER(abi.ExprEnv), I, P, P, P, I, P, P, P, I64]
    L.or_select.restype = I64
    L.or_select.argtypes = [C.POINTER(abi.ExprEnv), I, abi.Prog, I, P, P, P, I, P, P]
    L.or_join.restype = I64
    L.or_join.argtypes = [C.POINTER(abi.ColView), I64, C.POINTER(abi.ColView), I64, I, P, P, I64]
    L.or_sort_perm.argtypes = [P, P, I, I64, P]
    L.or_eval_typed.restype = D
    L.or_eval_typed.argtypes = [C.POINTER(abi.ExprEnv), P, abi.Prog, I64, C.POINTER(I)]
    L.or_eval_prog.restype = D
    L.or_eval_prog.argtypes = [C.POINTER(abi.ExprEnv), P, abi.Prog, I64]
    return L
H._oracle = bound_oracle()
ws._gen = None
ws.DATAGEN_PATH = sys.argv[3]
# generators
k, v, p = H.gen_c1(50_000)
user, region, date, url, dur = H.gen_c2(60_000, n_users=3000, n_urls=900)
fk, qty, price, flag, dkey, dattr = H.gen_c3(80_000, 5_000)
# query restatements
full = H.oracle_c2(user, region, date, url, dur)
rows = H.oracle_c3(fk, qty, price, flag, dkey, dattr)
assert len(full) > 0 and len(rows) > 0
# aggregation (hash and dense semantics), selection, join, sort
t = H.Table({"k": k, "v": v, "p": p}, [20_000, 0, 30_000])
b = abi.EnvBuilder(); sk, sv, sp = b.slot(abi.T_I32), b.slot(abi.T_I64), b.slot(abi.T_F64)
plan = H.plan_c1(b, sk, sv, sp)
host = H.Bound(t, [("k", abi.T_I32, 0), ("v", abi.T_I64, 0), ("p", abi.T_F64, 0)], False)
for dense in (False, True):
    H.oracle_aggregate(plan, b.env, host, dense=dense, max_groups=2000)
out = np.zeros(len(k), np.int32); per = np.zeros(3, np.int64)
sizes = np.array(t.sizes, np.int64)
O.or_select(C.byref(b.env), 1, plan.where, 3, ws.ptr(sizes), None, C.cast(host.views, C.c_void_p), 3,
            ws.ptr(out), ws.ptr(per))
bk = dkey[:2000].astype(np.int32); pk = fk[:20000].astype(np.int32)
bv, pv = abi.ColView(), abi.ColView()
bv.p, bv.t, pv.p, pv.t = bk.ctypes.data, abi.T_I32, pk.ctypes.data, abi.T_I32
ob = np.zeros(40000, np.int32); op = np.zeros(40000, np.int32)
for outer in (0, 1, 2):
    O.or_join(C.byref(bv), len(bk), C.byref(pv), len(pk), outer, ws.ptr(ob), ws.ptr(op), len(ob))
keys0 = np.round(p[:5000], 1); desc = np.array([1], np.uint8); perm = np.zeros(5000, np.int32)
O.or_sort_perm((C.c_void_p * 1)(keys0.ctypes.data), desc.ctypes.data, 1, 5000, perm.ctypes.data)
# typed and VM evaluation over the golden programs
g = json.load(open(os.path.join(sys.argv[1], "tests", "golden", "vm_golden.json")))
for case in g["vm"]:
    c = Case(case["cols"], [tuple(x) for x in case["code"]], case["sets"], case["maps"])
    for r in range(NROWS):
        fl = C.c_int()
        O.or_eval_typed(C.byref(c.env), None, c.prog, r, C.byref(fl))
        O.or_eval_prog(C.byref(c.env), None, c.prog, r)
# the CPU baseline's multithreaded restatements
P, I64, I = C.c_void_p, C.c_int64, C.c_int
B.cb_c3.restype = I64
B.cb_c3.argtypes = [P, P, P, P, I64, P, P, I64, C.c_int32, I, P, I64]
res = np.zeros((256, 4), np.int64)
n = B.cb_c3(fk.ctypes.data, qty.ctypes.data, price.ctypes.data, flag.ctypes.data, len(fk), dkey.ctypes.data, dattr.ctypes.data,
            len(dkey), 50, 4, res.ctypes.data, 256)
assert n == len(rows) and np.array_equal(res[:n], rows)
print("sanitized run ok", flush=True)
'''


def _asan_runtime():
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_and_host_code_under_sanitizers():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("no AddressSanitizer runtime for g++ here")
    flags = ["-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include")]
    with tempfile.TemporaryDirectory() as d:
        libs = {}
        for name, src, extra in (("oracle", "oracle/ws_oracle.cpp", []),
                                 ("datagen", "warpstore_amd/csrc/ws_datagen.cpp", ["-fopenmp"]),
                                 ("baseline", "oracle/cpu_baseline.cpp", ["-pthread"])):
            out = os.path.join(d, f"lib{name}_san.so")
            r = subprocess.run(["g++"] + flags + extra + [os.path.join(ROOT, src), "-o", out], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-3000:]
            libs[name] = out
        env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
        r = subprocess.run([sys.executable, "-c", CHILD, ROOT, libs["oracle"], libs["datagen"], libs["baseline"]],
                           capture_output=True, text=True, env=env, timeout=600)
        report = (r.stdout + r.stderr)[-4000:]
        assert r.returncode == 0 and "sanitized run ok" in r.stdout, report
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, report
