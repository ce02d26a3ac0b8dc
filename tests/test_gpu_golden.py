"""The reference's golden vectors replayed on the device: the 160 recorded VM programs (values the
reference's own evalProgC returned, reference src/devtypes.h:430-495) and the 9 decoder blobs
(encodings 1-6, devtypes.h:123-152), through
  - the device interpreter (ws_eval_at: the stack VM in a kernel), against the golden values;
  - compiled plans (the code generator's typed evaluator under hipRTC): 16 programs per plan, each
    program the argument of a MIN and a MAX slot of a GROUP BY over a row id, so every group holds
    one row and MIN = the program's value (MIN = +inf with MAX = -inf: NULL). A compiled plan
    types subexpressions as exact int64 where the reference's compiler does (src/jit.cu:107-272),
    so it is compared with the oracle's typed restatement (or_eval_typed), which equals the golden
    VM value on every row where no int64 node changes the value (pinned in test_oracle.py).
Bit-exact (-0.0 included; NULL is any NaN: the sign and payload of a NaN carry no meaning, and
subtracting a NULL flips the sign bit on this hardware), except programs that call a
transcendental scalar function (exp, ln, log10, log2, trigonometry, pow, round/truncate to
digits, log base): the device math library may differ from the host's in the last place, so
those compare within 1e-13 relative. Rows whose typed value went through an out-of-range double
-> int64 conversion (undefined in C++, answered differently by each processor) are counted and
left out. Also the checked-int64 overflow flag (devtypes.h:196-215): raised by compiled plans
only, as in the reference (its interpreter computes in doubles and has no flag)."""
import ctypes as C
import json
import os
import struct

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi
from golden.make_golden import Case, NROWS

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "vm_golden.json")
# scalar functions whose device results may differ from the host libm in the last place
TRANSCENDENTAL1 = {abi.FN_EXP, abi.FN_LN, abi.FN_LOG10, abi.FN_LOG2, abi.FN_SIN, abi.FN_COS, abi.FN_TAN, abi.FN_ASIN,
                   abi.FN_ACOS, abi.FN_ATAN}
TRANSCENDENTAL2 = {abi.FN_ROUND2, abi.FN_TRUNCATE2, abi.FN_POW, abi.FN_ATAN2, abi.FN_LOG_B}
REL_TOL = 1e-13


def fbits(x: float) -> str:
    return struct.pack("<d", x).hex()


def unbits(h: str) -> float:
    return struct.unpack("<d", bytes.fromhex(h))[0]


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def transcendental(code) -> bool:
    return any((op == abi.OP_CALL1 and arg in TRANSCENDENTAL1) or (op == abi.OP_CALL2 and arg in TRANSCENDENTAL2)
               for op, arg, _ in code if op != "c")


def agree(got: float, want_hex: str, loose: bool) -> bool:
    if fbits(got) == want_hex:
        return True
    want = unbits(want_hex)
    if np.isnan(got) and np.isnan(want):
        return True
    if not loose:
        return False
    if np.isinf(got) or np.isinf(want) or np.isnan(got) or np.isnan(want):
        return got == want
    return abs(got - want) <= REL_TOL * max(abs(want), abs(got))


class DeviceCase:
    """A golden case's columns, set and map uploaded; env pointers are device addresses."""

    def __init__(self, case):
        self.keep = []
        self.env = abi.ExprEnv()
        for s, (_, t, scale, vals) in enumerate(case["cols"]):
            arr = np.array(vals, np.float64 if t == abi.T_F64 else (np.int32 if t in (abi.T_I32, abi.T_I32N) else np.int64))
            d = ws.DevBuf.from_array(arr)
            self.keep.append(d)
            self.env.cols[s].p, self.env.cols[s].t, self.env.cols[s].scale = d.p, t, scale
        sw = np.array(case["sets"][0], np.uint32)
        ds = ws.DevBuf.from_array(sw)
        mp = np.array(case["maps"][0], np.float64)
        dm = ws.DevBuf.from_array(mp)
        self.keep += [ds, dm]
        self.env.sets[0], self.env.set_size[0] = ds.p, len(sw) * 32 - 5
        self.env.maps[0], self.env.map_size[0] = dm.p, len(mp)
        nconst = 0
        for pc, (op, arg, is_int) in enumerate(case["code"]):
            if op == "c":
                self.env.consts[nconst] = struct.unpack("<d", struct.pack("<q", int(arg)))[0] if is_int else arg
                self.env.code[pc] = abi.Instr(abi.OP_CONST, is_int, nconst)
                nconst += 1
            else:
                self.env.code[pc] = abi.Instr(op, is_int, arg)
        self.prog = abi.Prog(0, len(case["code"]))


def test_golden_vm_on_device_interpreter(gpu, golden):
    H.say("case golden VM programs, device interpreter")
    out = ws.DevBuf(NROWS * 8)
    bad = []
    for i, case in enumerate(golden["vm"]):
        dc = DeviceCase(case)
        ws.check(gpu.ws_eval_at(C.byref(dc.env), dc.prog, None, NROWS, out.p))
        got = out.to_array(np.float64)
        loose = transcendental(case["code"])
        for r in range(NROWS):
            if not agree(got[r], case["expect"][r], loose):
                bad.append((i, r, fbits(got[r]), case["expect"][r], case["code"]))
    assert not bad, f"{len(bad)} mismatches, first: {bad[:3]}"


def pack_plans(cases, per_plan=16):
    """Groups of golden cases as one compiled GROUP BY plan each: case j's rows are rows
    16 j .. 16 j + 15 of six shared columns plus a row id; its program (literals renumbered, set
    and map j) is the argument of slots 1 + 2 j (MIN) and 2 + 2 j (MAX)."""
    plans = []
    for g0 in range(0, len(cases), per_plan):
        group = cases[g0:g0 + per_plan]
        nrows = NROWS * len(group)
        keep = []
        b = abi.ExprEnv()
        for s in range(6):
            t, scale = group[0]["cols"][s][1], group[0]["cols"][s][2]
            dt = np.float64 if t == abi.T_F64 else (np.int32 if t in (abi.T_I32, abi.T_I32N) else np.int64)
            arr = np.concatenate([np.array(c["cols"][s][3], dt) for c in group])
            d = ws.DevBuf.from_array(arr)
            keep.append(d)
            b.cols[s].p, b.cols[s].t, b.cols[s].scale = d.p, t, scale
        rid = ws.DevBuf.from_array(np.arange(nrows, dtype=np.int32))
        keep.append(rid)
        b.cols[6].p, b.cols[6].t = rid.p, abi.T_I32
        plan = abi.AggPlan()
        plan.nkeys, plan.key_int[0] = 1, 1
        b.code[0] = abi.Instr(abi.OP_COL, 0, 6)
        plan.keys[0] = abi.Prog(0, 1)
        pc, nconst = 1, 0
        plan.naggs = 1 + 2 * len(group)
        plan.ops[0], plan.agg_int[0] = abi.AGG_COUNT, 1
        for j, case in enumerate(group):
            sw = ws.DevBuf.from_array(np.array(case["sets"][0], np.uint32))
            mp = ws.DevBuf.from_array(np.array(case["maps"][0], np.float64))
            keep += [sw, mp]
            b.sets[j], b.set_size[j] = sw.p, len(case["sets"][0]) * 32 - 5
            b.maps[j], b.map_size[j] = mp.p, len(case["maps"][0])
            start = pc
            for op, arg, is_int in case["code"]:
                if op == "c":
                    b.consts[nconst] = struct.unpack("<d", struct.pack("<q", int(arg)))[0] if is_int else arg
                    b.code[pc] = abi.Instr(abi.OP_CONST, is_int, nconst)
                    nconst += 1
                elif op in (abi.OP_INSET, abi.OP_DICTMAP, abi.OP_INSORTED):
                    b.code[pc] = abi.Instr(op, is_int, j)
                else:
                    b.code[pc] = abi.Instr(op, is_int, arg)
                pc += 1
            prog = abi.Prog(start, pc - start)
            for k, op in ((1 + 2 * j, abi.AGG_MIN), (2 + 2 * j, abi.AGG_MAX)):
                plan.ops[k], plan.aggs[k] = op, prog
        assert pc <= abi.MAX_CODE and nconst <= abi.MAX_CONSTS
        plans.append((group, plan, b, keep, nrows))
    return plans


def run_compiled(gpu, plan, env, nrows, kind):
    gpu.ws_jit_set_mode(abi.JIT_SYNC)
    jit = C.c_void_p()
    ws.check(gpu.ws_jit_get(C.byref(env), 7, C.byref(plan), abi.JF_GROUPED, kind, C.byref(jit)))
    assert jit.value, "the plan did not compile"
    h = C.c_void_p()
    ws.check(gpu.ws_agg_create(C.byref(plan), 1 << 12, None, C.byref(h)))
    try:
        ws.check(gpu.ws_agg_add_batch(h, C.byref(env), 7, nrows, None))
        ws.check(gpu.ws_agg_run(h, C.byref(env), jit))
        g = C.c_int64()
        ws.check(gpu.ws_agg_fetch_begin(h, C.byref(g)))
        keys = np.zeros(g.value)
        accs = np.zeros(g.value * plan.naggs)
        ws.check(gpu.ws_agg_fetch_copy(h, ws.ptr(keys), ws.ptr(accs)))
    finally:
        gpu.ws_agg_destroy(h)
    rid = keys.view(np.int64)
    accs = accs.reshape(g.value, plan.naggs)[np.argsort(rid)]
    assert np.array_equal(np.sort(rid), np.arange(nrows))
    return accs


def typed_expect(case):
    """The oracle's compiled-plan value of a golden program per row, as a hex bit pattern, or None
    where it went through an out-of-range double -> int64 conversion."""
    O = H.oracle()
    O.or_eval_typed.restype = C.c_double
    O.or_eval_typed.argtypes = [C.POINTER(abi.ExprEnv), C.c_void_p, abi.Prog, C.c_int64, C.POINTER(C.c_int)]
    c = Case(case["cols"], [tuple(x) for x in case["code"]], case["sets"], case["maps"])
    out = []
    for r in range(NROWS):
        fl = C.c_int()
        v = O.or_eval_typed(C.byref(c.env), None, c.prog, r, C.byref(fl))
        assert not fl.value & 2, ("unexplained typed difference", case["code"], r)
        out.append(None if fl.value & 4 else fbits(v))
    return out


def test_golden_vm_compiled_plans(gpu, golden):
    bad, skipped, typed_rows = [], 0, 0
    for pi, (group, plan, env, keep, nrows) in enumerate(pack_plans(golden["vm"])):
        H.say(f"case golden VM programs, compiled plan {pi}")
        accs = run_compiled(gpu, plan, env, nrows, b"golden programs")
        for j, case in enumerate(group):
            loose = transcendental(case["code"])
            want = typed_expect(case)
            for r in range(NROWS):
                if want[r] is None:
                    skipped += 1
                    continue
                typed_rows += want[r] != case["expect"][r]
                mn, mx = accs[j * NROWS + r, 1 + 2 * j], accs[j * NROWS + r, 2 + 2 * j]
                got = float("nan") if (mn == np.inf and mx == -np.inf) else mn
                if not agree(got, want[r], loose):
                    bad.append((pi * 16 + j, r, fbits(got), want[r], case["expect"][r], case["code"]))
    H.say(f"compiled plans: {len(golden['vm']) * NROWS - skipped} rows compared ({typed_rows} where int64 typing "
          f"differs from the VM), {skipped} not portable")
    assert not bad, f"{len(bad)} mismatches, first: {bad[:3]}"
    assert skipped <= 16


def test_golden_decoders_on_device(gpu, golden):
    """Encodings 1-6 decoded by the device VM's column load (ws_eval_at), by the encoded gather
    (ws_gather_enc: stored integers with NULL sentinels, decoded doubles) and by a compiled plan."""
    H.say("case golden decoder blobs on the device")
    for d in golden["decode"]:
        blob = ws.DevBuf.from_array(np.frombuffer(bytes.fromhex(d["blob"]), np.uint8).copy())
        rows = d["rows"]
        env = abi.ExprEnv()
        v = env.cols[0]
        v.p, v.t, v.scale, v.enc, v.off = blob.p, d["t"], d["scale"], d["enc"], d["off"]
        env.code[0] = abi.Instr(abi.OP_COL, 0, 0)
        out = ws.DevBuf(rows * 8)
        ws.check(gpu.ws_eval_at(C.byref(env), abi.Prog(0, 1), None, rows, out.p))
        got = out.to_array(np.float64)
        assert [fbits(x) for x in got] == d["load"], (d["enc"], d["t"])
        idx = ws.DevBuf.from_array(np.arange(rows, dtype=np.int32))
        if d["enc"] < 4:
            ptype = abi.P_I32 if d["t"] in (abi.T_I32, abi.T_I32N) else abi.P_I64
            ws.check(gpu.ws_gather_enc(C.byref(env.cols[0]), ptype, idx.p, rows, out.p))
            dt = np.int32 if ptype == abi.P_I32 else np.int64
            assert np.array_equal(out.to_array(dt, rows).astype(np.int64), np.array(d["colint"], np.int64)), d["enc"]
        else:
            ws.check(gpu.ws_gather_enc(C.byref(env.cols[0]), abi.P_F64, idx.p, rows, out.p))
            assert [fbits(x) for x in out.to_array(np.float64)] == d["load"], d["enc"]
        # compiled: GROUP BY row id, MIN / MAX of the column (COL reads through the typed decode)
        rid = ws.DevBuf.from_array(np.arange(rows, dtype=np.int32))
        env.cols[1].p, env.cols[1].t = rid.p, abi.T_I32
        env.code[1] = abi.Instr(abi.OP_COL, 0, 1)
        plan = abi.AggPlan()
        plan.nkeys, plan.key_int[0], plan.keys[0] = 1, 1, abi.Prog(1, 1)
        plan.naggs = 3
        plan.ops[0], plan.agg_int[0] = abi.AGG_COUNT, 1
        plan.ops[1], plan.aggs[1] = abi.AGG_MIN, abi.Prog(0, 1)
        plan.ops[2], plan.aggs[2] = abi.AGG_MAX, abi.Prog(0, 1)
        accs = run_compiled(gpu, plan, env, rows, b"golden decode")
        for r in range(rows):
            mn, mx = accs[r, 1], accs[r, 2]
            got = float("nan") if (mn == np.inf and mx == -np.inf) else mn
            assert fbits(got) == d["load"][r], (d["enc"], d["t"], r)


@pytest.mark.parametrize("mode", ["interp", "jit"])
def test_checked_overflow_raises_flag(gpu, mode):
    """A compiled plan whose int64 product overflows raises env.ovf (the engine then reports error
    1690, reference devtypes.h:196-215, jit.cu:159-160); a compiled plan without overflow leaves
    it clear, and so does the interpreter, which evaluates in doubles (devtypes.h:430-495)."""
    H.say(f"case overflow flag {mode}")
    n = 4096
    a = np.full(n, 3, np.int64)
    a[1234] = 2 ** 62
    da = ws.DevBuf.from_array(a)
    flag = ws.DevBuf.from_array(np.zeros(1, np.int32))
    for big, want in ((4, 1), (1, 0)):
        zero = np.zeros(1, np.int32)
        ws.check(gpu.ws_h2d(flag.p, ws.ptr(zero), 4))
        b = abi.EnvBuilder()
        s = b.slot(abi.T_I64)
        b.env.cols[s].p = da.p
        b.env.ovf = flag.p
        plan = abi.AggPlan()
        plan.naggs = 2
        plan.ops[0], plan.agg_int[0] = abi.AGG_COUNT, 1
        plan.ops[1], plan.agg_int[1] = abi.AGG_SUM, 1
        plan.aggs[1] = b.prog([(abi.OP_COL, s), b.ilit(big), (abi.OP_MUL,)])
        jit = None
        if mode == "jit":
            gpu.ws_jit_set_mode(abi.JIT_SYNC)
            jit = C.c_void_p()
            ws.check(gpu.ws_jit_get(C.byref(b.env), 1, C.byref(plan), abi.JF_NOKEYS, b"overflow", C.byref(jit)))
            assert jit.value
        h = C.c_void_p()
        ws.check(gpu.ws_agg_create(C.byref(plan), 1, None, C.byref(h)))
        try:
            ws.check(gpu.ws_agg_add_batch(h, C.byref(b.env), 1, n, None))
            ws.check(gpu.ws_agg_run(h, C.byref(b.env), jit))
            g = C.c_int64()
            ws.check(gpu.ws_agg_fetch_begin(h, C.byref(g)))
        finally:
            gpu.ws_agg_destroy(h)
        assert int(flag.to_array(np.int32)[0]) == (want if mode == "jit" else 0), (big, mode)
