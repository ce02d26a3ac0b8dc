"""Writes tests/golden/vm_golden.json: expression-VM golden vectors from the reference's own
semantics (its src/devtypes.h compiled on the host into oracle/_ref by oracle/build.py).

Each case is a bytecode program over a few typed columns (NULL sentinels, NaN, -0.0, encoded
columns), with the values the reference VM (evalProgC) returns for every row, stored as 64-bit
patterns so NaN and -0.0 are exact. Also decoder cases (loadCol / colInt over encodings 1-6),
civil-date round trips and checked int64 arithmetic. Run:  python3 tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import random
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from warpstore_amd import abi  # noqa: E402

NROWS = 16
NULL_I32, NULL_I64 = abi.NULL_I32, abi.NULL_I64


def bits(d: float) -> str:
    return struct.pack("<d", d).hex()


def unbits(h: str) -> float:
    return struct.unpack("<d", bytes.fromhex(h))[0]


def make_columns(rng: random.Random):
    """Six columns, one per scalar type code, with NULLs and edge values."""
    cols = []
    i32 = [rng.choice([rng.randint(-50, 50), 0, 1, -1, 17, 2 ** 31 - 1, -2 ** 31 + 1, 19723, 20087]) for _ in range(NROWS)]
    cols.append(("i32", abi.T_I32, 0, i32))
    i64 = [rng.choice([rng.randint(-10 ** 6, 10 ** 6), 0, 2 ** 40, -2 ** 40, 2 ** 62, 86400000000 * 19723 + 3600000000 * 5])
           for _ in range(NROWS)]
    cols.append(("i64", abi.T_I64, 0, i64))
    f64 = [rng.choice([rng.uniform(-100, 100), 0.0, -0.0, 1.5, float("nan"), 2.5, -3.25, 1e9]) for _ in range(NROWS)]
    cols.append(("f64", abi.T_F64, 0, f64))
    i32n = [rng.choice([rng.randint(-30, 30), NULL_I32, 5]) for _ in range(NROWS)]
    cols.append(("i32n", abi.T_I32N, 0, i32n))
    i64n = [rng.choice([rng.randint(-10 ** 9, 10 ** 9), NULL_I64, 19800]) for _ in range(NROWS)]
    cols.append(("i64n", abi.T_I64N, 0, i64n))
    dec = [rng.choice([rng.randint(-10 ** 7, 10 ** 7), NULL_I64, 12345]) for _ in range(NROWS)]
    cols.append(("dec", abi.T_DECIMAL, 2, dec))
    return cols


BIN = [abi.OP_ADD, abi.OP_SUB, abi.OP_MUL, abi.OP_DIV, abi.OP_MOD, abi.OP_EQ, abi.OP_NE, abi.OP_LT, abi.OP_LE, abi.OP_GT,
       abi.OP_GE, abi.OP_AND, abi.OP_OR, abi.OP_COALESCE, abi.OP_NULLIF]
UN = [abi.OP_NEG, abi.OP_NOT, abi.OP_ISNULL, abi.OP_YEAR]
FN1 = list(range(0, 40))
FN2 = [100, 101, 102, 103, 104, 105, 106, 107, 108, 109, 110]


def gen_expr(rng, depth, ncols, nsets, nmaps):
    """Random expression tree as postfix (op, arg, is_int) triples; consts as ('c', value, is_int)."""
    if depth <= 0 or rng.random() < 0.3:
        r = rng.random()
        if r < 0.55:
            return [(abi.OP_COL, rng.randrange(ncols), 0)]
        if r < 0.65:
            return [(abi.OP_COLRAW, rng.choice([0, 1, 3, 4, 5]), rng.randint(0, 1))]
        if r < 0.85:
            return [("c", rng.choice([0, 1, -1, 3, 7, 500, 19723, 2 ** 33, -5]), 1)]
        return [("c", rng.choice([0.5, -2.25, 1e-3, 3.0, float("nan"), -0.0]), 0)]
    r = rng.random()
    if r < 0.45:
        return gen_expr(rng, depth - 1, ncols, nsets, nmaps) + gen_expr(rng, depth - 1, ncols, nsets, nmaps) + [
            (rng.choice(BIN), 0, 0)]
    if r < 0.6:
        return gen_expr(rng, depth - 1, ncols, nsets, nmaps) + [(rng.choice(UN), 0, 0)]
    if r < 0.7:
        return gen_expr(rng, depth - 1, ncols, nsets, nmaps) + [(abi.OP_CALL1, rng.choice(FN1), 0)]
    if r < 0.78:
        return (gen_expr(rng, depth - 1, ncols, nsets, nmaps) + gen_expr(rng, depth - 1, ncols, nsets, nmaps) +
                [(abi.OP_CALL2, rng.choice(FN2), 0)])
    if r < 0.86:
        return (gen_expr(rng, depth - 1, ncols, nsets, nmaps) + gen_expr(rng, depth - 1, ncols, nsets, nmaps) +
                gen_expr(rng, depth - 1, ncols, nsets, nmaps) + [(abi.OP_SELECT, 0, 0)])
    if r < 0.92:
        return gen_expr(rng, depth - 1, ncols, nsets, nmaps) + [(abi.OP_INSET, rng.randrange(nsets), 0)]
    if r < 0.96:
        return gen_expr(rng, depth - 1, ncols, nsets, nmaps) + [(abi.OP_DICTMAP, rng.randrange(nmaps), 0)]
    return gen_expr(rng, depth - 1, ncols, nsets, nmaps) + [(abi.OP_INSORTED, rng.randrange(nmaps), 0)]


def max_depth(code):
    d = m = 0
    for op, _, _ in code:
        if op == "c" or op in (abi.OP_COL, abi.OP_COLRAW):
            d += 1
        elif op in BIN or op == abi.OP_CALL2:
            d -= 1
        elif op == abi.OP_SELECT:
            d -= 2
        m = max(m, d)
    return m


class Case:
    """An ExprEnv with host columns, sets and maps, ready for either VM."""

    def __init__(self, cols, code, sets, maps):
        self.env = abi.ExprEnv()
        self.keep = []
        for s, (_, t, scale, vals) in enumerate(cols):
            arr = np.array(vals, np.float64 if t == abi.T_F64 else (np.int32 if t in (abi.T_I32, abi.T_I32N) else np.int64))
            self.keep.append(arr)
            self.env.cols[s].p, self.env.cols[s].t, self.env.cols[s].scale = arr.ctypes.data, t, scale
        for i, words in enumerate(sets):
            a = np.array(words, np.uint32)
            self.keep.append(a)
            self.env.sets[i], self.env.set_size[i] = a.ctypes.data, len(words) * 32 - 5
        for i, vals in enumerate(maps):
            a = np.array(vals, np.float64)
            self.keep.append(a)
            self.env.maps[i], self.env.map_size[i] = a.ctypes.data, len(vals)
        nconst = 0
        for pc, (op, arg, is_int) in enumerate(code):
            if op == "c":
                v = arg
                if is_int:
                    self.env.consts[nconst] = struct.unpack("<d", struct.pack("<q", int(v)))[0]
                else:
                    self.env.consts[nconst] = v
                self.env.code[pc] = abi.Instr(abi.OP_CONST, is_int, nconst)
                nconst += 1
            else:
                self.env.code[pc] = abi.Instr(op, is_int, arg)
        self.prog = abi.Prog(0, len(code))


def gen_cases(seed=2024, n=160):
    rng = random.Random(seed)
    cases = []
    while len(cases) < n:
        cols = make_columns(rng)
        sets = [[rng.getrandbits(32) for _ in range(3)]]
        maps = [sorted(rng.uniform(-50, 50) for _ in range(8)) + [float("nan")]]
        maps[0] = [m for m in maps[0]]
        code = gen_expr(rng, rng.randint(1, 4), len(cols), len(sets), len(maps))
        if len(code) > 60 or max_depth(code) > 14:
            continue
        cases.append({"cols": [(nm, t, sc, v) for nm, t, sc, v in cols], "code": code, "sets": sets, "maps": maps})
    return cases


def encoded_cases(rng: random.Random):
    """Encoded segment columns (enc 1-6) with NULL top codes and a row offset."""
    out = []
    for enc, t in [(1, abi.T_I32N), (2, abi.T_I32), (3, abi.T_I64N), (1, abi.T_I64), (2, abi.T_DECIMAL), (3, abi.T_I32),
                   (4, abi.T_F64), (5, abi.T_F64), (6, abi.T_F64)]:
        n = 40
        w = {1: 1, 2: 2, 3: 4, 4: 1, 5: 2, 6: 4}[enc]
        top = (1 << (8 * w)) - 1
        base = rng.randint(-1000, 1000)
        codes = [rng.choice([rng.randrange(0, min(top, 300)), top]) for _ in range(n)]
        header = struct.pack("<q", base)
        if enc in (5, 6):
            header += struct.pack("<d", 100.0)
        else:
            header += b"\0" * 8
        if enc == 4:
            table = [rng.uniform(-5, 5) for _ in range(256)]
            table[3] = float("nan")
            payload = struct.pack("<256d", *table) + bytes(c % 256 for c in codes)
        else:
            fmt = {1: "B", 2: "H", 4: "I"}[w]
            payload = struct.pack(f"<{n}{fmt}", *codes)
        blob = header + payload + b"\0" * 16
        out.append({"enc": enc, "t": t, "scale": 2 if t == abi.T_DECIMAL else 0, "off": 3, "rows": n - 3,
                    "blob": blob.hex()})
    return out


def main():
    import helpers as H
    R = H.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref_devtypes.so is missing: python3 oracle/build.py (needs the reference tree)")
    out = {"vm": [], "decode": [], "dates": [], "checked": [], "call1": [], "call2": []}
    for case in gen_cases():
        c = Case(case["cols"], case["code"], case["sets"], case["maps"])
        vals = [bits(R.ref_eval_prog(C.addressof(c.env), None, 0, c.prog.len, r)) for r in range(NROWS)]
        out["vm"].append({**case, "expect": vals})
    rng = random.Random(7)
    for d in encoded_cases(rng):
        blob = np.frombuffer(bytes.fromhex(d["blob"]), np.uint8).copy()
        v = abi.ColView()
        v.p, v.t, v.scale, v.enc, v.off = blob.ctypes.data, d["t"], d["scale"], d["enc"], d["off"]
        d["load"] = [bits(R.ref_load_col(C.addressof(v), r)) for r in range(d["rows"])]
        d["colint"] = [int(R.ref_col_int(C.addressof(v), r)) if d["enc"] < 4 else 0 for r in range(d["rows"])]
        out["decode"].append(d)
    for z in list(range(-800000, 800000, 7919)) + [0, -1, 11016, 11017, 19723, 2932896, -719468, -719469]:
        y, m, dd = C.c_int64(), C.c_uint(), C.c_uint()
        R.ref_civil_from_days(z, C.byref(y), C.byref(m), C.byref(dd))
        out["dates"].append([z, y.value, m.value, dd.value, int(R.ref_days_from_civil(y.value, m.value, dd.value))])
    vals = [0, 1, -1, 2, 3, 2 ** 31, 2 ** 32 + 7, 2 ** 62, -(2 ** 62), 2 ** 63 - 1, -(2 ** 63), 123456789012, -98765]
    for op in range(3):
        for a in vals:
            for b in vals:
                f = C.c_int(0)
                r = int(R.ref_checked(op, a, b, C.byref(f)))
                out["checked"].append([op, a, b, r, f.value])
    for fn in range(0, 40):
        for x in [19723.0, 0.0, -1.0, 738000.0, 1.7e15, -2.5, 2.5, 0.5, float("nan"), 59.0, 60.0, 365.0, -36525.0]:
            out["call1"].append([fn, bits(x), bits(R.ref_call1(fn, x))])
    for fn in FN2:
        for a, b in [(19723.0, 1.0), (19723.0, -13.0), (2.5, 1.0), (-7.0, 2.0), (1.7e15, 5.0), (3.0, 0.0), (0.0, 0.0),
                     (float("nan"), 1.0), (11016.0, 0.0), (11016.0, 1.0), (2.0, 10.0)]:
            out["call2"].append([fn, bits(a), bits(b), bits(R.ref_call2(fn, a, b))])
    path = os.path.join(HERE, "vm_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"), default=lambda o: o)
    print(f"wrote {path}: {len(out['vm'])} VM cases, {len(out['decode'])} decoder cases")


if __name__ == "__main__":
    main()
