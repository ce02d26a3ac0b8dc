"""ctypes mirror of include/warpstore_amd/ws_abi.h (the plan structs the engine hands over).

Byte layouts equal the engine's own structs (reference src/devtypes.h:85-547); sizes are asserted
at import. The numeric codes (opcodes, type codes, aggregate ops) are the bytecode's values.
"""
from __future__ import annotations

import ctypes as C

MAX_STACK, MAX_COLS, MAX_CODE, MAX_CONSTS, MAX_SETS, MAX_KEYS, MAX_AGGS = 16, 128, 320, 96, 16, 8, 64
INTERP_KEYS, INTERP_AGGS = 4, 16
ENC_HEADER = 16
NULL_I32 = -(2 ** 31)
NULL_I64 = -(2 ** 63)

# column type codes (ColView.t)
T_I32, T_I64, T_F64, T_I32N, T_I64N, T_VEC, T_DECIMAL = range(7)
# physical types
P_I32, P_I64, P_F64, P_VEC = 0, 1, 2, 5
# opcodes
(OP_COL, OP_CONST, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_NEG, OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE,
 OP_AND, OP_OR, OP_NOT, OP_INSET, OP_SELECT, OP_YEAR, OP_ISNULL, OP_COALESCE, OP_NULLIF, OP_CALL1, OP_CALL2,
 OP_DICTMAP, OP_INSORTED, OP_VDOT, OP_VEUC, OP_COLRAW) = range(30)
# aggregate ops and dense kinds
AGG_SUM, AGG_COUNT, AGG_MIN, AGG_MAX = 0, 1, 2, 3
DK_COUNT, DK_SUM_I, DK_SUM_D, DK_MIN_I, DK_MAX_I, DK_MIN_D, DK_MAX_D = range(7)
# kernel kinds and codegen modes
JF_NOKEYS, JF_GROUPED, JF_DENSE, JF_FLAG, JF_FLAGM = range(5)
JIT_OFF, JIT_ON, JIT_SYNC = 0, 1, 2
# scalar function ids used in tests
FN_ABS, FN_CEIL, FN_FLOOR, FN_ROUND, FN_SQRT = 0, 1, 2, 3, 4
FN_EXP, FN_LN, FN_LOG10, FN_LOG2, FN_SIN, FN_COS, FN_TAN, FN_SIGN = 5, 6, 7, 8, 9, 10, 11, 12
FN_ASIN, FN_ACOS, FN_ATAN = 30, 31, 32
FN_MONTH, FN_DAY, FN_DAYOFWEEK, FN_DAYOFYEAR, FN_QUARTER, FN_WEEKDAY, FN_LAST_DAY = 13, 14, 15, 16, 17, 18, 19
FN_WEEK, FN_YEARWEEK = 35, 39
FN_ROUND2, FN_POW, FN_ATAN2, FN_LEAST, FN_GREATEST, FN_TRUNCATE2, FN_ADD_MONTHS = 100, 101, 102, 103, 104, 105, 106
FN_LOG_B, FN_IDIV, FN_DATE_TRUNC_MONTH = 108, 109, 110


class Instr(C.Structure):
    _fields_ = [("op", C.c_uint8), ("is_int", C.c_uint8), ("arg", C.c_uint16)]


class ColView(C.Structure):
    _fields_ = [("p", C.c_void_p), ("t", C.c_uint8), ("scale", C.c_uint8), ("enc", C.c_uint8), ("off", C.c_int32)]


class Prog(C.Structure):
    _fields_ = [("start", C.c_int16), ("len", C.c_int16)]


class ExprEnv(C.Structure):
    _fields_ = [("code", Instr * MAX_CODE), ("consts", C.c_double * MAX_CONSTS), ("sets", C.c_void_p * MAX_SETS),
                ("set_size", C.c_int32 * MAX_SETS), ("maps", C.c_void_p * MAX_SETS), ("map_size", C.c_int32 * MAX_SETS),
                ("cols", ColView * MAX_COLS), ("ovf", C.c_void_p)]


class AggPlan(C.Structure):
    _fields_ = [("has_where", C.c_uint8), ("where", Prog), ("nkeys", C.c_int32), ("keys", Prog * MAX_KEYS),
                ("naggs", C.c_int32), ("aggs", Prog * MAX_AGGS), ("ops", C.c_uint8 * MAX_AGGS),
                ("agg_int", C.c_uint8 * MAX_AGGS), ("agg_bits", C.c_uint8 * MAX_AGGS), ("agg_part", C.c_uint8 * MAX_AGGS),
                ("agg_null", C.c_uint8 * MAX_AGGS), ("dense_priv", C.c_uint8), ("key_int", C.c_uint8 * MAX_KEYS),
                ("genv", C.c_void_p)]


class DenseSpec(C.Structure):
    _fields_ = [("G", C.c_int64), ("copies", C.c_int32), ("copy_bytes", C.c_int32), ("priv", C.c_int32),
                ("kmin", C.c_int64 * MAX_KEYS), ("range", C.c_int64 * MAX_KEYS), ("stride", C.c_int64 * MAX_KEYS),
                ("slot_off", C.c_int32 * MAX_AGGS), ("kind", C.c_uint8 * MAX_AGGS)]


class BatchSet(C.Structure):
    _fields_ = [("nb", C.c_int32), ("ncols", C.c_int32), ("n", C.c_void_p), ("del_", C.c_void_p), ("cols", C.c_void_p)]


class TempColSpec(C.Structure):
    _fields_ = [("src", C.c_int8), ("idx", C.c_int32), ("hi", C.c_int32), ("cnt", C.c_int32), ("kind", C.c_uint8),
                ("key_int", C.c_uint8), ("sum_bits", C.c_uint8), ("out", C.c_uint8), ("div", C.c_double),
                ("mul", C.c_double)]


class ColStats(C.Structure):
    _fields_ = [("rows", C.c_int64), ("nulls", C.c_int64), ("imin", C.c_int64), ("imax", C.c_int64),
                ("dmin", C.c_double), ("dmax", C.c_double), ("hll", C.c_uint8 * 2048)]


class ProfRow(C.Structure):
    _fields_ = [("name", C.c_char * 96), ("launches", C.c_int64), ("total_ms", C.c_double)]


class JitQueryInfo(C.Structure):
    _fields_ = [("compiled", C.c_int), ("hits", C.c_int), ("from_disk", C.c_int), ("background", C.c_int),
                ("compile_ms", C.c_double), ("nids", C.c_int), ("ids", C.c_int * 16)]


class PlanCacheRow(C.Structure):
    _fields_ = [("id", C.c_int), ("kind", C.c_char * 64), ("origin", C.c_char * 16), ("compile_ms", C.c_double),
                ("execs", C.c_int64), ("avg_ms", C.c_double), ("last_ms", C.c_double), ("source_bytes", C.c_size_t),
                ("sql", C.c_char * 256)]


assert C.sizeof(Instr) == 4 and C.sizeof(ColView) == 16 and C.sizeof(Prog) == 4
assert C.sizeof(ExprEnv) == 4488, C.sizeof(ExprEnv)
assert C.sizeof(AggPlan) == 648, C.sizeof(AggPlan)
assert C.sizeof(DenseSpec) == 536, C.sizeof(DenseSpec)
assert C.sizeof(BatchSet) == 32
assert AggPlan.nkeys.offset == 8 and AggPlan.naggs.offset == 44 and AggPlan.key_int.offset == 625
assert ExprEnv.cols.offset == 2432 and DenseSpec.slot_off.offset == 216


def dense_kind(op: int, is_int: bool) -> int:
    """Dense accumulator kind of an aggregate slot (reference engine_select.cpp:1195-1200)."""
    if op == AGG_COUNT:
        return DK_COUNT
    if op == AGG_SUM:
        return DK_SUM_I if is_int else DK_SUM_D
    if op == AGG_MIN:
        return DK_MIN_I if is_int else DK_MIN_D
    return DK_MAX_I if is_int else DK_MAX_D


class EnvBuilder:
    """Builds an ExprEnv the way the engine's bytecode compiler does (engine_core.h:1485-1859):
    column slots with their type codes, deduplicated constants (integer literals keep their int64
    bits in the double slot), and programs as runs of instructions."""

    def __init__(self) -> None:
        self.env = ExprEnv()
        self.ncode = 0
        self.nconst = 0
        self.ncols = 0
        self.nsets = 0
        self.nmaps = 0

    def slot(self, t: int, scale: int = 0, enc: int = 0) -> int:
        s = self.ncols
        self.env.cols[s].t, self.env.cols[s].scale, self.env.cols[s].enc = t, scale, enc
        self.ncols += 1
        return s

    def const_int(self, v: int) -> int:
        d = C.c_double.from_buffer_copy(C.c_int64(v))
        for i in range(self.nconst):
            if bytes(C.c_double(self.env.consts[i])) == bytes(d):
                return i
        self.env.consts[self.nconst] = d.value
        self.nconst += 1
        return self.nconst - 1

    def const(self, v: float) -> int:
        for i in range(self.nconst):
            if self.env.consts[i] == v:
                return i
        self.env.consts[self.nconst] = v
        self.nconst += 1
        return self.nconst - 1

    def add_set(self, dev_ptr: int, nbits: int) -> int:
        i = self.nsets
        self.env.sets[i], self.env.set_size[i] = dev_ptr, nbits
        self.nsets += 1
        return i

    def add_map(self, dev_ptr: int, n: int) -> int:
        i = self.nmaps
        self.env.maps[i], self.env.map_size[i] = dev_ptr, n
        self.nmaps += 1
        return i

    def prog(self, instrs) -> Prog:
        """instrs: list of (op,) / (op, arg) / (op, arg, is_int)."""
        p = Prog(self.ncode, len(instrs))
        for ins in instrs:
            op = ins[0]
            arg = ins[1] if len(ins) > 1 else 0
            is_int = ins[2] if len(ins) > 2 else 0
            self.env.code[self.ncode] = Instr(op, is_int, arg)
            self.ncode += 1
        return p

    # convenience programs
    def col(self, s: int) -> Prog:
        return self.prog([(OP_COL, s)])

    def ilit(self, v: int):
        return (OP_CONST, self.const_int(v), 1)

    def dlit(self, v: float):
        return (OP_CONST, self.const(v), 0)
