// Row-expression semantics shared by the interpreter kernels and the runtime-compiled kernels.
//
// Everything here is a restatement of the engine's bytecode semantics (reference src/devtypes.h):
// NULL is NaN inside expressions and a sentinel in integer storage, comparisons and boolean
// connectives are three-valued, integer constants carry their exact int64 bits in a double slot,
// and encoded segment columns are decoded inline. Compiles as HIP (hipcc), under hipRTC and as
// plain host C++ (the device qualifiers collapse to `inline`).
#pragma once
#include "ws_abi.h"

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define WS_HD __host__ __device__ __forceinline__
#else
#define WS_HD inline
#include <cmath>
#endif

namespace ws {

typedef long long i64;
typedef unsigned long long u64;

constexpr i64 kI64Max = 0x7fffffffffffffffll;
constexpr i64 kI64Min = -kI64Max - 1;

WS_HD i64 dbits(double d) { return __builtin_bit_cast(i64, d); }
WS_HD double bitsd(i64 v) { return __builtin_bit_cast(double, v); }
WS_HD double nullv() { return bitsd(0x7ff8000000000000ll); }
WS_HD bool isnan_(double v) { return v != v; }
WS_HD bool truthy(double v) { return v == v && v != 0; }  // NULL and 0 are not true

// ---- civil calendar <-> days since 1970-01-01 (H. Hinnant's public-domain algorithms;
//      reference devtypes.h:41-60)
WS_HD i64 days_from_civil(i64 y, unsigned m, unsigned d) {
  y -= m <= 2 ? 1 : 0;
  const i64 era = (y >= 0 ? y : y - 399) / 400;
  const unsigned yoe = (unsigned)(y - era * 400);
  const unsigned mp = m > 2 ? m - 3 : m + 9;
  const unsigned doy = (153 * mp + 2) / 5 + d - 1;
  const unsigned doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + (i64)doe - 719468;
}
WS_HD void civil_from_days(i64 z, i64& y, unsigned& m, unsigned& d) {
  z += 719468;
  const i64 era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = (unsigned)(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const unsigned mp = (5 * doy + 2) / 153;
  d = doy - (153 * mp + 2) / 5 + 1;
  m = mp < 10 ? mp + 3 : mp - 9;
  y = (i64)yoe + era * 400 + (m <= 2 ? 1 : 0);
}

// ---- encoded segment columns (reference devtypes.h:123-152; encoders storage.cu:1097-1194)
WS_HD int enc_width(uint8_t enc) { return (enc == 1 || enc == 4) ? 1 : (enc == 2 || enc == 5) ? 2 : 4; }

// Stored integer of an encoded integer column: base + code, the top code being the type's NULL.
WS_HD i64 enc_int_base(const ws_colview& c, i64 base, i64 r) {
  const unsigned char* data = (const unsigned char*)c.p + WS_ENC_HEADER;
  const i64 i = r + c.off;
  const i64 nul = (c.t == WS_T_I32 || c.t == WS_T_I32N) ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
  unsigned code, top;
  if (c.enc == 1) {
    code = data[i];
    top = 0xFFu;
  } else if (c.enc == 2) {
    code = ((const unsigned short*)data)[i];
    top = 0xFFFFu;
  } else {
    code = ((const unsigned*)data)[i];
    top = 0xFFFFFFFFu;
  }
  return code == top ? nul : base + (i64)code;
}
WS_HD i64 enc_int(const ws_colview& c, i64 r) { return enc_int_base(c, *(const i64*)c.p, r); }
WS_HD double enc_dbl(const ws_colview& c, i64 r) {
  const unsigned char* data = (const unsigned char*)c.p + WS_ENC_HEADER;
  if (c.enc == 4) return ((const double*)data)[data[256 * 8 + r + c.off]];  // 256-entry value table, u8 codes
  const i64 base = *(const i64*)c.p;
  const double p10 = *(const double*)((const unsigned char*)c.p + 8);
  const i64 i = r + c.off;
  if (c.enc == 5) {
    const unsigned v = ((const unsigned short*)data)[i];
    return v == 0xFFFFu ? nullv() : (double)(base + (i64)v) / p10;
  }
  const unsigned v = ((const unsigned*)data)[i];
  return v == 0xFFFFFFFFu ? nullv() : (double)(base + (i64)v) / p10;
}
// Stored value of an integer-typed view, raw or encoded (NULL as the type's sentinel).
WS_HD i64 col_int(const ws_colview& c, i64 r) {
  if (c.enc) return enc_int(c, r);
  return (c.t == WS_T_I32 || c.t == WS_T_I32N) ? (i64)((const int32_t*)c.p)[r] : (i64)((const int64_t*)c.p)[r];
}
WS_HD double col_dbl(const ws_colview& c, i64 r) { return c.enc ? enc_dbl(c, r) : ((const double*)c.p)[r]; }
WS_HD double pow10i(int s) {
  double p = 1;
  for (int i = 0; i < s; ++i) p *= 10;
  return p;
}
WS_HD double null_if_i32(i64 v) { return v == (i64)WS_NULL_I32 ? nullv() : (double)v; }
WS_HD double null_if_i64(i64 v) { return v == (i64)WS_NULL_I64 ? nullv() : (double)v; }
WS_HD double dec_of(i64 v, double p10) { return v == (i64)WS_NULL_I64 ? nullv() : (double)v / p10; }

// A column value as the VM sees it: a double, NULL as NaN (reference devtypes.h:217-245).
WS_HD double load_col(const ws_colview& c, i64 r) {
  if (c.enc) {
    if (c.t == WS_T_F64) return enc_dbl(c, r);
    const i64 v = enc_int(c, r);
    if (c.t == WS_T_I32N) return null_if_i32(v);
    if (c.t == WS_T_I64N) return null_if_i64(v);
    if (c.t == WS_T_DECIMAL) return v == (i64)WS_NULL_I64 ? nullv() : (double)v / pow10i(c.scale);
    return (double)v;
  }
  switch (c.t) {
    case WS_T_I32: return (double)((const int32_t*)c.p)[r];
    case WS_T_I64: return (double)((const int64_t*)c.p)[r];
    case WS_T_I32N: return null_if_i32(((const int32_t*)c.p)[r]);
    case WS_T_I64N: return null_if_i64(((const int64_t*)c.p)[r]);
    case WS_T_DECIMAL: {
      const i64 v = ((const int64_t*)c.p)[r];
      return v == (i64)WS_NULL_I64 ? nullv() : (double)v / pow10i(c.scale);
    }
    default: return ((const double*)c.p)[r];
  }
}

// ---- checked int64 arithmetic: wraps, and raises *f on signed overflow (devtypes.h:196-215)
WS_HD i64 add_chk(i64 a, i64 b, int* f) {
  const i64 r = (i64)((u64)a + (u64)b);
  if ((((a ^ r) & (b ^ r)) < 0) && f) *f = 1;
  return r;
}
WS_HD i64 sub_chk(i64 a, i64 b, int* f) {
  const i64 r = (i64)((u64)a - (u64)b);
  if ((((a ^ b) & (a ^ r)) < 0) && f) *f = 1;
  return r;
}
WS_HD i64 mul_chk(i64 a, i64 b, int* f) {
  i64 r;
  const bool o = __builtin_mul_overflow(a, b, &r);
  if (o && f) *f = 1;
  return r;
}

// ---- three-valued helpers (NULL = NaN; devtypes.h:247-304)
// Comparison k of WS_OP_EQ + k (equal, unequal, less, less-or-equal, greater, greater-or-equal).
WS_HD double tri_cmp(int op, double a, double b) {
  if (a != a || b != b) return nullv();
  bool r;
  switch (op) {
    case 0: r = a == b; break;
    case 1: r = a != b; break;
    case 2: r = a < b; break;
    case 3: r = a <= b; break;
    case 4: r = a > b; break;
    default: r = a >= b; break;
  }
  return r ? 1.0 : 0.0;
}
WS_HD double tri_and(double a, double b) {
  const bool af = a == a && a == 0, bf = b == b && b == 0;
  if (af || bf) return 0.0;
  return (a != a || b != b) ? nullv() : 1.0;
}
WS_HD double tri_or(double a, double b) {
  const bool at = a == a && a != 0, bt = b == b && b != 0;
  if (at || bt) return 1.0;
  return (a != a || b != b) ? nullv() : 0.0;
}
WS_HD double tri_not(double a) { return a != a ? a : (a == 0 ? 1.0 : 0.0); }
WS_HD double first_non_null(double a, double b) { return a != a ? b : a; }
WS_HD double null_if_equal(double a, double b) { return a == b ? nullv() : a; }
WS_HD double div_or_null(double a, double b) { return b == 0 ? nullv() : a / b; }  // x / 0 = NULL
WS_HD i64 inset_i(const uint32_t* bits, int size, i64 c) {
  return (c >= 0 && c < size) ? (i64)((bits[c >> 5] >> (c & 31)) & 1u) : 0;
}
WS_HD double inset_d(const uint32_t* bits, int size, double x) {
  if (x != x) return x;
  return (double)inset_i(bits, size, (i64)x);
}
WS_HD double vdot_d(const float* v, i64 row, const double* q, int dim) {
  const float* x = v + row * (i64)dim;
  if (x[0] != x[0]) return nullv();
  float acc = 0;
  for (int i = 0; i < dim; ++i) acc += x[i] * (float)q[i];
  return acc;
}
WS_HD double veuc_d(const float* v, i64 row, const double* q, int dim) {
  const float* x = v + row * (i64)dim;
  if (x[0] != x[0]) return nullv();
  float acc = 0;
  for (int i = 0; i < dim; ++i) {
    const float d = x[i] - (float)q[i];
    acc += d * d;
  }
  return sqrt((double)acc);
}
WS_HD double insorted_d(const double* a, int n, double x) {
  if (x != x) return x;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && a[lo] == x) ? 1.0 : 0.0;
}
WS_HD double dictmap_d(const double* map, int size, double x) {
  if (x != x) return x;
  const i64 c = (i64)x;
  return (c >= 0 && c < size) ? map[c] : nullv();
}
WS_HD i64 year_of_days(i64 days) {
  i64 y;
  unsigned m, d;
  civil_from_days(days, y, m, d);
  return y;
}

// ---- scalar functions (devtypes.h:319-427)
constexpr i64 kUsPerDay = 86400000000ll;
WS_HD i64 floor_div(i64 a, i64 b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
WS_HD i64 add_months_days(i64 days, i64 n) {
  i64 y;
  unsigned m, d;
  civil_from_days(days, y, m, d);
  const i64 months = y * 12 + (i64)(m - 1) + n;
  y = months >= 0 ? months / 12 : (months - 11) / 12;
  m = (unsigned)(months - y * 12) + 1;
  const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
  unsigned lim = 31;
  if (m == 4 || m == 6 || m == 9 || m == 11) lim = 30;
  else if (m == 2) lim = leap ? 29 : 28;
  return days_from_civil(y, m, d < lim ? d : lim);
}
// MySQL WEEK(d, 0): weeks start on Sunday; days before the year's first Sunday are week 0.
WS_HD double week0(i64 x) {
  i64 y;
  unsigned m, d;
  civil_from_days(x, y, m, d);
  const i64 jan1 = days_from_civil(y, 1, 1);
  const i64 dow = (jan1 + 4) - floor_div(jan1 + 4, 7) * 7;  // 0 = Sunday
  const i64 first_sun = jan1 + (7 - dow) % 7;
  return x < first_sun ? 0.0 : (double)((x - first_sun) / 7 + 1);
}

WS_HD double call1(int fn, double x) {
  if (x != x) return x;
  i64 y;
  unsigned m, d;
  const i64 xi = (i64)x;
  switch (fn) {
    case WS_FN_ABS: return x < 0 ? -x : x;
    case WS_FN_CEIL: return ceil(x);
    case WS_FN_FLOOR: return floor(x);
    case WS_FN_ROUND: return round(x);
    case WS_FN_SQRT: return x < 0 ? nullv() : sqrt(x);
    case WS_FN_EXP: return exp(x);
    case WS_FN_LN: return x <= 0 ? nullv() : log(x);
    case WS_FN_LOG10: return x <= 0 ? nullv() : log10(x);
    case WS_FN_LOG2: return x <= 0 ? nullv() : log2(x);
    case WS_FN_SIN: return sin(x);
    case WS_FN_COS: return cos(x);
    case WS_FN_TAN: return tan(x);
    case WS_FN_ASIN: return asin(x);
    case WS_FN_ACOS: return acos(x);
    case WS_FN_ATAN: return atan(x);
    case WS_FN_DEGREES: return x * 57.29577951308232;
    case WS_FN_RADIANS: return x * 0.017453292519943295;
    case WS_FN_SIGN: return x > 0 ? 1.0 : x < 0 ? -1.0 : 0.0;
    case WS_FN_MONTH: civil_from_days(xi, y, m, d); return (double)m;
    case WS_FN_DAY: civil_from_days(xi, y, m, d); return (double)d;
    case WS_FN_DAYOFWEEK: return (double)(xi + 4 - floor_div(xi + 4, 7) * 7 + 1);  // 1 = Sunday
    case WS_FN_WEEKDAY: return (double)((xi + 3) - floor_div(xi + 3, 7) * 7);      // 0 = Monday
    case WS_FN_DAYOFYEAR: civil_from_days(xi, y, m, d); return (double)(xi - days_from_civil(y, 1, 1) + 1);
    case WS_FN_QUARTER: civil_from_days(xi, y, m, d); return (double)((m - 1) / 3 + 1);
    case WS_FN_WEEK: return week0(xi);
    case WS_FN_YEARWEEK: civil_from_days(xi, y, m, d); return (double)(y * 100) + week0(xi);
    case WS_FN_LAST_DAY: civil_from_days(xi, y, m, d); return (double)(add_months_days(days_from_civil(y, m, 1), 1) - 1);
    case WS_FN_TO_DAYS: return x + 719528;
    case WS_FN_FROM_DAYS: return x - 719528;
    case WS_FN_DT_DATE: return (double)floor_div(xi, kUsPerDay);
    case WS_FN_DT_HOUR: return (double)((xi - floor_div(xi, kUsPerDay) * kUsPerDay) / 3600000000ll);
    case WS_FN_DT_MINUTE: return (double)(((xi - floor_div(xi, kUsPerDay) * kUsPerDay) / 60000000ll) % 60);
    case WS_FN_DT_SECOND: return (double)(((xi - floor_div(xi, kUsPerDay) * kUsPerDay) / 1000000ll) % 60);
    case WS_FN_DT_MICRO: return (double)(xi - floor_div(xi, 1000000ll) * 1000000ll);
    case WS_FN_DATE_TO_DT: return x * (double)kUsPerDay;
    case WS_FN_DT_UNIX: return floor(x / 1e6);
    case WS_FN_UNIX_DT: return x * 1e6;
    case WS_FN_DT_YEAR: civil_from_days(floor_div(xi, kUsPerDay), y, m, d); return (double)y;
    case WS_FN_DT_MONTH: civil_from_days(floor_div(xi, kUsPerDay), y, m, d); return (double)m;
    case WS_FN_DT_DAY: civil_from_days(floor_div(xi, kUsPerDay), y, m, d); return (double)d;
    default: return nullv();
  }
}

WS_HD double call2(int fn, double a, double b) {
  if (a != a || b != b) return nullv();
  switch (fn) {
    case WS_FN_LEAST: return a < b ? a : b;
    case WS_FN_GREATEST: return a > b ? a : b;
    case WS_FN_ROUND2: {
      const double p = pow(10.0, b);
      return round(a * p) / p;
    }
    case WS_FN_TRUNCATE2: {
      const double p = pow(10.0, b);
      return trunc(a * p) / p;
    }
    case WS_FN_POW: return pow(a, b);
    case WS_FN_ATAN2: return atan2(a, b);
    case WS_FN_LOG_B: return (a <= 0 || b <= 0) ? nullv() : log(b) / log(a);
    case WS_FN_IDIV: return b == 0 ? nullv() : trunc(a / b);
    case WS_FN_ADD_MONTHS: return (double)add_months_days((i64)a, (i64)b);
    case WS_FN_DT_ADD_MONTHS: {
      const i64 us = (i64)a, days = floor_div(us, kUsPerDay);
      return (double)(add_months_days(days, (i64)b) * kUsPerDay + (us - days * kUsPerDay));
    }
    case WS_FN_DATE_TRUNC_MONTH: {
      i64 y;
      unsigned m, d;
      civil_from_days((i64)a, y, m, d);
      return (double)days_from_civil(y, b > 0 ? 1u : m, 1);  // b > 0: truncate to the year
    }
    default: return nullv();
  }
}

// ---- the stack VM: one program over one row (reference devtypes.h:430-495)
// kFull = false leaves out the scalar-function and vector opcodes (CALL1, CALL2, MOD, VDOT, VEUC):
// the math library they pull in sets the register budget of every kernel that can reach it, so
// plans without them run on a lean variant. Such an opcode then yields NULL; callers only pick the
// lean variant for programs that do not contain one (see prog_is_simple).
template <bool kFull>
WS_HD double eval_prog_t(const ws_expr_env& env, const ws_colview* cols, ws_prog pg, i64 row) {
  if (pg.len == 1) {  // a lone column or literal: no stack
    const ws_instr in = env.code[pg.start];
    if (in.op == WS_OP_COL) return load_col(cols[in.arg], row);
    if (in.op == WS_OP_CONST) return in.is_int ? (double)dbits(env.consts[in.arg]) : env.consts[in.arg];
  }
  double st[WS_MAX_STACK];
  int sp = 0;
  const int end = pg.start + pg.len;
  for (int pc = pg.start; pc < end; ++pc) {
    const ws_instr in = env.code[pc];
    switch (in.op) {
      case WS_OP_COL: st[sp++] = load_col(cols[in.arg], row); break;
      case WS_OP_COLRAW: st[sp++] = null_if_i64(col_int(cols[in.arg], row)); break;
      case WS_OP_CONST: st[sp++] = in.is_int ? (double)dbits(env.consts[in.arg]) : env.consts[in.arg]; break;
      case WS_OP_ADD: --sp; st[sp - 1] += st[sp]; break;
      case WS_OP_SUB: --sp; st[sp - 1] -= st[sp]; break;
      case WS_OP_MUL: --sp; st[sp - 1] *= st[sp]; break;
      case WS_OP_DIV: --sp; st[sp - 1] = div_or_null(st[sp - 1], st[sp]); break;
      case WS_OP_MOD: --sp; st[sp - 1] = kFull ? fmod(st[sp - 1], st[sp]) : nullv(); break;
      case WS_OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case WS_OP_EQ: case WS_OP_NE: case WS_OP_LT: case WS_OP_LE: case WS_OP_GT: case WS_OP_GE:
        --sp;
        st[sp - 1] = tri_cmp(in.op - WS_OP_EQ, st[sp - 1], st[sp]);
        break;
      case WS_OP_AND: --sp; st[sp - 1] = tri_and(st[sp - 1], st[sp]); break;
      case WS_OP_OR: --sp; st[sp - 1] = tri_or(st[sp - 1], st[sp]); break;
      case WS_OP_NOT: st[sp - 1] = tri_not(st[sp - 1]); break;
      case WS_OP_INSET: st[sp - 1] = inset_d(env.sets[in.arg], env.set_size[in.arg], st[sp - 1]); break;
      case WS_OP_SELECT:
        sp -= 2;
        st[sp - 1] = truthy(st[sp - 1]) ? st[sp] : st[sp + 1];
        break;
      case WS_OP_ISNULL: st[sp - 1] = st[sp - 1] != st[sp - 1] ? 1.0 : 0.0; break;
      case WS_OP_COALESCE: --sp; st[sp - 1] = first_non_null(st[sp - 1], st[sp]); break;
      case WS_OP_NULLIF: --sp; if (st[sp - 1] == st[sp]) st[sp - 1] = nullv(); break;
      case WS_OP_CALL1: st[sp - 1] = kFull ? call1(in.arg, st[sp - 1]) : nullv(); break;
      case WS_OP_CALL2: --sp; st[sp - 1] = kFull ? call2(in.arg, st[sp - 1], st[sp]) : nullv(); break;
      case WS_OP_DICTMAP: st[sp - 1] = dictmap_d(env.maps[in.arg], env.map_size[in.arg], st[sp - 1]); break;
      case WS_OP_INSORTED: st[sp - 1] = insorted_d(env.maps[in.arg], env.map_size[in.arg], st[sp - 1]); break;
      case WS_OP_VDOT: case WS_OP_VEUC: {
        const int slot = in.arg & 255, mi = in.arg >> 8;
        const float* v = (const float*)cols[slot].p;
        if (!kFull) st[sp++] = nullv();
        else st[sp++] = in.op == WS_OP_VDOT ? vdot_d(v, row, env.maps[mi], env.map_size[mi])
                                            : veuc_d(v, row, env.maps[mi], env.map_size[mi]);
        break;
      }
      case WS_OP_YEAR:
        if (st[sp - 1] == st[sp - 1]) st[sp - 1] = (double)year_of_days((i64)st[sp - 1]);
        break;
      default: break;
    }
  }
  return st[0];
}
WS_HD double eval_prog(const ws_expr_env& env, const ws_colview* cols, ws_prog pg, i64 row) {
  return eval_prog_t<true>(env, cols, pg, row);
}
// True when the program uses none of the opcodes the lean VM leaves out.
WS_HD bool prog_is_simple(const ws_expr_env& env, ws_prog pg) {
  for (int pc = pg.start; pc < pg.start + pg.len; ++pc) {
    const int op = env.code[pc].op;
    if (op == WS_OP_MOD || op == WS_OP_CALL1 || op == WS_OP_CALL2 || op == WS_OP_VDOT || op == WS_OP_VEUC) return false;
  }
  return true;
}

}  // namespace ws
