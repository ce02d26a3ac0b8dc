// CPU oracle of the analytic hot path. TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke
// check and as the benchmark's CPU baseline; the product path never calls it.
//
// A plain, independent restatement of the reference's semantics (not of this repository's GPU
// code), written for clarity over speed:
//   - bytecode VM, NULL / three-valued logic, checked int64, dates, scalar functions, encoded
//     column decode .................................. reference src/devtypes.h:27-495
//   - aggregation slot semantics (identities, NULL-skipping combine, int_part / split BIGINT sum,
//     agg_bits int64 results, 128-bit no-keys sums, dense accumulator kinds)
//                                                     src/kernels.cuh:35-82, 188-633
//   - GROUP BY identity = the 64-bit key hash; here groups are keyed by key values instead, which
//     differs only on a 64-bit hash collision (noted in the design notes)
//   - row selection = ascending global row indices of rows passing WHERE and not deleted
//                                                     src/gpu.cu:664-755, kernels.cuh:636-670
//   - inner / outer equi-join pair sets (NULL never matches)  src/gpu.cu:1089-1293
//   - stable multi-key ORDER BY permutation (NULL -> -inf, -0 -> 0)  src/gpu.cu:1318-1356
//   - exclusive sums of integer SUM halves recombined in 128 bits   src/engine_core.h:1972-1983
//   - column statistics for sizing (no reference counterpart): live / NULL counts, min / max and
//     HyperLogLog registers (Flajolet et al. 2007; p = 11, splitmix64 finalizer over the value's
//     bits: exact int64 for integer types, bits of v + 0.0 for doubles)
// Pinned by: golden vectors from the reference's own devtypes.h compiled on the host (the VM,
// decoders, dates, checked arithmetic), brute force and closed forms in tests/test_oracle.py.
#include <algorithm>
#include <array>
#include <set>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "warpstore_amd/ws_abi.h"

namespace {

typedef long long i64;
typedef unsigned long long u64;
const i64 I64MAX = 0x7fffffffffffffffll, I64MIN = -I64MAX - 1;

double NaN() {
  u64 b = 0x7ff8000000000000ull;
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}
i64 bits_of(double d) {
  i64 v;
  std::memcpy(&v, &d, 8);
  return v;
}
double from_bits(i64 v) {
  double d;
  std::memcpy(&d, &v, 8);
  return d;
}

// Double -> int64 truncation as x86-64 performs it (cvttsd2si: out of range or NaN -> INT64_MIN).
// C++ leaves the out-of-range case undefined and other processors answer differently, so every
// such conversion is recorded in g_out_of_range: a result that depended on one is not portable.
thread_local int g_out_of_range = 0;
i64 to_i64(double x) {
  if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) {
    g_out_of_range = 1;
    return I64MIN;
  }
  return (i64)x;
}

// Two's-complement wrapping int64 arithmetic: date and datetime arithmetic on out-of-range inputs
// wraps on the processors the engine runs on; spelled out so the restatement has no signed overflow.
i64 wadd(i64 a, i64 b) { return (i64)((u64)a + (u64)b); }
i64 wsub(i64 a, i64 b) { return (i64)((u64)a - (u64)b); }
i64 wmul(i64 a, i64 b) { return (i64)((u64)a * (u64)b); }

// ------------------------------------------------------------ dates (Hinnant, public domain)
i64 civil_to_days(i64 y, unsigned m, unsigned d) {
  if (m <= 2) y = wsub(y, 1);
  i64 era = (y >= 0 ? y : wsub(y, 399)) / 400;
  unsigned yoe = (unsigned)wsub(y, wmul(era, 400));
  unsigned doy = (153 * (m > 2 ? m - 3 : m + 9) + 2) / 5 + d - 1;
  unsigned doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return wsub(wadd(wmul(era, 146097), (i64)doe), 719468);
}
void days_to_civil(i64 z, i64& y, unsigned& m, unsigned& d) {
  z = wadd(z, 719468);
  i64 era = (z >= 0 ? z : wsub(z, 146096)) / 146097;
  unsigned doe = (unsigned)wsub(z, wmul(era, 146097));
  unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  y = wadd((i64)yoe, wmul(era, 400));
  unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  unsigned mp = (5 * doy + 2) / 153;
  d = doy - (153 * mp + 2) / 5 + 1;
  m = mp < 10 ? mp + 3 : mp - 9;
  if (m <= 2) y = wadd(y, 1);
}
// Floor division for b > 0, as -((-a + b - 1) / b) for negative a, with the negations and the sum
// wrapping in two's complement (the value x86-64 gives for a = INT64_MIN, without signed overflow).
i64 fdiv(i64 a, i64 b) {
  if (a >= 0) return a / b;
  const i64 t = (i64)((0ull - (u64)a) + (u64)(b - 1));
  return (i64)(0ull - (u64)(t / b));
}
i64 add_months(i64 days, i64 n) {
  i64 y;
  unsigned m, d;
  days_to_civil(days, y, m, d);
  i64 months = wadd(wadd(wmul(y, 12), (i64)(m - 1)), n);
  y = months >= 0 ? months / 12 : wsub(months, 11) / 12;
  m = (unsigned)wsub(months, wmul(y, 12)) + 1;
  static const unsigned len[] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
  unsigned lim = len[m - 1] + (m == 2 && leap ? 1 : 0);
  return civil_to_days(y, m, std::min(d, lim));
}
double mysql_week(i64 x) {
  i64 y;
  unsigned m, d;
  days_to_civil(x, y, m, d);
  i64 jan1 = civil_to_days(y, 1, 1);
  i64 dow = wsub(wadd(jan1, 4), wmul(fdiv(wadd(jan1, 4), 7), 7));
  i64 sun = wadd(jan1, (7 - dow) % 7);
  return x < sun ? 0.0 : (double)(wsub(x, sun) / 7 + 1);
}

// ------------------------------------------------------------ column decode
i64 stored_int(const ws_colview& c, i64 r) {
  if (c.enc) {
    const unsigned char* base = (const unsigned char*)c.p;
    i64 b;
    std::memcpy(&b, base, 8);
    const unsigned char* data = base + WS_ENC_HEADER;
    i64 i = r + c.off;
    bool narrow = c.t == WS_T_I32 || c.t == WS_T_I32N;
    i64 nul = narrow ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
    u64 code, top;
    if (c.enc == 1) code = data[i], top = 0xFF;
    else if (c.enc == 2) code = ((const uint16_t*)data)[i], top = 0xFFFF;
    else code = ((const uint32_t*)data)[i], top = 0xFFFFFFFFull;
    return code == top ? nul : b + (i64)code;
  }
  if (c.t == WS_T_I32 || c.t == WS_T_I32N) return ((const int32_t*)c.p)[r];
  return ((const int64_t*)c.p)[r];
}
double stored_dbl(const ws_colview& c, i64 r) {
  if (!c.enc) return ((const double*)c.p)[r];
  const unsigned char* base = (const unsigned char*)c.p;
  const unsigned char* data = base + WS_ENC_HEADER;
  if (c.enc == 4) return ((const double*)data)[data[2048 + r + c.off]];
  i64 b;
  double p10;
  std::memcpy(&b, base, 8);
  std::memcpy(&p10, base + 8, 8);
  i64 i = r + c.off;
  if (c.enc == 5) {
    unsigned v = ((const uint16_t*)data)[i];
    return v == 0xFFFF ? NaN() : (double)(b + v) / p10;
  }
  unsigned v = ((const uint32_t*)data)[i];
  return v == 0xFFFFFFFFu ? NaN() : (double)(b + (i64)v) / p10;
}
double p10(int s) {
  double p = 1;
  while (s-- > 0) p *= 10;
  return p;
}
double value(const ws_colview& c, i64 r) {
  if (c.t == WS_T_F64) return stored_dbl(c, r);
  if (c.t == WS_T_VEC) return NaN();
  i64 v = stored_int(c, r);
  switch (c.t) {
    case WS_T_I32N: return v == (i64)WS_NULL_I32 ? NaN() : (double)v;
    case WS_T_I64N: return v == (i64)WS_NULL_I64 ? NaN() : (double)v;
    case WS_T_DECIMAL: return v == (i64)WS_NULL_I64 ? NaN() : (double)v / p10(c.scale);
    default: return (double)v;
  }
}

// ------------------------------------------------------------ scalar functions
double fn1(int f, double x) {
  if (std::isnan(x)) return x;
  i64 y;
  unsigned m, d;
  i64 xi;
  switch (f) {
    case WS_FN_ABS: return std::fabs(x) == x ? x : -x;
    case WS_FN_CEIL: return std::ceil(x);
    case WS_FN_FLOOR: return std::floor(x);
    case WS_FN_ROUND: return std::round(x);
    case WS_FN_SQRT: return x < 0 ? NaN() : std::sqrt(x);
    case WS_FN_EXP: return std::exp(x);
    case WS_FN_LN: return x <= 0 ? NaN() : std::log(x);
    case WS_FN_LOG10: return x <= 0 ? NaN() : std::log10(x);
    case WS_FN_LOG2: return x <= 0 ? NaN() : std::log2(x);
    case WS_FN_SIN: return std::sin(x);
    case WS_FN_COS: return std::cos(x);
    case WS_FN_TAN: return std::tan(x);
    case WS_FN_ASIN: return std::asin(x);
    case WS_FN_ACOS: return std::acos(x);
    case WS_FN_ATAN: return std::atan(x);
    case WS_FN_DEGREES: return x * 57.29577951308232;
    case WS_FN_RADIANS: return x * 0.017453292519943295;
    case WS_FN_SIGN: return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0);
    default: break;
  }
  xi = to_i64(x);
  switch (f) {
    case WS_FN_MONTH: days_to_civil(xi, y, m, d); return m;
    case WS_FN_DAY: days_to_civil(xi, y, m, d); return d;
    case WS_FN_DAYOFWEEK: return (double)(wsub(wadd(xi, 4), wmul(fdiv(wadd(xi, 4), 7), 7)) + 1);
    case WS_FN_WEEKDAY: return (double)wsub(wadd(xi, 3), wmul(fdiv(wadd(xi, 3), 7), 7));
    case WS_FN_DAYOFYEAR: days_to_civil(xi, y, m, d); return (double)wadd(wsub(xi, civil_to_days(y, 1, 1)), 1);
    case WS_FN_QUARTER: days_to_civil(xi, y, m, d); return (double)((m - 1) / 3 + 1);
    case WS_FN_WEEK: return mysql_week(xi);
    case WS_FN_YEARWEEK: days_to_civil(xi, y, m, d); return (double)wmul(y, 100) + mysql_week(xi);
    case WS_FN_LAST_DAY: days_to_civil(xi, y, m, d); return (double)wsub(add_months(civil_to_days(y, m, 1), 1), 1);
    case WS_FN_TO_DAYS: return x + 719528;
    case WS_FN_FROM_DAYS: return x - 719528;
    case WS_FN_DT_DATE: return (double)fdiv(xi, 86400000000ll);
    case WS_FN_DT_HOUR: return (double)(wsub(xi, wmul(fdiv(xi, 86400000000ll), 86400000000ll)) / 3600000000ll);
    case WS_FN_DT_MINUTE: return (double)((wsub(xi, wmul(fdiv(xi, 86400000000ll), 86400000000ll)) / 60000000ll) % 60);
    case WS_FN_DT_SECOND: return (double)((wsub(xi, wmul(fdiv(xi, 86400000000ll), 86400000000ll)) / 1000000ll) % 60);
    case WS_FN_DT_MICRO: return (double)wsub(xi, wmul(fdiv(xi, 1000000ll), 1000000ll));
    case WS_FN_DATE_TO_DT: return x * 86400000000.0;
    case WS_FN_DT_UNIX: return std::floor(x / 1e6);
    case WS_FN_UNIX_DT: return x * 1e6;
    case WS_FN_DT_YEAR: days_to_civil(fdiv(xi, 86400000000ll), y, m, d); return (double)y;
    case WS_FN_DT_MONTH: days_to_civil(fdiv(xi, 86400000000ll), y, m, d); return m;
    case WS_FN_DT_DAY: days_to_civil(fdiv(xi, 86400000000ll), y, m, d); return d;
  }
  return NaN();
}
double fn2(int f, double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NaN();
  switch (f) {
    case WS_FN_LEAST: return a < b ? a : b;
    case WS_FN_GREATEST: return a > b ? a : b;
    case WS_FN_ROUND2: return std::round(a * std::pow(10.0, b)) / std::pow(10.0, b);
    case WS_FN_TRUNCATE2: return std::trunc(a * std::pow(10.0, b)) / std::pow(10.0, b);
    case WS_FN_POW: return std::pow(a, b);
    case WS_FN_ATAN2: return std::atan2(a, b);
    case WS_FN_LOG_B: return (a <= 0 || b <= 0) ? NaN() : std::log(b) / std::log(a);
    case WS_FN_IDIV: return b == 0 ? NaN() : std::trunc(a / b);
    case WS_FN_ADD_MONTHS: return (double)add_months(to_i64(a), to_i64(b));
    case WS_FN_DT_ADD_MONTHS: {
      i64 us = to_i64(a), days = fdiv(us, 86400000000ll);
      return (double)wadd(wmul(add_months(days, to_i64(b)), 86400000000ll), wsub(us, wmul(days, 86400000000ll)));
    }
    case WS_FN_DATE_TRUNC_MONTH: {
      i64 y;
      unsigned m, d;
      days_to_civil(to_i64(a), y, m, d);
      return (double)civil_to_days(y, b > 0 ? 1 : m, 1);
    }
  }
  return NaN();
}

// ------------------------------------------------------------ VM
double tv(bool b) { return b ? 1.0 : 0.0; }
double run(const ws_expr_env& env, const ws_colview* cols, ws_prog pg, i64 row) {
  std::vector<double> st;
  st.reserve(16);
  auto pop = [&] {
    double v = st.back();
    st.pop_back();
    return v;
  };
  for (int pc = pg.start; pc < pg.start + pg.len; ++pc) {
    const ws_instr in = env.code[pc];
    switch (in.op) {
      case WS_OP_COL: st.push_back(value(cols[in.arg], row)); break;
      case WS_OP_COLRAW: {
        i64 v = stored_int(cols[in.arg], row);
        st.push_back(v == (i64)WS_NULL_I64 ? NaN() : (double)v);
        break;
      }
      case WS_OP_CONST: st.push_back(in.is_int ? (double)bits_of(env.consts[in.arg]) : env.consts[in.arg]); break;
      case WS_OP_ADD: { double b = pop(), a = pop(); st.push_back(a + b); break; }
      case WS_OP_SUB: { double b = pop(), a = pop(); st.push_back(a - b); break; }
      case WS_OP_MUL: { double b = pop(), a = pop(); st.push_back(a * b); break; }
      case WS_OP_DIV: { double b = pop(), a = pop(); st.push_back(b == 0 ? NaN() : a / b); break; }
      case WS_OP_MOD: { double b = pop(), a = pop(); st.push_back(std::fmod(a, b)); break; }
      case WS_OP_NEG: st.back() = -st.back(); break;
      case WS_OP_EQ: case WS_OP_NE: case WS_OP_LT: case WS_OP_LE: case WS_OP_GT: case WS_OP_GE: {
        double b = pop(), a = pop();
        if (std::isnan(a) || std::isnan(b)) { st.push_back(NaN()); break; }
        bool r = in.op == WS_OP_EQ ? a == b : in.op == WS_OP_NE ? a != b : in.op == WS_OP_LT ? a < b
               : in.op == WS_OP_LE ? a <= b : in.op == WS_OP_GT ? a > b : a >= b;
        st.push_back(tv(r));
        break;
      }
      case WS_OP_AND: {
        double b = pop(), a = pop();
        if ((!std::isnan(a) && a == 0) || (!std::isnan(b) && b == 0)) st.push_back(0.0);
        else st.push_back(std::isnan(a) || std::isnan(b) ? NaN() : 1.0);
        break;
      }
      case WS_OP_OR: {
        double b = pop(), a = pop();
        if ((!std::isnan(a) && a != 0) || (!std::isnan(b) && b != 0)) st.push_back(1.0);
        else st.push_back(std::isnan(a) || std::isnan(b) ? NaN() : 0.0);
        break;
      }
      case WS_OP_NOT: { double a = pop(); st.push_back(std::isnan(a) ? a : tv(a == 0)); break; }
      case WS_OP_INSET: {
        double x = pop();
        if (std::isnan(x)) { st.push_back(x); break; }
        i64 c = to_i64(x);
        const uint32_t* bits = env.sets[in.arg];
        st.push_back(c >= 0 && c < env.set_size[in.arg] ? (double)((bits[c >> 5] >> (c & 31)) & 1u) : 0.0);
        break;
      }
      case WS_OP_SELECT: {
        double e = pop(), t = pop(), c = pop();
        st.push_back(!std::isnan(c) && c != 0 ? t : e);
        break;
      }
      case WS_OP_YEAR: {
        double x = pop();
        if (std::isnan(x)) { st.push_back(x); break; }
        i64 y;
        unsigned m, d;
        days_to_civil(to_i64(x), y, m, d);
        st.push_back((double)y);
        break;
      }
      case WS_OP_ISNULL: { double a = pop(); st.push_back(tv(std::isnan(a))); break; }
      case WS_OP_COALESCE: { double b = pop(), a = pop(); st.push_back(std::isnan(a) ? b : a); break; }
      case WS_OP_NULLIF: { double b = pop(), a = pop(); st.push_back(a == b ? NaN() : a); break; }
      case WS_OP_CALL1: { double a = pop(); st.push_back(fn1(in.arg, a)); break; }
      case WS_OP_CALL2: { double b = pop(), a = pop(); st.push_back(fn2(in.arg, a, b)); break; }
      case WS_OP_DICTMAP: {
        double x = pop();
        if (std::isnan(x)) { st.push_back(x); break; }
        i64 c = to_i64(x);
        st.push_back(c >= 0 && c < env.map_size[in.arg] ? env.maps[in.arg][c] : NaN());
        break;
      }
      case WS_OP_INSORTED: {
        double x = pop();
        if (std::isnan(x)) { st.push_back(x); break; }
        // lower bound by '<', then an exact '==' (a NaN entry never matches)
        const double* a = env.maps[in.arg];
        const double* e = a + env.map_size[in.arg];
        const double* it = std::lower_bound(a, e, x);
        st.push_back(it != e && *it == x ? 1.0 : 0.0);
        break;
      }
      case WS_OP_VDOT: case WS_OP_VEUC: {
        int slot = in.arg & 255, mi = in.arg >> 8, dim = env.map_size[mi];
        const float* x = (const float*)cols[slot].p + row * (i64)dim;
        const double* q = env.maps[mi];
        if (std::isnan(x[0])) { st.push_back(NaN()); break; }
        float acc = 0;
        for (int i = 0; i < dim; ++i) {
          if (in.op == WS_OP_VDOT) acc += x[i] * (float)q[i];
          else { float dd = x[i] - (float)q[i]; acc += dd * dd; }
        }
        st.push_back(in.op == WS_OP_VDOT ? (double)acc : std::sqrt((double)acc));
        break;
      }
    }
  }
  return st.empty() ? NaN() : st[0];
}

// ------------------------------------------------------------ compiled-plan (typed) evaluation
// The per-query compiler types every subexpression as exact int64 or as double (NULL = NaN) and
// evaluates int64 nodes in integer arithmetic (reference src/jit.cu:107-272): integer columns and
// literals are int64; + - * of two int64 wrap (and raise the overflow flag); comparisons, AND /
// OR / NOT, IN-set, CASE, YEAR and COALESCE of int64 operands stay int64; a DECIMAL's raw stored
// integer is int64 unless read as a nullable value; YEAR of a double goes through the datetime
// functions (days -> microseconds -> year). Everything else is the VM's double operator. On most
// rows this equals the VM's value; it differs where double arithmetic cannot be exact (int64
// operands or results beyond 2^53, wrapped overflow), where a raw NULL sentinel is read as an
// integer, or where YEAR's datetime path truncates differently. Each node checks its typed value
// against the VM operator on the same operands and classifies a difference:
//   kTypedExplained:   the difference has one of the causes above (checked from the operands);
//   kTypedUnexplained: a difference without such a cause;
//   kTypedNotPortable: the value went through an out-of-range double -> int64 conversion.
enum { kTypedExplained = 1, kTypedUnexplained = 2, kTypedNotPortable = 4 };
struct TV {
  bool is_int;
  i64 i;
  double d;
  double as_d() const { return is_int ? (double)i : d; }
};
constexpr double kExactLimit = 9007199254740992.0;  // 2^53
bool beyond_exact(i64 v) { return v > (i64)kExactLimit || v < -(i64)kExactLimit; }
bool same_double(double a, double b) { return (std::isnan(a) && std::isnan(b)) || bits_of(a) == bits_of(b); }

TV typed_run(const ws_expr_env& env, const ws_colview* cols, ws_prog pg, i64 row, int* flags) {
  std::vector<TV> st;
  auto pop = [&] {
    TV v = st.back();
    st.pop_back();
    return v;
  };
  auto I = [](i64 v) { return TV{true, v, 0.0}; };
  auto D = [](double v) { return TV{false, 0, v}; };
  auto differ = [&](double typed, double vm, bool cause) {
    if (!same_double(typed, vm)) *flags |= cause ? kTypedExplained : kTypedUnexplained;
  };
  for (int pc = pg.start; pc < pg.start + pg.len; ++pc) {
    const ws_instr in = env.code[pc];
    switch (in.op) {
      case WS_OP_COL: {
        const ws_colview& c = cols[in.arg];
        if (c.t == WS_T_I32 || c.t == WS_T_I64) st.push_back(I(stored_int(c, row)));
        else st.push_back(D(value(c, row)));
        break;
      }
      case WS_OP_COLRAW: {
        const ws_colview& c = cols[in.arg];
        const i64 v = stored_int(c, row);
        if (c.t == WS_T_DECIMAL && !in.is_int) {
          st.push_back(D(v == (i64)WS_NULL_I64 ? NaN() : (double)v));
        } else {
          st.push_back(I(v));
          differ((double)v, v == (i64)WS_NULL_I64 ? NaN() : (double)v, v == (i64)WS_NULL_I64);
        }
        break;
      }
      case WS_OP_CONST: st.push_back(in.is_int ? I(bits_of(env.consts[in.arg])) : D(env.consts[in.arg])); break;
      case WS_OP_ADD: case WS_OP_SUB: case WS_OP_MUL: {
        TV b = pop(), a = pop();
        const double x = a.as_d(), y = b.as_d();
        const double vm = in.op == WS_OP_ADD ? x + y : in.op == WS_OP_SUB ? x - y : x * y;
        if (a.is_int && b.is_int) {
          const __int128 exact = in.op == WS_OP_ADD ? (__int128)a.i + b.i : in.op == WS_OP_SUB ? (__int128)a.i - b.i
                                                                                              : (__int128)a.i * b.i;
          const i64 r = (i64)(u64)(unsigned __int128)exact;
          const bool inexact = beyond_exact(a.i) || beyond_exact(b.i) || exact > (__int128)kExactLimit ||
                               exact < -(__int128)kExactLimit;
          // an integer zero has no sign: 0 * -5 is +0 in int64 and -0.0 in doubles
          differ((double)r, vm, inexact || (r == 0 && vm == 0));
          st.push_back(I(r));
        } else {
          st.push_back(D(vm));
        }
        break;
      }
      case WS_OP_DIV: { TV b = pop(), a = pop(); st.push_back(D(b.as_d() == 0 ? NaN() : a.as_d() / b.as_d())); break; }
      case WS_OP_MOD: { TV b = pop(), a = pop(); st.push_back(D(std::fmod(a.as_d(), b.as_d()))); break; }
      case WS_OP_NEG: {
        TV a = pop();
        if (a.is_int) {
          const i64 r = (i64)(0ull - (u64)a.i);
          differ((double)r, -(double)a.i, a.i == I64MIN || a.i == 0);  // wraps; -0 is +0 in int64
          st.push_back(I(r));
        } else {
          st.push_back(D(-a.d));
        }
        break;
      }
      case WS_OP_EQ: case WS_OP_NE: case WS_OP_LT: case WS_OP_LE: case WS_OP_GT: case WS_OP_GE: {
        TV b = pop(), a = pop();
        auto cmp = [&](auto x, auto y) {
          return in.op == WS_OP_EQ ? x == y : in.op == WS_OP_NE ? x != y : in.op == WS_OP_LT ? x < y
               : in.op == WS_OP_LE ? x <= y : in.op == WS_OP_GT ? x > y : x >= y;
        };
        const double x = a.as_d(), y = b.as_d();
        const double vm = std::isnan(x) || std::isnan(y) ? NaN() : tv(cmp(x, y));
        if (a.is_int && b.is_int) {
          const i64 r = cmp(a.i, b.i) ? 1 : 0;
          differ((double)r, vm, beyond_exact(a.i) || beyond_exact(b.i));
          st.push_back(I(r));
        } else {
          st.push_back(D(vm));
        }
        break;
      }
      case WS_OP_AND: case WS_OP_OR: {
        TV b = pop(), a = pop();
        if (a.is_int && b.is_int) {
          st.push_back(I(in.op == WS_OP_AND ? (a.i != 0 && b.i != 0) : (a.i != 0 || b.i != 0)));
          break;
        }
        const double x = a.as_d(), y = b.as_d();
        if (in.op == WS_OP_AND) {
          if ((!std::isnan(x) && x == 0) || (!std::isnan(y) && y == 0)) st.push_back(D(0.0));
          else st.push_back(D(std::isnan(x) || std::isnan(y) ? NaN() : 1.0));
        } else {
          if ((!std::isnan(x) && x != 0) || (!std::isnan(y) && y != 0)) st.push_back(D(1.0));
          else st.push_back(D(std::isnan(x) || std::isnan(y) ? NaN() : 0.0));
        }
        break;
      }
      case WS_OP_NOT: {
        TV a = pop();
        st.push_back(a.is_int ? I(a.i == 0) : D(std::isnan(a.d) ? a.d : tv(a.d == 0)));
        break;
      }
      case WS_OP_INSET: {
        TV a = pop();
        const uint32_t* bits = env.sets[in.arg];
        const i64 n = env.set_size[in.arg];
        auto bit = [&](i64 c) { return c >= 0 && c < n ? (i64)((bits[c >> 5] >> (c & 31)) & 1u) : 0; };
        if (a.is_int) {
          const i64 r = bit(a.i);
          const int saved = g_out_of_range;
          differ((double)r, (double)bit(to_i64((double)a.i)), beyond_exact(a.i));
          g_out_of_range = saved;
          st.push_back(I(r));
        } else {
          st.push_back(D(std::isnan(a.d) ? a.d : (double)bit(to_i64(a.d))));
        }
        break;
      }
      case WS_OP_SELECT: {
        TV e = pop(), t = pop(), c = pop();
        const bool cond = c.is_int ? c.i != 0 : (!std::isnan(c.d) && c.d != 0);
        if (t.is_int && e.is_int) st.push_back(cond ? t : e);
        else st.push_back(D(cond ? t.as_d() : e.as_d()));
        break;
      }
      case WS_OP_YEAR: {
        TV a = pop();
        i64 y;
        unsigned m, d;
        if (a.is_int) {
          days_to_civil(a.i, y, m, d);
          const i64 r = y;
          const int saved = g_out_of_range;
          days_to_civil(to_i64((double)a.i), y, m, d);
          g_out_of_range = saved;
          differ((double)r, (double)y, beyond_exact(a.i));
          st.push_back(I(r));
        } else if (std::isnan(a.d)) {
          st.push_back(a);
        } else {
          const int saved = g_out_of_range;
          g_out_of_range = 0;
          const double r = fn1(WS_FN_DT_YEAR, fn1(WS_FN_DATE_TO_DT, a.d));
          const bool out_of_range = g_out_of_range;
          i64 vm_year = 0;
          if (a.d >= -9223372036854775808.0 && a.d < 9223372036854775808.0) {
            days_to_civil((i64)a.d, y, m, d);
            vm_year = y;
          }
          g_out_of_range = saved || out_of_range;
          differ(r, (double)vm_year, out_of_range || a.d != std::trunc(a.d));
          st.push_back(D(r));
        }
        break;
      }
      case WS_OP_ISNULL: { TV a = pop(); st.push_back(I(a.is_int ? 0 : std::isnan(a.d))); break; }
      case WS_OP_COALESCE: {
        TV b = pop(), a = pop();
        if (a.is_int) st.push_back(a);
        else st.push_back(D(std::isnan(a.d) ? b.as_d() : a.d));
        break;
      }
      case WS_OP_NULLIF: { TV b = pop(), a = pop(); st.push_back(D(a.as_d() == b.as_d() ? NaN() : a.as_d())); break; }
      case WS_OP_CALL1: { TV a = pop(); st.push_back(D(fn1(in.arg, a.as_d()))); break; }
      case WS_OP_CALL2: { TV b = pop(), a = pop(); st.push_back(D(fn2(in.arg, a.as_d(), b.as_d()))); break; }
      default: {
        // DICTMAP, INSORTED, VDOT, VEUC: double operators on double operands, as in the VM
        ws_expr_env one = env;
        const int nin = (in.op == WS_OP_VDOT || in.op == WS_OP_VEUC) ? 0 : 1;
        double arg = 0;
        if (nin) arg = pop().as_d();
        one.consts[0] = arg;
        one.code[0] = ws_instr{WS_OP_CONST, 0, 0};
        one.code[1] = in;
        st.push_back(D(run(one, cols, ws_prog{(int16_t)(nin ? 0 : 1), (int16_t)(nin + 1)}, row)));
        break;
      }
    }
  }
  if (g_out_of_range) *flags |= kTypedNotPortable;
  return st.empty() ? TV{false, 0, NaN()} : st[0];
}

bool is_col(const ws_expr_env& env, ws_prog p, int* slot) {
  if (p.len == 1 && env.code[p.start].op == WS_OP_COL) {
    *slot = env.code[p.start].arg;
    return true;
  }
  return false;
}
// Integer value of a key / aggregate argument: a plain integer column is read exactly, anything
// else goes through the double VM and truncates (reference kernels.cuh:122-136).
i64 int_value(const ws_expr_env& env, const ws_colview* cols, ws_prog p, i64 r) {
  int s;
  if (is_col(env, p, &s) && cols[s].t != WS_T_F64) return stored_int(cols[s], r);
  return to_i64(run(env, cols, p, r));
}

// ------------------------------------------------------------ aggregation
struct Slot {
  __int128 isum = 0;  // integer sums
  i64 imm = 0;        // integer min / max (int64 combine)
  bool imm_set = false;
  long double fsum = 0;  // double sums (extended precision)
  double fmm = 0;        // double min / max
  i64 count = 0;
};
struct Group {
  std::vector<double> keys;
  std::vector<Slot> slots;
  i64 rows = 0;
};

i64 int_part(const ws_agg_plan& p, int a, i64 v) {
  if (p.agg_null[a]) {
    i64 nul = p.agg_null[a] == WS_T_I32N ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
    if (v == nul) return p.ops[a] == WS_AGG_MIN ? I64MAX : p.ops[a] == WS_AGG_MAX ? I64MIN : 0;
  }
  if (!p.agg_bits[a]) return v;
  if (p.agg_part[a] == 1) return v & 0xffffffffll;
  if (p.agg_part[a] == 2) return v >> 32;
  return v;
}

}  // namespace

extern "C" {

double or_eval_prog(const ws_expr_env* env, const ws_colview* cols, ws_prog p, int64_t row) {
  return run(*env, cols ? cols : env->cols, p, row);
}
// A program's value as a compiled plan computes it (int64 nodes in integer arithmetic, see
// typed_run); *flags gets kTypedExplained / kTypedUnexplained / kTypedNotPortable.
double or_eval_typed(const ws_expr_env* env, const ws_colview* cols, ws_prog p, int64_t row, int* flags) {
  g_out_of_range = 0;
  *flags = 0;
  const double v = typed_run(*env, cols ? cols : env->cols, p, row, flags).as_d();
  g_out_of_range = 0;
  return v;
}
int64_t or_stored_int(const ws_colview* c, int64_t r) { return stored_int(*c, r); }
double or_stored_dbl(const ws_colview* c, int64_t r) { return stored_dbl(*c, r); }
double or_value(const ws_colview* c, int64_t r) { return value(*c, r); }
double or_call1(int f, double x) { return fn1(f, x); }
double or_call2(int f, double a, double b) { return fn2(f, a, b); }
int64_t or_days_from_civil(int64_t y, unsigned m, unsigned d) { return civil_to_days(y, m, d); }
void or_civil_from_days(int64_t z, int64_t* y, unsigned* m, unsigned* d) {
  i64 yy;
  days_to_civil(z, yy, *m, *d);
  *y = yy;
}
// Checked int64 arithmetic: the wrapped result, *ovf = 1 on signed overflow.
int64_t or_checked(int op, int64_t a, int64_t b, int* ovf) {
  __int128 exact = op == 0 ? (__int128)a + b : op == 1 ? (__int128)a - b : (__int128)a * b;
  i64 r = (i64)(u64)(unsigned __int128)exact;
  if ((__int128)r != exact) *ovf = 1;
  return r;
}

// Filter + aggregate over batches. Batch b's columns are cols[b * ncols ...] (host memory), rows
// n[b], deleted-row bitmap del[b] (may be null). dense != 0: the dense-kernel slot semantics (int
// MIN / MAX over integer kinds without agg_bits become (double) of the int64 combine).
// Output: groups in ascending order of their key bit patterns (row-major keys / accs); returns
// the number of groups, or -needed when max_groups is too small. Without GROUP BY: one group,
// accs[a] the low word or value, hi[a] the high word of 128-bit integer sums.
int64_t or_aggregate(const ws_agg_plan* planp, int dense, const ws_expr_env* env, int nb, const int64_t* n,
                     const uint32_t* const* del, const ws_colview* cols, int ncols, double* keys_out, double* accs_out,
                     int64_t* hi_out, int64_t max_groups) {
  const ws_agg_plan& plan = *planp;
  const int NK = plan.nkeys, NA = plan.naggs;
  std::map<std::vector<i64>, Group> groups;
  auto new_group = [&](const std::vector<double>& kv) {
    Group g;
    g.keys = kv;
    g.slots.resize(NA);
    return g;
  };
  Group nokeys = new_group({});
  for (int b = 0; b < nb; ++b) {
    const ws_colview* bc = cols + (size_t)b * ncols;
    for (i64 r = 0; r < n[b]; ++r) {
      if (del && del[b] && ((del[b][r >> 5] >> (r & 31)) & 1u)) continue;
      if (plan.has_where) {
        double w = run(*env, bc, plan.where, r);
        if (std::isnan(w) || w == 0) continue;
      }
      Group* g = &nokeys;
      if (NK) {
        std::vector<double> kv(NK);
        std::vector<i64> kb(NK);
        for (int k = 0; k < NK; ++k) {
          kv[k] = plan.key_int[k] ? from_bits(int_value(*env, bc, plan.keys[k], r)) : run(*env, bc, plan.keys[k], r) + 0.0;
          kb[k] = bits_of(kv[k]);
        }
        auto it = groups.find(kb);
        if (it == groups.end()) it = groups.emplace(kb, new_group(kv)).first;
        g = &it->second;
      }
      g->rows++;
      for (int a = 0; a < NA; ++a) {
        Slot& s = g->slots[a];
        const int op = plan.ops[a];
        if (op == WS_AGG_COUNT) continue;
        if (plan.agg_int[a] && (plan.agg_bits[a] || dense || !NK)) {
          i64 v = int_part(plan, a, int_value(*env, bc, plan.aggs[a], r));
          if (op == WS_AGG_SUM) s.isum += v;
          else if (!s.imm_set) s.imm = v, s.imm_set = true;
          else s.imm = op == WS_AGG_MIN ? std::min(s.imm, v) : std::max(s.imm, v);
          s.count++;
        } else {
          double v = op == WS_AGG_COUNT ? 1.0 : run(*env, bc, plan.aggs[a], r);
          if (plan.agg_null[a]) {  // nullable int column read raw: sentinel skipped
            i64 iv = int_value(*env, bc, plan.aggs[a], r);
            i64 nul = plan.agg_null[a] == WS_T_I32N ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
            v = iv == nul ? NaN() : (double)iv;
          }
          if (std::isnan(v)) continue;
          if (op == WS_AGG_SUM) s.fsum += v;
          else if (s.count == 0) s.fmm = v;
          else s.fmm = op == WS_AGG_MIN ? std::fmin(s.fmm, v) : std::fmax(s.fmm, v);
          s.count++;
        }
      }
    }
  }
  auto emit = [&](const Group& g, double* acc, int64_t* hi) {
    for (int a = 0; a < NA; ++a) {
      const Slot& s = g.slots[a];
      const int op = plan.ops[a];
      if (hi) hi[a] = 0;
      if (op == WS_AGG_COUNT) {
        acc[a] = (double)g.rows;
        continue;
      }
      const bool int_path = plan.agg_int[a] && (plan.agg_bits[a] || dense || !NK);
      if (int_path) {
        i64 mm = s.imm_set ? s.imm : (op == WS_AGG_MIN ? I64MAX : I64MIN);
        if (!NK) {  // no GROUP BY: 128-bit sums, int64 min / max bits
          if (op == WS_AGG_SUM) {
            acc[a] = from_bits((i64)(u64)(unsigned __int128)s.isum);
            if (hi) hi[a] = (int64_t)(s.isum >> 64);
          } else {
            acc[a] = from_bits(mm);
          }
        } else if (plan.agg_bits[a]) {
          acc[a] = from_bits(op == WS_AGG_SUM ? (i64)(u64)(unsigned __int128)s.isum : mm);
        } else {
          acc[a] = op == WS_AGG_SUM ? (double)(i64)s.isum : (double)mm;
        }
      } else {
        if (op == WS_AGG_SUM) acc[a] = (double)s.fsum;
        else acc[a] = s.count ? s.fmm : (op == WS_AGG_MIN ? INFINITY : -INFINITY);
      }
    }
  };
  if (!NK) {
    emit(nokeys, accs_out, hi_out);
    return 1;
  }
  const i64 ng = (i64)groups.size();
  if (ng > max_groups) return -ng;
  i64 gi = 0;
  for (auto& [kb, g] : groups) {  // std::map order = ascending key bit patterns (as signed int64)
    for (int k = 0; k < NK; ++k) keys_out[gi * NK + k] = g.keys[k];
    emit(g, accs_out + gi * NA, nullptr);
    ++gi;
  }
  return ng;
}

// Ascending global indices (row0 = prefix of n) of rows passing WHERE and not deleted; per_batch
// gets each batch's count. Returns the count.
int64_t or_select(const ws_expr_env* env, int has_where, ws_prog where, int nb, const int64_t* n,
                  const uint32_t* const* del, const ws_colview* cols, int ncols, int32_t* out, int64_t* per_batch) {
  i64 m = 0, row0 = 0;
  for (int b = 0; b < nb; ++b) {
    const ws_colview* bc = cols + (size_t)b * ncols;
    i64 c = 0;
    for (i64 r = 0; r < n[b]; ++r) {
      if (del && del[b] && ((del[b][r >> 5] >> (r & 31)) & 1u)) continue;
      if (has_where) {
        double w = run(*env, bc, where, r);
        if (std::isnan(w) || w == 0) continue;
      }
      out[m++] = (int32_t)(row0 + r);
      ++c;
    }
    if (per_batch) per_batch[b] = c;
    row0 += n[b];
  }
  return m;
}

// Join key of a key column (int types widened to int64; doubles truncated; NULL -> no key).
static bool join_key(const ws_colview& c, i64 r, i64* k) {
  if (c.t == WS_T_F64) {
    double v = ((const double*)c.p)[r];
    if (std::isnan(v)) return false;
    *k = (i64)v;
    return true;
  }
  i64 v = (c.t == WS_T_I32 || c.t == WS_T_I32N) ? ((const int32_t*)c.p)[r] : ((const int64_t*)c.p)[r];
  i64 nul = (c.t == WS_T_I32 || c.t == WS_T_I32N) ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
  if (v == nul) return false;
  *k = v;
  return true;
}

// Inner (outer = 0), left-outer (1: unmatched probe rows as (-1, j)) or full (2: plus unmatched
// build rows as (i, -1)) equi-join pairs, sorted by (probe, build). Returns the pair count, or
// -needed when cap is too small.
int64_t or_join(const ws_colview* bkey, int64_t nb, const ws_colview* pkey, int64_t np, int outer, int32_t* ob,
                int32_t* op, int64_t cap) {
  std::multimap<i64, i64> table;
  for (i64 i = 0; i < nb; ++i) {
    i64 k;
    if (join_key(*bkey, i, &k)) table.emplace(k, i);
  }
  std::vector<std::pair<i64, i64>> pairs;  // (probe, build)
  std::vector<char> matched(nb, 0);
  for (i64 j = 0; j < np; ++j) {
    i64 k;
    bool any = false;
    if (join_key(*pkey, j, &k)) {
      auto rg = table.equal_range(k);
      for (auto it = rg.first; it != rg.second; ++it) {
        pairs.push_back({j, it->second});
        matched[it->second] = 1;
        any = true;
      }
    }
    if (!any && outer) pairs.push_back({j, -1});
  }
  std::sort(pairs.begin(), pairs.end());
  if (outer == 2)
    for (i64 i = 0; i < nb; ++i)
      if (!matched[i]) pairs.push_back({-1, i});
  if ((i64)pairs.size() > cap) return -(i64)pairs.size();
  for (size_t i = 0; i < pairs.size(); ++i) op[i] = (int32_t)pairs[i].first, ob[i] = (int32_t)pairs[i].second;
  return (i64)pairs.size();
}

// Stable multi-key sort permutation: keys[0] most significant; NULL (NaN) sorts as -inf, -0 as 0.
void or_sort_perm(const double* const* keys, const uint8_t* desc, int nkeys, int64_t m, int32_t* perm) {
  std::vector<int32_t> p(m);
  std::iota(p.begin(), p.end(), 0);
  auto norm = [](double v) { return std::isnan(v) ? -INFINITY : (v == 0 ? 0.0 : v); };
  std::stable_sort(p.begin(), p.end(), [&](int32_t x, int32_t y) {
    for (int k = 0; k < nkeys; ++k) {
      double a = norm(keys[k][x]), b = norm(keys[k][y]);
      if (a == b) continue;
      return desc && desc[k] ? a > b : a < b;
    }
    return false;
  });
  std::copy(p.begin(), p.end(), perm);
}

// out[i] = map[in[i]] (NULL stays NULL, codes outside the map -> -2).
void or_remap_codes(const int32_t* in, int64_t n, const int32_t* map, int64_t map_n, int32_t* out) {
  for (i64 j = 0; j < n; ++j) {
    int32_t c = in[j];
    out[j] = c == WS_NULL_I32 ? WS_NULL_I32 : (c >= 0 && c < map_n ? map[c] : -2);
  }
}

// Query-level restatements of the measured shapes (plain host loops, one thread).

// SELECT user_id, COUNT(*), SUM(dur), COUNT(DISTINCT url) FROM h WHERE event_date < date_lt AND
// region < region_lt GROUP BY user_id: every group, ordered by count desc then user_id asc, as
// rows (user_id, count, sum(dur), distinct urls). The two-phase plan of the engine
// (engine_select.cpp:1403-1492) is restated as (user, url) pairs sorted and counted.
// Column statistics of one batch, the fields of ws_col_stats in order: rows, nulls, imin, imax
// (as int64), dmin, dmax (as doubles), then 2048 register bytes.
void or_column_stats(const ws_colview* c, int64_t n, const uint32_t* del, int64_t* counts, double* minmax,
                     uint8_t* regs) {
  i64 rows = 0, nulls = 0, imin = I64MAX, imax = I64MIN;
  double dmin = NaN(), dmax = NaN();
  std::memset(regs, 0, 2048);
  const bool is_int = c->t != WS_T_F64;
  const bool nullable = c->t == WS_T_I32N || c->t == WS_T_I64N || c->t == WS_T_DECIMAL;
  const i64 nul = c->t == WS_T_I32N ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
  for (i64 r = 0; r < n; ++r) {
    if (del && ((del[r >> 5] >> (r & 31)) & 1u)) continue;
    ++rows;
    u64 bits;
    double dv;
    if (is_int) {
      const i64 v = stored_int(*c, r);
      if (nullable && v == nul) {
        ++nulls;
        continue;
      }
      bits = (u64)v;
      imin = std::min(imin, v);
      imax = std::max(imax, v);
      dv = c->t == WS_T_DECIMAL ? (double)v / std::pow(10.0, c->scale) : (double)v;
    } else {
      dv = stored_dbl(*c, r);
      if (std::isnan(dv)) {
        ++nulls;
        continue;
      }
      const double z = dv + 0.0;
      std::memcpy(&bits, &z, 8);
    }
    dv += 0.0;
    dmin = std::isnan(dmin) ? dv : std::min(dmin, dv);
    dmax = std::isnan(dmax) ? dv : std::max(dmax, dv);
    u64 x = bits;  // splitmix64 finalizer
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    const unsigned idx = (unsigned)(x >> 53);
    u64 w = (x << 11) | (1ull << 10);
    unsigned rank = 1;
    while (!(w >> 63)) {
      w <<= 1;
      ++rank;
    }
    if (rank > regs[idx]) regs[idx] = (uint8_t)rank;
  }
  counts[0] = rows, counts[1] = nulls, counts[2] = imin, counts[3] = imax;
  minmax[0] = dmin, minmax[1] = dmax;
}

int64_t or_c2_query(const int64_t* user, const int32_t* region, const int32_t* date, const int32_t* url, const int32_t* dur,
                    int64_t n, int32_t date_lt, int32_t region_lt, int64_t* out, int64_t cap) {
  struct R {
    int64_t user;
    int32_t url, dur;
  };
  std::vector<R> rows;
  rows.reserve((size_t)n);
  for (int64_t i = 0; i < n; ++i)
    if (date[i] < date_lt && region[i] < region_lt) rows.push_back({user[i], url[i], dur[i]});
  std::sort(rows.begin(), rows.end(), [](const R& a, const R& b) { return a.user != b.user ? a.user < b.user : a.url < b.url; });
  std::vector<std::array<int64_t, 4>> g;
  for (size_t i = 0; i < rows.size();) {
    size_t j = i;
    int64_t cnt = 0, sum = 0, distinct = 0;
    while (j < rows.size() && rows[j].user == rows[i].user) {
      if (j == i || rows[j].url != rows[j - 1].url) ++distinct;
      ++cnt;
      sum += rows[j].dur;
      ++j;
    }
    g.push_back({rows[i].user, cnt, sum, distinct});
    i = j;
  }
  std::sort(g.begin(), g.end(), [](const std::array<int64_t, 4>& a, const std::array<int64_t, 4>& b) {
    return a[1] != b[1] ? a[1] > b[1] : a[0] < b[0];
  });
  if ((int64_t)g.size() > cap) return -(int64_t)g.size();
  for (size_t i = 0; i < g.size(); ++i)
    for (int k = 0; k < 4; ++k) out[i * 4 + k] = g[i][k];
  return (int64_t)g.size();
}

// SELECT f.flag, COUNT(*), SUM(f.price), SUM(f.qty) FROM fact f JOIN dim d ON f.fk = d.key WHERE
// d.attr < attr_lt GROUP BY f.flag: rows (flag, count, raw price sum, qty sum) ordered by flag.
int64_t or_c3_query(const int32_t* fk, const int64_t* qty, const int64_t* price, const int32_t* flag, int64_t nf,
                    const int32_t* dkey, const int32_t* dattr, int64_t nd, int32_t attr_lt, int64_t* out, int64_t cap) {
  std::multiset<int32_t> keys;
  for (int64_t i = 0; i < nd; ++i)
    if (dattr[i] < attr_lt) keys.insert(dkey[i]);
  std::vector<int32_t> sorted(keys.begin(), keys.end());
  std::map<int32_t, std::array<__int128, 3>> g;
  for (int64_t i = 0; i < nf; ++i) {
    auto rg = std::equal_range(sorted.begin(), sorted.end(), fk[i]);
    const int64_t mult = rg.second - rg.first;
    if (!mult) continue;
    auto& a = g[flag[i]];
    a[0] += mult;
    a[1] += (__int128)price[i] * mult;
    a[2] += (__int128)qty[i] * mult;
  }
  if ((int64_t)g.size() > cap) return -(int64_t)g.size();
  int64_t r = 0;
  for (auto& [f, a] : g) {
    out[r * 4 + 0] = f;
    out[r * 4 + 1] = (int64_t)a[0];
    out[r * 4 + 2] = (int64_t)a[1];
    out[r * 4 + 3] = (int64_t)a[2];
    ++r;
  }
  return r;
}

}  // extern "C"
