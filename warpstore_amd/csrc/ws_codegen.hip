// Code generation for per-query compiled plans (host code only): the engine's stack bytecode
// (reference src/engine_core.h:1485-1859 compiles it; src/devtypes.h:430-495 defines its meaning)
// becomes a C++ evaluator struct that the shared kernel bodies (ws_kernels.h) are instantiated
// with under hipRTC.
//
// Two phases:
//   1. lowering: the bytecode of every program of the plan is replayed on a symbolic stack into
//      one arena of typed IR nodes. Values are either exact int64 (integer columns and literals,
//      never NULL) or double (NULL = NaN, three-valued logic); identical nodes are shared
//      (hash-consing), so a column read by several programs is loaded once per row;
//   2. emission: each evaluator function (where, key, agg, ...) is emitted as straight-line
//      static single assignments, one `const` temporary per node in dependency order. Checked
//      int64 arithmetic can raise the overflow flag, so a node that depends on such a flag-raising
//      node under a lazy construct (CASE, AND / OR over such operands) is emitted inside the
//      branch that needs it, keeping the flag exactly as a lazy evaluation raises it; everything
//      else is computed branch-free, so all of a row's loads issue together.
// Literal values are never part of the source: the evaluator reads them from the env when it is
// constructed (scalar registers), so one compiled shape serves every literal.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ws_codegen.h"
#include "ws_internal.h"

namespace ws {
namespace codegen {
namespace {

enum class Ty : uint8_t { I64, F64 };

// IR opcodes: the bytecode's, plus RAWINT (a column's stored integer with its NULL sentinel).
constexpr int kRawInt = 200;

struct Node {
  int op = 0;
  int arg = 0;
  int kid[3] = {-1, -1, -1};
  Ty ty = Ty::F64;
  bool flags = false;  // evaluating it may raise the checked-arithmetic overflow flag
};

// Operand count of each bytecode opcode, from the VM's stack effect.
int arity(int op) {
  switch (op) {
    case WS_OP_COL: case WS_OP_COLRAW: case WS_OP_CONST: case WS_OP_VDOT: case WS_OP_VEUC: case kRawInt: return 0;
    case WS_OP_NEG: case WS_OP_NOT: case WS_OP_INSET: case WS_OP_YEAR: case WS_OP_ISNULL: case WS_OP_CALL1:
    case WS_OP_DICTMAP: case WS_OP_INSORTED: return 1;
    case WS_OP_SELECT: return 3;
    default: return 2;
  }
}

// How a plan's column slots are read: directly by row, or from words a kernel loaded earlier.
enum class Source { Direct, Row };

struct Lowering {
  const ws_expr_env& env;
  int ncols;
  std::vector<Node> nodes;
  std::map<std::tuple<int, int, int, int, int, int>, int> interned;
  std::set<int> cols, ilits, dlits, encoded, sets;

  Lowering(const ws_expr_env& e, int n) : env(e), ncols(n) {}

  int intern(Node n) {
    const auto key = std::make_tuple(n.op, n.arg, n.kid[0], n.kid[1], n.kid[2], (int)n.ty);
    auto it = interned.find(key);
    if (it != interned.end()) return it->second;
    for (int i = 0; i < 3; ++i)
      if (n.kid[i] >= 0) n.flags = n.flags || nodes[(size_t)n.kid[i]].flags;
    nodes.push_back(n);
    interned.emplace(key, (int)nodes.size() - 1);
    return (int)nodes.size() - 1;
  }
  const Node& at(int i) const { return nodes[(size_t)i]; }
  bool is_int(int i) const { return at(i).ty == Ty::I64; }

  int column(int slot) {
    if (slot < 0 || slot >= ncols || slot >= WS_MAX_COLS) throw Error("codegen: column slot out of range", WS_ERR_INTERNAL);
    cols.insert(slot);
    if (env.cols[slot].enc && env.cols[slot].t != WS_T_F64) encoded.insert(slot);
    return slot;
  }
  // A node for one instruction given its operands (already lowered).
  int make(const ws_instr in, const int* k) {
    Node n;
    n.op = in.op;
    auto both_int = [&] { return is_int(k[0]) && is_int(k[1]); };
    switch (in.op) {
      case WS_OP_COL: {
        const ws_colview& v = env.cols[column(in.arg)];
        if (v.t == WS_T_VEC) throw Error("codegen: vector column used as a scalar", WS_ERR_INTERNAL);
        n.arg = in.arg;
        n.ty = (v.t == WS_T_I32 || v.t == WS_T_I64) ? Ty::I64 : Ty::F64;
        return intern(n);
      }
      case WS_OP_COLRAW:
        n.arg = column(in.arg);
        n.ty = (env.cols[in.arg].t == WS_T_DECIMAL && !in.is_int) ? Ty::F64 : Ty::I64;
        n.op = n.ty == Ty::I64 ? kRawInt : WS_OP_COLRAW;
        return intern(n);
      case WS_OP_CONST:
        if (in.arg >= WS_MAX_CONSTS) throw Error("codegen: literal out of range", WS_ERR_INTERNAL);
        n.arg = in.arg;
        n.ty = in.is_int ? Ty::I64 : Ty::F64;
        (in.is_int ? ilits : dlits).insert(in.arg);
        return intern(n);
      case WS_OP_VDOT: case WS_OP_VEUC:
        if ((in.arg >> 8) >= WS_MAX_SETS) throw Error("codegen: vector operand out of range", WS_ERR_INTERNAL);
        column(in.arg & 255);
        n.arg = in.arg;
        return intern(n);
      case WS_OP_INSET:
        if (in.arg >= WS_MAX_SETS) throw Error("codegen: set out of range", WS_ERR_INTERNAL);
        sets.insert(in.arg);
        n.arg = in.arg;
        n.kid[0] = k[0];
        n.ty = at(k[0]).ty;
        return intern(n);
      case WS_OP_DICTMAP: case WS_OP_INSORTED:
        if (in.arg >= WS_MAX_SETS) throw Error("codegen: map out of range", WS_ERR_INTERNAL);
        n.arg = in.arg;
        n.kid[0] = k[0];
        return intern(n);
      case WS_OP_COALESCE:
        if (is_int(k[0])) return k[0];  // an integer value is never NULL: the fallback is not read
        n.kid[0] = k[0], n.kid[1] = k[1];
        return intern(n);
      default: break;
    }
    for (int i = 0; i < arity(in.op); ++i) n.kid[i] = k[i];
    n.arg = (in.op == WS_OP_CALL1 || in.op == WS_OP_CALL2) ? in.arg : 0;
    switch (in.op) {
      case WS_OP_ADD: case WS_OP_SUB: case WS_OP_MUL:
        n.ty = both_int() ? Ty::I64 : Ty::F64;
        n.flags = n.ty == Ty::I64;
        break;
      case WS_OP_EQ: case WS_OP_NE: case WS_OP_LT: case WS_OP_LE: case WS_OP_GT: case WS_OP_GE:
      case WS_OP_AND: case WS_OP_OR:
        n.ty = both_int() ? Ty::I64 : Ty::F64;
        break;
      case WS_OP_NEG: case WS_OP_NOT: case WS_OP_YEAR: n.ty = at(k[0]).ty; break;
      case WS_OP_ISNULL: n.ty = Ty::I64; break;
      case WS_OP_SELECT: n.ty = is_int(k[1]) && is_int(k[2]) ? Ty::I64 : Ty::F64; break;
      case WS_OP_DIV: case WS_OP_MOD: case WS_OP_NULLIF: case WS_OP_CALL1: case WS_OP_CALL2: n.ty = Ty::F64; break;
      default: throw Error("codegen: unknown opcode", WS_ERR_INTERNAL);
    }
    return intern(n);
  }
  // Replays a program on a symbolic stack; returns its root node.
  int lower(ws_prog p) {
    std::vector<int> stack;
    for (int pc = p.start; pc < p.start + p.len; ++pc) {
      if (pc < 0 || pc >= WS_MAX_CODE) throw Error("codegen: program out of range", WS_ERR_INTERNAL);
      const ws_instr in = env.code[pc];
      const int a = arity(in.op);
      if ((int)stack.size() < a) throw Error("codegen: malformed expression", WS_ERR_INTERNAL);
      int k[3] = {-1, -1, -1};
      for (int i = a - 1; i >= 0; --i) k[i] = stack.back(), stack.pop_back();
      stack.push_back(make(in, k));
    }
    if (stack.size() != 1) throw Error("codegen: malformed expression", WS_ERR_INTERNAL);
    return stack[0];
  }
  // The stored integer of a nullable integer column (sentinel included): aggregate arguments the
  // kernels skip on the sentinel themselves.
  int raw_column(int slot) {
    Node n;
    n.op = kRawInt;
    n.arg = column(slot);
    n.ty = Ty::I64;
    return intern(n);
  }
};

// Member names of the generated evaluator.
std::string col_ptr(int c) { return "col" + std::to_string(c); }
std::string col_view(int c) { return "view" + std::to_string(c); }
std::string enc_view(int c) { return "enc" + std::to_string(c); }
std::string row_word(int c) { return "x.w" + std::to_string(c); }

const char* elem_type(int t) { return t == WS_T_F64 ? "double" : (t == WS_T_I32 || t == WS_T_I32N) ? "int" : "long long"; }
const char* global_ptr_type(int t) {
  return (t == WS_T_I32 || t == WS_T_I32N) ? "const int WS_GLOBAL*"
         : t == WS_T_F64                   ? "const double WS_GLOBAL*"
         : t == WS_T_VEC                   ? "const float WS_GLOBAL*"
                                           : "const long long WS_GLOBAL*";
}
const char* cmp_token(int op) {
  switch (op) {
    case WS_OP_EQ: return "==";
    case WS_OP_NE: return "!=";
    case WS_OP_LT: return "<";
    case WS_OP_LE: return "<=";
    case WS_OP_GT: return ">";
    default: return ">=";
  }
}

// Emits one evaluator function body: statements into `out`, returning the root's temporary.
struct Emitter {
  const Lowering& lw;
  Source src;
  std::ostringstream& out;
  std::vector<std::set<int>> scopes{{}};  // temporaries visible in each nested block
  std::set<int> row_cols;                 // Row mode: columns whose words the body reads
  bool row_ok = true;                     // false: a read with no row form (a dictionary double)

  Emitter(const Lowering& l, Source s, std::ostringstream& o) : lw(l), src(s), out(o) {}

  static std::string tmp(int i) { return "t" + std::to_string(i); }
  bool visible(int i) const {
    for (const auto& s : scopes)
      if (s.count(i)) return true;
    return false;
  }
  std::string dbl(int i) const { return lw.is_int(i) ? "(double)" + tmp(i) : tmp(i); }
  std::string truth(int i) const { return lw.is_int(i) ? "(" + tmp(i) + " != 0)" : "ws::truthy(" + tmp(i) + ")"; }
  std::string indent() const { return std::string(4 + 2 * scopes.size(), ' '); }

  std::string stored_int(int c) {
    const ws_colview& v = lw.env.cols[c];
    if (src == Source::Row) {
      row_cols.insert(c);
      return v.enc ? "ws::enc_decode(" + enc_view(c) + "," + row_word(c) + ",r)" : "(long long)" + row_word(c);
    }
    return v.enc ? "ws::enc_get(" + enc_view(c) + ",r)" : "(long long)" + col_ptr(c) + "[r]";
  }
  std::string column_value(int c) {
    const ws_colview& v = lw.env.cols[c];
    switch (v.t) {
      case WS_T_F64:
        if (v.enc) {
          if (src == Source::Row) row_ok = false;
          return "ws::col_dbl(" + col_view(c) + ",r)";
        }
        if (src == Source::Row) {
          row_cols.insert(c);
          return row_word(c);
        }
        return col_ptr(c) + "[r]";
      case WS_T_I32N: return "ws::null_if_i32(" + stored_int(c) + ")";
      case WS_T_I64N: return "ws::null_if_i64(" + stored_int(c) + ")";
      case WS_T_DECIMAL: {
        char p10[48];
        snprintf(p10, sizeof p10, "%.1f", pow10i(v.scale));
        return "ws::dec_of(" + stored_int(c) + "," + p10 + ")";
      }
      default: return stored_int(c);
    }
  }

  // Emits node i (and what it needs) into the current block; returns its temporary.
  std::string emit(int i) {
    if (visible(i)) return tmp(i);
    const Node& n = lw.at(i);
    const bool lazy_bool = (n.op == WS_OP_AND || n.op == WS_OP_OR) && lw.is_int(i) && lw.at(n.kid[1]).flags;
    const bool lazy_case = n.op == WS_OP_SELECT && (lw.at(n.kid[1]).flags || lw.at(n.kid[2]).flags);
    if (lazy_bool) {
      // (a != 0) && / || (b != 0): b, which can raise the overflow flag, only when it decides
      const std::string a = emit(n.kid[0]);
      out << indent() << "long long " << tmp(i) << " = " << a << " != 0;\n";
      out << indent() << "if (" << (n.op == WS_OP_AND ? "" : "!") << tmp(i) << ") {\n";
      scopes.emplace_back();
      const std::string b = emit(n.kid[1]);
      out << indent() << tmp(i) << " = " << b << " != 0;\n";
      scopes.pop_back();
      out << indent() << "}\n";
      scopes.back().insert(i);
      return tmp(i);
    }
    if (lazy_case) {
      const std::string c = emit(n.kid[0]);
      const char* ty = lw.is_int(i) ? "long long" : "double";
      out << indent() << ty << " " << tmp(i) << ";\n" << indent() << "if (" << truth(n.kid[0]) << ") {\n";
      for (int side = 1; side <= 2; ++side) {
        scopes.emplace_back();
        emit(n.kid[side]);
        out << indent() << tmp(i) << " = " << (lw.is_int(i) ? tmp(n.kid[side]) : dbl(n.kid[side])) << ";\n";
        scopes.pop_back();
        out << indent() << (side == 1 ? "} else {\n" : "}\n");
      }
      (void)c;
      scopes.back().insert(i);
      return tmp(i);
    }
    std::string k[3];
    for (int j = 0; j < arity(n.op); ++j) k[j] = emit(n.kid[j]);
    const bool I = lw.is_int(i);
    std::string e;
    switch (n.op) {
      case WS_OP_COL: e = column_value(n.arg); break;
      case kRawInt: e = stored_int(n.arg); break;
      case WS_OP_COLRAW: e = "ws::null_if_i64(" + stored_int(n.arg) + ")"; break;
      case WS_OP_CONST: e = n.arg < 0 ? std::string("1ll") : (I ? "li" : "ld") + std::to_string(n.arg); break;
      case WS_OP_ADD: case WS_OP_SUB: case WS_OP_MUL:
        if (I) {
          const char* f = n.op == WS_OP_ADD ? "ws::add_chk" : n.op == WS_OP_SUB ? "ws::sub_chk" : "ws::mul_chk";
          e = std::string(f) + "(" + k[0] + "," + k[1] + ",&ovf_)";
        } else {
          e = dbl(n.kid[0]) + (n.op == WS_OP_ADD ? " + " : n.op == WS_OP_SUB ? " - " : " * ") + dbl(n.kid[1]);
        }
        break;
      case WS_OP_DIV: e = "ws::div_or_null(" + dbl(n.kid[0]) + "," + dbl(n.kid[1]) + ")"; break;
      case WS_OP_MOD: e = "fmod(" + dbl(n.kid[0]) + "," + dbl(n.kid[1]) + ")"; break;
      case WS_OP_NEG: e = I ? "(long long)(0ull - (unsigned long long)" + k[0] + ")" : "-" + k[0]; break;
      case WS_OP_EQ: case WS_OP_NE: case WS_OP_LT: case WS_OP_LE: case WS_OP_GT: case WS_OP_GE:
        e = I ? "(long long)(" + k[0] + " " + cmp_token(n.op) + " " + k[1] + ")"
              : "ws::tri_cmp(" + std::to_string(n.op - WS_OP_EQ) + "," + dbl(n.kid[0]) + "," + dbl(n.kid[1]) + ")";
        break;
      case WS_OP_AND: case WS_OP_OR:
        // both sides are evaluated: their loads issue together (neither can raise a flag here)
        e = I ? "(long long)((" + k[0] + " != 0) " + (n.op == WS_OP_AND ? "&" : "|") + " (" + k[1] + " != 0))"
              : std::string(n.op == WS_OP_AND ? "ws::tri_and(" : "ws::tri_or(") + dbl(n.kid[0]) + "," + dbl(n.kid[1]) + ")";
        break;
      case WS_OP_NOT: e = I ? "(long long)(" + k[0] + " == 0)" : "ws::tri_not(" + k[0] + ")"; break;
      case WS_OP_INSET: {
        const std::string s = std::to_string(n.arg);
        e = std::string(I ? "ws::inset_s(" : "ws::inset_sd(") + "setl" + s + ",sets" + s + ",setg" + s + ",setn" + s + "," + k[0] +
            ")";
        break;
      }
      case WS_OP_SELECT:  // branch-free: neither side can raise a flag
        e = truth(n.kid[0]) + " ? " + (I ? k[1] : dbl(n.kid[1])) + " : " + (I ? k[2] : dbl(n.kid[2]));
        break;
      case WS_OP_YEAR:
        e = I ? "ws::year_of_days(" + k[0] + ")" : "ws::call1(WS_FN_DT_YEAR,ws::call1(WS_FN_DATE_TO_DT," + k[0] + "))";
        break;
      case WS_OP_ISNULL: e = lw.is_int(n.kid[0]) ? std::string("0ll") : "(long long)ws::isnan_(" + k[0] + ")"; break;
      case WS_OP_COALESCE: e = "ws::first_non_null(" + k[0] + "," + dbl(n.kid[1]) + ")"; break;
      case WS_OP_NULLIF: e = "ws::null_if_equal(" + dbl(n.kid[0]) + "," + dbl(n.kid[1]) + ")"; break;
      case WS_OP_CALL1: e = "ws::call1(" + std::to_string(n.arg) + "," + dbl(n.kid[0]) + ")"; break;
      case WS_OP_CALL2: e = "ws::call2(" + std::to_string(n.arg) + "," + dbl(n.kid[0]) + "," + dbl(n.kid[1]) + ")"; break;
      case WS_OP_DICTMAP: case WS_OP_INSORTED: {
        const std::string m = std::to_string(n.arg);
        e = std::string(n.op == WS_OP_DICTMAP ? "ws::dictmap_d(" : "ws::insorted_d(") + "env->maps[" + m + "],env->map_size[" + m +
            "]," + dbl(n.kid[0]) + ")";
        break;
      }
      case WS_OP_VDOT: case WS_OP_VEUC: {
        const std::string m = std::to_string(n.arg >> 8);
        if (src == Source::Row) row_ok = false;
        e = std::string(n.op == WS_OP_VDOT ? "ws::vdot_d(" : "ws::veuc_d(") + "(const float*)" + col_view(n.arg & 255) +
            ".p,r,env->maps[" + m + "],env->map_size[" + m + "])";
        break;
      }
      default: throw Error("codegen: unknown opcode", WS_ERR_INTERNAL);
    }
    out << indent() << "const " << (I ? "long long " : "double ") << tmp(i) << " = " << e << ";\n";
    scopes.back().insert(i);
    return tmp(i);
  }
};

// One evaluator function: `signature { body; return ret(root); }`, or a switch over several roots.
void function(std::ostringstream& o, const Lowering& lw, const std::string& signature, const std::vector<int>& roots,
              const std::string& kind /* "bool", "double", "int" */) {
  o << "  __device__ " << signature << " const {\n";
  auto ret = [&](int root) {
    std::ostringstream body;
    Emitter em(lw, Source::Direct, body);
    em.emit(root);
    std::string v = Emitter::tmp(root);
    if (kind == "bool") v = em.truth(root);
    else if (kind == "double") v = em.dbl(root);
    else if (!lw.is_int(root)) v = "(long long)" + v;
    return body.str() + std::string(4, ' ') + "  return " + v + ";\n";
  };
  if (roots.size() == 1 && kind == "bool") {
    o << ret(roots[0]);
  } else {
    o << "    switch (i) {\n";
    for (size_t i = 0; i < roots.size(); ++i) o << "    case " << i << ": {\n" << ret(roots[i]) << "    }\n";
    o << "    }\n    return 0;\n";
  }
  o << "  }\n";
}

// The same function over a row's loaded column words (x, of struct VRow): adds the columns it
// reads to `cols`; clears `ok` when a read has no row form.
std::string function_row(const Lowering& lw, const std::string& signature, const std::vector<int>& roots,
                         const std::string& kind /* "double", "int" */, std::set<int>& cols, bool& ok) {
  std::ostringstream o;
  o << "  __device__ " << signature << " const {\n    switch (i) {\n";
  for (size_t i = 0; i < roots.size(); ++i) {
    std::ostringstream body;
    Emitter em(lw, Source::Row, body);
    em.emit(roots[i]);
    std::string v = Emitter::tmp(roots[i]);
    if (kind == "double") v = em.dbl(roots[i]);
    else if (!lw.is_int(roots[i])) v = "(long long)" + v;
    cols.insert(em.row_cols.begin(), em.row_cols.end());
    ok = ok && em.row_ok;
    o << "    case " << i << ": {\n" << body.str() << std::string(4, ' ') << "  return " << v << ";\n    }\n";
  }
  o << "    }\n    return 0;\n  }\n";
  return o.str();
}

int dense_kind(int op, bool is_int) {
  if (op == WS_AGG_COUNT) return WS_DK_COUNT;
  if (op == WS_AGG_SUM) return is_int ? WS_DK_SUM_I : WS_DK_SUM_D;
  if (op == WS_AGG_MIN) return is_int ? WS_DK_MIN_I : WS_DK_MIN_D;
  return is_int ? WS_DK_MAX_I : WS_DK_MAX_D;
}

// A `switch (i)` returning per-index constants.
template <class F>
void constant_table(std::ostringstream& o, const char* signature, int n, F value) {
  o << "  __device__ " << signature << " const { switch (i) {";
  for (int i = 0; i < n; ++i) o << " case " << i << ": return " << value(i) << ";";
  o << " } return 0; }\n";
}

}  // namespace

bool program_is_int(const ws_expr_env& env, ws_prog p) {
  Lowering lw(env, WS_MAX_COLS);
  return lw.is_int(lw.lower(p));
}

std::string evaluator_source(const ws_expr_env& env, int ncols, const ws_agg_plan& plan, int* rows_out) {
  if (plan.nkeys < 0 || plan.nkeys > WS_MAX_KEYS || plan.naggs < 0 || plan.naggs > WS_MAX_AGGS || plan.agg_null[0])
    throw Error("codegen: plan out of range", WS_ERR_INTERNAL);
  Lowering lw(env, ncols);
  const int where = plan.has_where ? lw.lower(plan.where) : -1;
  std::vector<int> keys, kfull, aggs, iaggs;
  for (int k = 0; k < plan.nkeys; ++k) keys.push_back(lw.lower(plan.keys[k]));
  for (int a = 0; a < plan.naggs; ++a) {
    if (plan.ops[a] == WS_AGG_COUNT) {
      Node one;  // COUNT(*) counts rows: the integer 1 (arg -1: no literal slot)
      one.op = WS_OP_CONST, one.arg = -1, one.ty = Ty::I64;
      const int c = lw.intern(one);
      aggs.push_back(c), iaggs.push_back(c);
      continue;
    }
    const int root = lw.lower(plan.aggs[a]);
    aggs.push_back(root);
    if (plan.agg_null[a]) {
      const ws_instr in = env.code[plan.aggs[a].start];
      if (!(plan.aggs[a].len == 1 && in.op == WS_OP_COL)) throw Error("codegen: nullable aggregate argument is not a column", WS_ERR_INTERNAL);
      iaggs.push_back(lw.raw_column(in.arg));
    } else {
      iaggs.push_back(root);
    }
  }
  const int nk = std::max(plan.nkeys, 1), na = std::max(plan.naggs, 1);
  const int rows = std::max(1, std::min(4, 12 / (nk + na)));  // rows per thread per step (registers)
  *rows_out = rows;
  // a program that is one column or literal has no side effect and may be evaluated for rows the
  // WHERE then drops (the partition passes do)
  auto single = [&](ws_prog p) {
    return p.len == 1 && (env.code[p.start].op == WS_OP_COL || env.code[p.start].op == WS_OP_CONST);
  };
  bool pure_keys = true, pure_vals = true;
  for (int k = 0; k < plan.nkeys; ++k) pure_keys = pure_keys && single(plan.keys[k]);
  for (int a = 0; a < plan.naggs; ++a) pure_vals = pure_vals && (plan.ops[a] == WS_AGG_COUNT || single(plan.aggs[a]));
  const bool pure_where = where < 0 || !lw.at(where).flags;

  std::ostringstream o;
  o << "struct QEval {\n";
  o << "  static constexpr int kKeys = " << nk << ", kAggs = " << na << ", kRows = " << rows << ";\n";
  o << "  static constexpr bool kPureKeys = " << (pure_keys ? "true" : "false") << ", kPureVals = " << (pure_vals ? "true" : "false")
    << ", kPureWhere = " << (pure_where ? "true" : "false") << ", kGlobalCols = true;\n";
  o << "  const ws_expr_env* env;\n";
  // Overflow of checked arithmetic: a per-thread flag stored to env->ovf once, when the evaluator
  // goes out of scope (a conditional store inside the row loop would order every later load).
  o << "  mutable int ovf_ = 0;\n  __device__ ~QEval() { if (ovf_ && env->ovf) *env->ovf = 1; }\n";
  for (int c : lw.cols) o << "  " << global_ptr_type(env.cols[c].t) << " " << col_ptr(c) << ";\n  ws_colview " << col_view(c) << ";\n";
  for (int c : lw.encoded) o << "  ws::EncView " << enc_view(c) << ";\n";
  for (int k : lw.ilits) o << "  long long li" << k << ";\n";
  for (int k : lw.dlits) o << "  double ld" << k << ";\n";
  for (int s : lw.sets)
    o << "  const unsigned WS_GLOBAL* setg" << s << ";\n  long long setn" << s << ";\n  const unsigned WS_LOCAL* setl" << s
      << ";\n  int sets" << s << ";\n";
  o << "  __device__ QEval(const ws::ScanDesc& d) : env(d.env) {\n";
  for (int k : lw.ilits) o << "    li" << k << " = ws::uni((long long)ws::dbits(env->consts[" << k << "]));\n";
  for (int k : lw.dlits) o << "    ld" << k << " = ws::bitsd(ws::uni((long long)ws::dbits(env->consts[" << k << "])));\n";
  // an empty set points at its own size word (always readable; every bit is masked off)
  for (int s : lw.sets) {
    const std::string n = std::to_string(s);
    o << "    setn" << n << " = ws::uni((long long)(env->set_size[" << n << "] > 0 ? env->set_size[" << n << "] : 0));\n";
    o << "    setg" << n << " = (const unsigned WS_GLOBAL*)ws::uni(setn" << n << " > 0 ? env->sets[" << n
      << "] : (const unsigned*)&env->set_size[" << n << "]);\n";
    o << "    setl" << n << " = ws::lds_spare_word(); sets" << n << " = 0;\n";
  }
  o << "  }\n";
  // leading words of the sets copied to LDS, in set order, within `words`
  o << "  __device__ void stage_sets(unsigned* lds, int words) {\n    int off = 0;\n";
  for (int s : lw.sets) {
    const std::string n = std::to_string(s);
    o << "    { const int w = ws::stage_set(lds, off, words, setg" << n << ", setn" << n << "); if (w > 0) { setl" << n
      << " = (const unsigned WS_LOCAL*)(lds + off); sets" << n << " = w; } off += w; }\n";
  }
  o << "    (void)off;\n  }\n";
  o << "  __device__ void bind(const ws_colview* c) {\n";
  for (int c : lw.cols)
    o << "    " << col_view(c) << " = ws::uniform_view(c[" << c << "]); " << col_ptr(c) << " = (" << global_ptr_type(env.cols[c].t) << ")"
      << col_view(c) << ".p;\n";
  for (int c : lw.encoded) o << "    " << enc_view(c) << " = ws::enc_view(" << col_view(c) << ");\n";
  o << "  }\n";
  o << "  __device__ int nkeys() const { return " << plan.nkeys << "; }\n";
  o << "  __device__ int naggs() const { return " << plan.naggs << "; }\n";
  o << "  __device__ bool has_where() const { return " << (plan.has_where ? "true" : "false") << "; }\n";
  constant_table(o, "int op(int i)", plan.naggs, [&](int a) { return std::to_string((int)plan.ops[a]); });
  constant_table(o, "bool agg_int(int i)", plan.naggs, [&](int a) { return plan.agg_int[a] ? "true" : "false"; });
  constant_table(o, "bool agg_bits(int i)", plan.naggs, [&](int a) { return plan.agg_bits[a] ? "true" : "false"; });
  constant_table(o, "int agg_part(int i)", plan.naggs, [&](int a) { return std::to_string((int)plan.agg_part[a]); });
  constant_table(o, "int agg_null(int i)", plan.naggs, [&](int a) { return std::to_string((int)plan.agg_null[a]); });
  constant_table(o, "bool key_int(int i)", plan.nkeys, [&](int k) { return plan.key_int[k] ? "true" : "false"; });
  constant_table(o, "uint8_t dkind(int i, const ws_dense_spec&)", plan.naggs,
                 [&](int a) { return std::to_string(dense_kind(plan.ops[a], plan.agg_int[a] != 0)); });
  if (where >= 0) function(o, lw, "bool where(long long r)", {where}, "bool");
  else o << "  __device__ bool where(long long) const { return true; }\n";

  // The WHERE in a load half (load_row: the row's column words) and a compute half (where_row),
  // for kernels that issue a later step's loads before computing this one.
  {
    std::ostringstream body;
    Emitter em(lw, Source::Row, body);
    std::string cond = "true";
    if (where >= 0) {
      em.emit(where);
      cond = em.truth(where);
    }
    const bool split = where >= 0 && em.row_ok && !em.row_cols.empty();
    o << "  static constexpr bool kSplit = " << (split ? "true" : "false") << ";\n";
    // where_row needs no load of its own: a split WHERE, or none
    o << "  static constexpr bool kRowSplit = " << (split || where < 0 ? "true" : "false") << ";\n";
    o << "  struct Row {";
    if (split)
      for (int c : em.row_cols) o << " " << (env.cols[c].enc ? "unsigned" : elem_type(env.cols[c].t)) << " w" << c << ";";
    o << " int pad_; };\n";
    o << "  __device__ void load_row(long long r, Row& x) const {";
    if (split)
      for (int c : em.row_cols)
        o << " " << row_word(c) << " = " << (env.cols[c].enc ? "ws::enc_word(" + enc_view(c) + ",r)" : col_ptr(c) + "[r]") << ";";
    o << " }\n";
    // four consecutive rows (r a multiple of 4) with 16-byte loads of the raw columns
    o << "  __device__ bool rows4_ok() const { return true";
    if (split)
      for (int c : em.row_cols)
        if (!env.cols[c].enc) o << " && (((unsigned long long)" << col_ptr(c) << " & 15) == 0)";
    o << "; }\n";
    o << "  __device__ void load_rows4(long long r, Row* x) const {";
    if (split)
      for (int c : em.row_cols) {
        const ws_colview& v = env.cols[c];
        const std::string w = ".w" + std::to_string(c);
        if (v.enc) {
          for (int q = 0; q < 4; ++q) o << " x[" << q << "]" << w << " = ws::enc_word(" << enc_view(c) << ",r+" << q << ");";
        } else if (v.t == WS_T_I32 || v.t == WS_T_I32N) {
          o << " { const ws::ws_i32x4 q = *(const ws::ws_i32x4 WS_GLOBAL*)(" << col_ptr(c) << " + r); x[0]" << w << " = q.x; x[1]" << w
            << " = q.y; x[2]" << w << " = q.z; x[3]" << w << " = q.w; }";
        } else {
          const char* vt = v.t == WS_T_F64 ? "ws::ws_f64x2" : "ws::ws_i64x2";
          o << " { const " << vt << " q0 = *(const " << vt << " WS_GLOBAL*)(" << col_ptr(c) << " + r), q1 = *(const " << vt
            << " WS_GLOBAL*)(" << col_ptr(c) << " + r + 2); x[0]" << w << " = q0.x; x[1]" << w << " = q0.y; x[2]" << w << " = q1.x; x[3]"
            << w << " = q1.y; }";
        }
      }
    o << " }\n";
    // Four consecutive rows kept as the 16-byte vectors they load as (a kernel that carries the
    // next rows across a loop iteration keeps each load's registers whole, with no copy that would
    // wait for the load); row_of4 takes row s of them.
    o << "  struct Row4 {";
    if (split)
      for (int c : em.row_cols) {
        const ws_colview& v = env.cols[c];
        if (v.enc) o << " ws::ws_u32x4 w" << c << ";";
        else if (v.t == WS_T_I32 || v.t == WS_T_I32N) o << " ws::ws_i32x4 w" << c << ";";
        else o << " " << (v.t == WS_T_F64 ? "ws::ws_f64x2" : "ws::ws_i64x2") << " w" << c << "a, w" << c << "b;";
      }
    o << " int pad_; };\n";
    o << "  __device__ void load_row4(long long r, Row4& x) const {";
    if (split)
      for (int c : em.row_cols) {
        const ws_colview& v = env.cols[c];
        const std::string w = "x.w" + std::to_string(c);
        if (v.enc) {
          o << " " << w << " = ws::ws_u32x4{";
          for (int q = 0; q < 4; ++q) o << (q ? "," : "") << "ws::enc_word(" << enc_view(c) << ",r+" << q << ")";
          o << "};";
        } else if (v.t == WS_T_I32 || v.t == WS_T_I32N) {
          o << " " << w << " = *(const ws::ws_i32x4 WS_GLOBAL*)(" << col_ptr(c) << " + r);";
        } else {
          const char* vt = v.t == WS_T_F64 ? "ws::ws_f64x2" : "ws::ws_i64x2";
          o << " " << w << "a = *(const " << vt << " WS_GLOBAL*)(" << col_ptr(c) << " + r); " << w << "b = *(const " << vt
            << " WS_GLOBAL*)(" << col_ptr(c) << " + r + 2);";
        }
      }
    o << " }\n";
    o << "  __device__ Row row_of4(const Row4& x, int s) const { Row y;";
    if (split)
      for (int c : em.row_cols) {
        const ws_colview& v = env.cols[c];
        const std::string w = "x.w" + std::to_string(c);
        if (v.enc || v.t == WS_T_I32 || v.t == WS_T_I32N) o << " y.w" << c << " = " << w << "[s];";
        else o << " y.w" << c << " = s < 2 ? " << w << "a[s] : " << w << "b[s - 2];";
      }
    o << " y.pad_ = 0; return y; }\n";
    if (split) o << "  __device__ bool where_row(const Row& x, long long r) const {\n" << body.str() << "      return " << cond << ";\n  }\n";
    else o << "  __device__ bool where_row(const Row&, long long r) const { return where(r); }\n";
  }
  function(o, lw, "double key(int i, long long r)", keys, "double");
  function(o, lw, "long long ikey(int i, long long r)", keys, "int");
  function(o, lw, "double agg(int i, long long r)", aggs, "double");
  function(o, lw, "long long iagg(int i, long long r)", iaggs, "int");
  // Keys and aggregate arguments in a load half (load_vrow: the row's column words) and a compute
  // half (ikey_v, iagg_v, agg_v), for kernels that issue a row's loads a step before they use them.
  {
    std::set<int> cols;
    bool ok = true;
    const std::string fk = function_row(lw, "long long ikey_v(int i, const VRow& x, long long r)", keys, "int", cols, ok);
    const std::string fd = function_row(lw, "double key_v(int i, const VRow& x, long long r)", keys, "double", cols, ok);
    const std::string fi = function_row(lw, "long long iagg_v(int i, const VRow& x, long long r)", iaggs, "int", cols, ok);
    const std::string fa = function_row(lw, "double agg_v(int i, const VRow& x, long long r)", aggs, "double", cols, ok);
    o << "  static constexpr bool kVSplit = " << (ok ? "true" : "false") << ";\n";
    o << "  struct VRow {";
    if (ok)
      for (int c : cols) o << " " << (env.cols[c].enc ? "unsigned" : elem_type(env.cols[c].t)) << " w" << c << ";";
    o << " int pad_; };\n";
    o << "  __device__ void load_vrow(long long r, VRow& x) const {";
    if (ok)
      for (int c : cols)
        o << " " << row_word(c) << " = " << (env.cols[c].enc ? "ws::enc_word(" + enc_view(c) + ",r)" : col_ptr(c) + "[r]") << ";";
    o << " }\n";
    // the keys alone (KRow: the key columns' words), for passes that need no aggregate argument
    std::set<int> kcols;
    bool kok = true;
    const std::string kk = function_row(lw, "long long ikey_k(int i, const KRow& x, long long r)", keys, "int", kcols, kok);
    const std::string kd = function_row(lw, "double key_k(int i, const KRow& x, long long r)", keys, "double", kcols, kok);
    kok = kok && ok;
    o << "  static constexpr bool kKSplit = " << (kok ? "true" : "false") << ";\n";
    o << "  struct KRow {";
    if (kok)
      for (int c : kcols) o << " " << (env.cols[c].enc ? "unsigned" : elem_type(env.cols[c].t)) << " w" << c << ";";
    o << " int pad_; };\n";
    o << "  __device__ void load_krow(long long r, KRow& x) const {";
    if (kok)
      for (int c : kcols)
        o << " " << row_word(c) << " = " << (env.cols[c].enc ? "ws::enc_word(" + enc_view(c) + ",r)" : col_ptr(c) + "[r]") << ";";
    o << " }\n";
    if (kok) {
      o << kk << kd;
    } else {
      o << "  __device__ long long ikey_k(int i, const KRow&, long long r) const { return ikey(i, r); }\n";
      o << "  __device__ double key_k(int i, const KRow&, long long r) const { return key(i, r); }\n";
    }
    if (ok) {
      o << fk << fd << fi << fa;
    } else {
      o << "  __device__ long long ikey_v(int i, const VRow&, long long r) const { return ikey(i, r); }\n";
      o << "  __device__ double key_v(int i, const VRow&, long long r) const { return key(i, r); }\n";
      o << "  __device__ long long iagg_v(int i, const VRow&, long long r) const { return iagg(i, r); }\n";
      o << "  __device__ double agg_v(int i, const VRow&, long long r) const { return agg(i, r); }\n";
    }
  }
  o << "};\n";
  return o.str();
}

}  // namespace codegen
}  // namespace ws
