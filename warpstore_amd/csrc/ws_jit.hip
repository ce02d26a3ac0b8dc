// Per-query compiled plans: the engine's bytecode is translated into a typed C++ evaluator, compiled
// for the device's gfx950 target with hipRTC together with the shared kernel bodies, and cached in
// memory (keyed by kernel kind + evaluator source) and optionally on disk (keyed by hipRTC version,
// options, target and full source, with a checksum of the code object). Literal values are not part
// of the source; they are read from the env at run time, so one shape serves all its literals.
// Reference: src/jit.cu (NVRTC), src/jit.h.
#include <hip/hiprtc.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <thread>

#include "ws_internal.h"
#include "ws_jit_prelude.inc"  // kWsJitPrelude: ws_abi.h + ws_vm.h + ws_kernels.h (generated at build)

struct ws_jit_program {
  int id = 0;
  int fn_kind = WS_JF_FLAG;
  std::string kind, sql, source;
  double compile_ms = 0;
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t fn_part[2] = {nullptr, nullptr};  // grouped plans: partition count / scatter kernels
  hipFunction_t fn_batch = nullptr;               // dense plans: the batched (many queries) kernel
  hipFunction_t fn_chunk = nullptr;               // selection plans: the chunked kernel
  hipFunction_t fn_staged = nullptr;              // selection plans: the persistent chunked kernel (sets in LDS)
  int staged_static_lds = 0;                      // its static LDS bytes
  std::atomic<int> state{0};  // 0 compiling, 1 ready, 2 failed
  bool from_disk = false;
  int rows = 1;
  int64_t execs = 0;
  double total_ms = 0, last_ms = 0;
};

namespace ws {
namespace {

std::atomic<int> g_default_mode{WS_JIT_ON};
thread_local int t_mode = -1;
std::mutex g_mu;
std::map<std::string, std::shared_ptr<ws_jit_program>> g_cache;
// The same programs by their structural key (what the generated source depends on: kernel kind,
// plan, bytecode, column types), so a repeated plan skips generating the source to look it up.
std::unordered_map<std::string, std::shared_ptr<ws_jit_program>> g_by_shape;
std::string shape_key(const ws_expr_env& env, int ncols, const ws_agg_plan& plan, int fn) {
  std::string k;
  k.reserve(2048);
  auto put = [&](const void* q, size_t n) { k.append((const char*)q, n); };
  put(&fn, sizeof fn);
  put(&ncols, sizeof ncols);
  put(&plan.has_where, 1);
  put(&plan.where, sizeof plan.where);
  put(&plan.nkeys, sizeof plan.nkeys);
  put(plan.keys, sizeof plan.keys);
  put(&plan.naggs, sizeof plan.naggs);
  put(plan.aggs, sizeof plan.aggs);
  put(plan.ops, sizeof plan.ops);
  put(plan.agg_int, sizeof plan.agg_int);
  put(plan.agg_bits, sizeof plan.agg_bits);
  put(plan.agg_part, sizeof plan.agg_part);
  put(plan.agg_null, sizeof plan.agg_null);
  put(&plan.dense_priv, 1);
  put(plan.key_int, sizeof plan.key_int);
  put(env.code, sizeof env.code);
  for (int c = 0; c < ncols && c < WS_MAX_COLS; ++c) {
    const uint8_t t[3] = {env.cols[c].t, env.cols[c].scale, env.cols[c].enc};
    put(t, 3);
  }
  return k;
}
std::vector<std::shared_ptr<ws_jit_program>> g_by_id;
std::string g_disk_dir;
thread_local std::string t_sql;
thread_local ws_jit_query_info t_info;
thread_local std::vector<ws_jit_program*> t_used;

const char* kKernelMacro[] = {"WS_KERNEL_NOKEYS", "WS_KERNEL_GROUPED", "WS_KERNEL_DENSE", "WS_KERNEL_SELECT_JIT",
                              "WS_KERNEL_SELECT_JIT"};
const char* kKernelName[] = {"ws_q_nokeys", "ws_q_grouped", "ws_q_dense", "ws_q_select", "ws_q_select"};

// ---------------------------------------------------------------- persistent plan cache

uint64_t fnv64(const std::string& s, uint64_t h) {
  for (unsigned char c : s) h = (h ^ c) * 0x100000001b3ull;
  return h;
}
std::string disk_path(const std::string& key) {
  char name[64];
  snprintf(name, sizeof name, "%016llx%016llx.wsplan", (unsigned long long)fnv64(key, 0xcbf29ce484222325ull),
           (unsigned long long)fnv64(key, 0x84222325cbf29ce4ull));
  return g_disk_dir + "/" + name;
}
constexpr char kMagic[9] = "WSPLAN01";
bool disk_load(const std::string& path, const std::string& key, std::string& image) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::string magic(8, '\0');
  uint64_t kl = 0, il = 0;
  f.read(magic.data(), 8);
  f.read((char*)&kl, 8);
  if (!f || magic != kMagic || kl != key.size()) return false;
  std::string k(kl, '\0');
  f.read(k.data(), (std::streamsize)kl);
  f.read((char*)&il, 8);
  if (!f || k != key || il > (1ull << 30)) return false;
  image.assign(il, '\0');
  f.read(image.data(), (std::streamsize)il);
  uint64_t sum = 0;
  f.read((char*)&sum, 8);
  if (!f || sum != fnv64(image, 0xcbf29ce484222325ull)) {
    fprintf(stderr, "[warpstore_amd] plan cache file %s is corrupt; recompiling\n", path.c_str());
    return false;
  }
  return true;
}
void disk_store(const std::string& path, const std::string& key, const std::string& image) {
  const std::string tmp = path + ".tmp" + std::to_string((unsigned long long)std::hash<std::thread::id>{}(std::this_thread::get_id()));
  {
    std::ofstream f(tmp, std::ios::binary);
    const uint64_t kl = key.size(), il = image.size();
    f.write(kMagic, 8);
    f.write((const char*)&kl, 8);
    f.write(key.data(), (std::streamsize)kl);
    f.write((const char*)&il, 8);
    f.write(image.data(), (std::streamsize)il);
    const uint64_t sum = fnv64(image, 0xcbf29ce484222325ull);
    f.write((const char*)&sum, 8);
    if (!f) return;
  }
  std::error_code ec;
  std::filesystem::rename(tmp, path, ec);  // atomic publish
}

// ---------------------------------------------------------------- bytecode -> typed C++

struct Node {
  std::string e;
  bool is_int;
};
std::string as_d(const Node& n) { return n.is_int ? "((double)" + n.e + ")" : n.e; }
std::string as_i(const Node& n) { return n.is_int ? n.e : "((long long)" + n.e + ")"; }

// Rebuilds typed expression trees from the stack bytecode. Integer columns (INT, BIGINT, DATE,
// dictionary codes) and integer literals stay in int64 arithmetic (+ - * checked for overflow into
// env->ovf); anything NULL-able or fractional is double with three-valued helpers.
struct Translator {
  const ws_expr_env& env;
  int ncols;
  std::set<int> cols, iconsts, dconsts, encs, sets;  // encs: encoded integer columns read; sets: INSET bitmaps
  // Row mode: column reads come from a Row of values loaded beforehand (x.v<c>: the element, or the
  // 32-bit word holding an encoded code), so a kernel can issue the loads of a later step early.
  bool row_mode = false, split_ok = true;
  std::set<int> row_cols;

  std::string raw_int(int c) {
    const ws_colview& v = env.cols[c];
    if (v.enc) encs.insert(c);
    const std::string n = std::to_string(c);
    if (row_mode) {
      row_cols.insert(c);
      return v.enc ? "ws::enc_decode(e" + n + ",x.v" + n + ",r)" : "((long long)x.v" + n + ")";
    }
    return v.enc ? "ws::enc_get(e" + n + ",r)" : "((long long)c" + n + "[r])";
  }
  Node run(ws_prog p) {
    std::vector<Node> st;
    auto pop = [&]() {
      if (st.empty()) throw Error("codegen: malformed expression", WS_ERR_INTERNAL);
      Node n = st.back();
      st.pop_back();
      return n;
    };
    for (int pc = p.start; pc < p.start + p.len; ++pc) {
      if (pc < 0 || pc >= WS_MAX_CODE) throw Error("codegen: program out of range", WS_ERR_INTERNAL);
      const ws_instr in = env.code[pc];
      switch (in.op) {
        case WS_OP_COL: {
          if (in.arg >= ncols) throw Error("codegen: column slot out of range", WS_ERR_INTERNAL);
          cols.insert(in.arg);
          const ws_colview& v = env.cols[in.arg];
          const std::string n = std::to_string(in.arg);
          if (v.t == WS_T_F64) {
            if (row_mode && !v.enc) row_cols.insert(in.arg);
            if (row_mode && v.enc) split_ok = false;  // (a dictionary read: no row form)
            st.push_back({v.enc ? "ws::col_dbl(v" + n + ",r)" : row_mode ? "x.v" + n : "c" + n + "[r]", false});
          }
          else if (v.t == WS_T_I32N) st.push_back({"ws::null_if_i32(" + raw_int(in.arg) + ")", false});
          else if (v.t == WS_T_I64N) st.push_back({"ws::null_if_i64(" + raw_int(in.arg) + ")", false});
          else if (v.t == WS_T_DECIMAL) {
            char p10[64];
            snprintf(p10, sizeof p10, "%.1f", pow10i(v.scale));
            st.push_back({"ws::dec_of(" + raw_int(in.arg) + "," + p10 + ")", false});
          } else if (v.t == WS_T_VEC) {
            throw Error("codegen: vector column used as a scalar", WS_ERR_INTERNAL);
          } else st.push_back({raw_int(in.arg), true});
          break;
        }
        case WS_OP_COLRAW: {
          if (in.arg >= ncols) throw Error("codegen: column slot out of range", WS_ERR_INTERNAL);
          cols.insert(in.arg);
          if (env.cols[in.arg].t == WS_T_DECIMAL && !in.is_int) st.push_back({"ws::null_if_i64(" + raw_int(in.arg) + ")", false});
          else st.push_back({raw_int(in.arg), true});
          break;
        }
        case WS_OP_CONST:
          if (in.arg >= WS_MAX_CONSTS) throw Error("codegen: constant out of range", WS_ERR_INTERNAL);
          if (in.is_int) {
            iconsts.insert(in.arg);
            st.push_back({"ik" + std::to_string(in.arg), true});
          } else {
            dconsts.insert(in.arg);
            st.push_back({"k" + std::to_string(in.arg), false});
          }
          break;
        case WS_OP_ADD: case WS_OP_SUB: case WS_OP_MUL: {
          Node b = pop(), a = pop();
          if (a.is_int && b.is_int) {
            const char* f = in.op == WS_OP_ADD ? "ws::add_chk(" : in.op == WS_OP_SUB ? "ws::sub_chk(" : "ws::mul_chk(";
            st.push_back({std::string(f) + a.e + "," + b.e + ",&ovf_)", true});
          } else {
            const char* o = in.op == WS_OP_ADD ? "+" : in.op == WS_OP_SUB ? "-" : "*";
            st.push_back({"(" + as_d(a) + o + as_d(b) + ")", false});
          }
          break;
        }
        case WS_OP_DIV: {
          Node b = pop(), a = pop();
          st.push_back({"ws::divz(" + as_d(a) + "," + as_d(b) + ")", false});
          break;
        }
        case WS_OP_MOD: {
          Node b = pop(), a = pop();
          st.push_back({"fmod(" + as_d(a) + "," + as_d(b) + ")", false});
          break;
        }
        case WS_OP_NEG: {
          Node a = pop();
          if (a.is_int) st.push_back({"((long long)(0ull-(unsigned long long)" + a.e + "))", true});
          else st.push_back({"(-" + a.e + ")", false});
          break;
        }
        case WS_OP_EQ: case WS_OP_NE: case WS_OP_LT: case WS_OP_LE: case WS_OP_GT: case WS_OP_GE: {
          Node b = pop(), a = pop();
          static const char* ops[] = {"==", "!=", "<", "<=", ">", ">="};
          const int k = in.op - WS_OP_EQ;
          if (a.is_int && b.is_int) st.push_back({"((long long)(" + a.e + ops[k] + b.e + "))", true});  // never NULL
          else st.push_back({"ws::cmp3(" + std::to_string(k) + "," + as_d(a) + "," + as_d(b) + ")", false});
          break;
        }
        case WS_OP_AND: case WS_OP_OR: {
          Node b = pop(), a = pop();
          // Both operands are evaluated (bitwise & / |) unless one can raise the overflow flag:
          // their column loads then issue together instead of one memory round trip after the
          // other. Operands that can set env->ovf keep the short circuit.
          const bool fx = a.e.find("&ovf_") != std::string::npos || b.e.find("&ovf_") != std::string::npos;
          const char* join = in.op == WS_OP_AND ? (fx ? "&&" : "&") : (fx ? "||" : "|");
          if (a.is_int && b.is_int)
            st.push_back({"((long long)(((" + a.e + ")!=0)" + join + "((" + b.e + ")!=0)))", true});
          else st.push_back({std::string(in.op == WS_OP_AND ? "ws::and3(" : "ws::or3(") + as_d(a) + "," + as_d(b) + ")", false});
          break;
        }
        case WS_OP_NOT: {
          Node a = pop();
          if (a.is_int) st.push_back({"((long long)(" + a.e + "==0))", true});
          else st.push_back({"ws::not3(" + a.e + ")", false});
          break;
        }
        case WS_OP_INSET: {
          Node a = pop();
          const std::string s = std::to_string(in.arg);
          if (in.arg >= WS_MAX_SETS) throw Error("codegen: set out of range", WS_ERR_INTERNAL);
          sets.insert(in.arg);
          const std::string args = "(sl" + s + ",sw" + s + ",s" + s + ",sn" + s + ",";
          if (a.is_int) st.push_back({"ws::inset_s" + args + a.e + ")", true});
          else st.push_back({"ws::inset_sd" + args + a.e + ")", false});
          break;
        }
        case WS_OP_SELECT: {
          Node e = pop(), t = pop(), c = pop();
          const std::string cond = c.is_int ? "(" + c.e + "!=0)" : "ws::truthy(" + c.e + ")";
          if (t.is_int && e.is_int) st.push_back({"(" + cond + "?" + t.e + ":" + e.e + ")", true});
          else st.push_back({"(" + cond + "?" + as_d(t) + ":" + as_d(e) + ")", false});
          break;
        }
        case WS_OP_YEAR: {
          Node a = pop();
          if (a.is_int) st.push_back({"ws::year_of_days(" + a.e + ")", true});
          else st.push_back({"ws::call1(WS_FN_DT_YEAR,ws::call1(WS_FN_DATE_TO_DT," + a.e + "))", false});
          break;
        }
        case WS_OP_ISNULL: {
          Node a = pop();
          st.push_back({a.is_int ? std::string("0ll") : "((long long)ws::isnan_(" + a.e + "))", true});
          break;
        }
        case WS_OP_COALESCE: {
          Node b = pop(), a = pop();
          if (a.is_int) st.push_back(a);
          else st.push_back({"ws::coalesce2(" + a.e + "," + as_d(b) + ")", false});
          break;
        }
        case WS_OP_NULLIF: {
          Node b = pop(), a = pop();
          st.push_back({"ws::nullif2(" + as_d(a) + "," + as_d(b) + ")", false});
          break;
        }
        case WS_OP_CALL1: {
          Node a = pop();
          st.push_back({"ws::call1(" + std::to_string(in.arg) + "," + as_d(a) + ")", false});
          break;
        }
        case WS_OP_CALL2: {
          Node b = pop(), a = pop();
          st.push_back({"ws::call2(" + std::to_string(in.arg) + "," + as_d(a) + "," + as_d(b) + ")", false});
          break;
        }
        case WS_OP_VDOT: case WS_OP_VEUC: {
          const int slot = in.arg & 255, m = in.arg >> 8;
          if (slot >= ncols || m >= WS_MAX_SETS) throw Error("codegen: vector operand out of range", WS_ERR_INTERNAL);
          cols.insert(slot);
          const std::string ms = std::to_string(m);
          st.push_back({std::string(in.op == WS_OP_VDOT ? "ws::vdot_d(" : "ws::veuc_d(") + "(const float*)v" +
                            std::to_string(slot) + ".p,r,env->maps[" + ms + "],env->map_size[" + ms + "])",
                        false});
          break;
        }
        case WS_OP_INSORTED: {
          Node a = pop();
          const std::string s = std::to_string(in.arg);
          st.push_back({"ws::insorted_d(env->maps[" + s + "],env->map_size[" + s + "]," + as_d(a) + ")", false});
          break;
        }
        case WS_OP_DICTMAP: {
          Node a = pop();
          const std::string s = std::to_string(in.arg);
          st.push_back({"ws::dictmap_d(env->maps[" + s + "],env->map_size[" + s + "]," + as_d(a) + ")", false});
          break;
        }
        default: throw Error("codegen: unknown opcode", WS_ERR_INTERNAL);
      }
    }
    if (st.size() != 1) throw Error("codegen: malformed expression", WS_ERR_INTERNAL);
    return st[0];
  }
};

const char* elem_of(int t) {  // element type of a raw column
  return t == WS_T_F64 ? "double" : (t == WS_T_I32 || t == WS_T_I32N) ? "int" : "long long";
}
const char* ctype_of(int t) {  // global-memory pointers (loads issue as global, not flat)
  return (t == WS_T_I32 || t == WS_T_I32N) ? "const int WS_GLOBAL*"
         : (t == WS_T_F64)                 ? "const double WS_GLOBAL*"
         : t == WS_T_VEC                   ? "const float WS_GLOBAL*"
                                           : "const long long WS_GLOBAL*";
}

int dense_kind_of(int op, bool is_int) {
  return op == WS_AGG_COUNT ? WS_DK_COUNT
         : op == WS_AGG_SUM ? (is_int ? WS_DK_SUM_I : WS_DK_SUM_D)
         : op == WS_AGG_MIN ? (is_int ? WS_DK_MIN_I : WS_DK_MIN_D)
                            : (is_int ? WS_DK_MAX_I : WS_DK_MAX_D);
}

// The evaluator struct for one plan: shape queries are compile-time constants, so the unrolled
// loops of the kernel bodies collapse to the work this query needs.
std::string gen_evaluator(const ws_expr_env& env, int ncols, const ws_agg_plan& plan, int* rows_out) {
  Translator tr{env, ncols, {}, {}, {}};
  const Node where = plan.has_where ? tr.run(plan.where) : Node{"1", true};
  std::vector<Node> keys, aggs;
  for (int k = 0; k < plan.nkeys; ++k) keys.push_back(tr.run(plan.keys[k]));
  for (int a = 0; a < plan.naggs; ++a) aggs.push_back(plan.ops[a] == WS_AGG_COUNT ? Node{"1ll", true} : tr.run(plan.aggs[a]));
  const int nk = std::max(plan.nkeys, 1), na = std::max(plan.naggs, 1);
  const int rows = std::max(1, std::min(4, 12 / (nk + na)));  // rows per thread per step (registers)
  *rows_out = rows;
  std::ostringstream o;
  o << "struct QEval {\n";
  // A program that is one column or one literal has no side effect (no checked-arithmetic flag),
  // so the partition passes may evaluate it for rows the WHERE then drops.
  auto pure = [&](ws_prog p) {
    return p.len == 1 && (env.code[p.start].op == WS_OP_COL || env.code[p.start].op == WS_OP_CONST);
  };
  bool pure_keys = true, pure_vals = true;
  for (int k = 0; k < plan.nkeys; ++k) pure_keys = pure_keys && pure(plan.keys[k]);
  for (int a = 0; a < plan.naggs; ++a) pure_vals = pure_vals && (plan.ops[a] == WS_AGG_COUNT || pure(plan.aggs[a]));
  o << "  static constexpr int kKeys = " << nk << ", kAggs = " << na << ", kRows = " << rows << ";\n";
  o << "  static constexpr bool kPureKeys = " << (pure_keys ? "true" : "false") << ", kPureVals = "
    << (pure_vals ? "true" : "false") << ";\n";
  o << "  static constexpr bool kGlobalCols = true;\n";
  o << "  static constexpr bool kPureWhere = " << (where.e.find("&ovf_") == std::string::npos ? "true" : "false") << ";\n";
  o << "  const ws_expr_env* env;\n";
  // Overflow of checked integer arithmetic: a per-thread flag, stored to env->ovf once when the
  // evaluator goes out of scope (a possible store in the row loop would pin every later load
  // behind it).
  o << "  mutable int ovf_ = 0;\n";
  o << "  __device__ ~QEval() { if (ovf_ && env->ovf) *env->ovf = 1; }\n";
  for (int c : tr.cols) o << "  " << ctype_of(env.cols[c].t) << " c" << c << ";\n  ws_colview v" << c << ";\n";
  for (int c : tr.encs) o << "  ws::EncView e" << c << ";\n";
  for (int k : tr.iconsts) o << "  long long ik" << k << ";\n";
  for (int k : tr.dconsts) o << "  double k" << k << ";\n";
  for (int k : tr.sets)
    o << "  const unsigned WS_GLOBAL* s" << k << ";\n  long long sn" << k << ";\n  const unsigned WS_LOCAL* sl" << k << ";\n  int sw" << k
      << ";\n";
  o << "  __device__ QEval(const ws::ScanDesc& d) : env(d.env) {\n";
  // literals and set bitmaps are the same for every lane: scalar registers
  for (int k : tr.iconsts) o << "    ik" << k << " = ws::uni((long long)ws::dbits(env->consts[" << k << "]));\n";
  for (int k : tr.dconsts) o << "    k" << k << " = ws::bitsd(ws::uni((long long)ws::dbits(env->consts[" << k << "])));\n";
  // an empty set reads its own size word (always readable; every bit is masked off)
  for (int k : tr.sets)
    o << "    sn" << k << " = ws::uni((long long)(env->set_size[" << k << "] > 0 ? env->set_size[" << k << "] : 0)); s" << k << " = (const unsigned WS_GLOBAL*)ws::uni(sn" << k
      << " > 0 ? env->sets[" << k << "] : (const unsigned*)&env->set_size[" << k << "]); sl" << k << " = ws::lds_spare_word(); sw" << k
      << " = 0;\n";
  o << "  }\n";
  // A kernel with LDS to spare copies the sets' leading words there, in order, within `words`.
  o << "  __device__ void stage_sets(unsigned* lds, int words) {\n    int off = 0, w = 0;\n";
  for (int k : tr.sets)
    o << "    w = ws::stage_set(lds, off, words, s" << k << ", sn" << k << "); if (w > 0) { sl" << k
      << " = (const unsigned WS_LOCAL*)(lds + off); sw" << k << " = w; } off += w;\n";
  o << "    (void)off; (void)w;\n  }\n  __device__ void bind(const ws_colview* c) {\n";
  for (int c : tr.cols)
    o << "    v" << c << " = ws::uniform_view(c[" << c << "]); c" << c << " = (" << ctype_of(env.cols[c].t) << ")v" << c << ".p;\n";
  for (int c : tr.encs) o << "    e" << c << " = ws::enc_view(v" << c << ");\n";
  o << "  }\n";
  o << "  __device__ int nkeys() const { return " << plan.nkeys << "; }\n";
  o << "  __device__ int naggs() const { return " << plan.naggs << "; }\n";
  o << "  __device__ bool has_where() const { return " << (plan.has_where ? "true" : "false") << "; }\n";
  o << "  __device__ int op(int a) const { switch (a) {";
  for (int a = 0; a < plan.naggs; ++a) o << " case " << a << ": return " << (int)plan.ops[a] << ";";
  o << " } return 0; }\n";
  o << "  __device__ bool agg_int(int a) const { switch (a) {";
  for (int a = 0; a < plan.naggs; ++a) o << " case " << a << ": return " << (plan.agg_int[a] ? "true" : "false") << ";";
  o << " } return false; }\n";
  o << "  __device__ bool agg_bits(int a) const { switch (a) {";
  for (int a = 0; a < plan.naggs; ++a) o << " case " << a << ": return " << (plan.agg_bits[a] ? "true" : "false") << ";";
  o << " } return false; }\n";
  o << "  __device__ int agg_part(int a) const { switch (a) {";
  for (int a = 0; a < plan.naggs; ++a) o << " case " << a << ": return " << (int)plan.agg_part[a] << ";";
  o << " } return 0; }\n";
  o << "  __device__ int agg_null(int a) const { switch (a) {";
  for (int a = 0; a < plan.naggs; ++a) o << " case " << a << ": return " << (int)plan.agg_null[a] << ";";
  o << " } return 0; }\n";
  o << "  __device__ bool key_int(int k) const { switch (k) {";
  for (int k = 0; k < plan.nkeys; ++k) o << " case " << k << ": return " << (plan.key_int[k] ? "true" : "false") << ";";
  o << " } return false; }\n";
  o << "  __device__ uint8_t dkind(int a, const ws_dense_spec&) const { switch (a) {";
  for (int a = 0; a < plan.naggs; ++a) o << " case " << a << ": return " << dense_kind_of(plan.ops[a], plan.agg_int[a] != 0) << ";";
  o << " } return 0; }\n";
  o << "  __device__ bool where(long long r) const { return " << (where.is_int ? where.e + " != 0" : "ws::truthy(" + where.e + ")")
    << "; }\n";
  // The WHERE split in a load half (load_row: the row's column words) and a compute half
  // (where_row), for kernels that load a later step's rows before computing this one.
  {
    Translator trr{env, ncols, {}, {}, {}};
    trr.row_mode = true;
    const Node w2 = plan.has_where ? trr.run(plan.where) : Node{"1", true};
    // (its sets, constants and encoded columns are among the main translation's members)
    const bool split = plan.has_where && trr.split_ok && !trr.row_cols.empty();
    o << "  static constexpr bool kSplit = " << (split ? "true" : "false") << ";\n";
    o << "  struct Row {";
    if (split)
      for (int c : trr.row_cols) o << " " << (env.cols[c].enc ? "unsigned" : elem_of(env.cols[c].t)) << " v" << c << ";";
    o << " int unused_; };\n";
    o << "  __device__ void load_row(long long r, Row& x) const {";
    if (split)
      for (int c : trr.row_cols)
        o << " x.v" << c << " = " << (env.cols[c].enc ? "ws::enc_word(e" + std::to_string(c) + ",r)" : "c" + std::to_string(c) + "[r]")
          << ";";
    o << " }\n";
    // Four consecutive rows (r a multiple of 4) with 16-byte loads of raw columns: a wave's load
    // instruction then carries 1 KiB; rows4_ok tells whether every raw column base is 16-byte
    // aligned (encoded columns read their code words one by one).
    o << "  __device__ bool rows4_ok() const { return true";
    if (split)
      for (int c : trr.row_cols)
        if (!env.cols[c].enc) o << " && (((unsigned long long)c" << c << " & 15) == 0)";
    o << "; }\n";
    o << "  __device__ void load_rows4(long long r, Row* x) const {";
    if (split)
      for (int c : trr.row_cols) {
        const ws_colview& v = env.cols[c];
        const std::string n = std::to_string(c);
        if (v.enc) {
          for (int q = 0; q < 4; ++q) o << " x[" << q << "].v" << n << " = ws::enc_word(e" << n << ",r+" << q << ");";
        } else if (v.t == WS_T_I32 || v.t == WS_T_I32N) {
          o << " { const ws::ws_i32x4 q = *(const ws::ws_i32x4 WS_GLOBAL*)(c" << n << " + r); x[0].v" << n << " = q.x; x[1].v" << n
            << " = q.y; x[2].v" << n << " = q.z; x[3].v" << n << " = q.w; }";
        } else {
          const char* vt = v.t == WS_T_F64 ? "ws::ws_f64x2" : "ws::ws_i64x2";
          o << " { const " << vt << " q0 = *(const " << vt << " WS_GLOBAL*)(c" << n << " + r), q1 = *(const " << vt
            << " WS_GLOBAL*)(c" << n << " + r + 2); x[0].v" << n << " = q0.x; x[1].v" << n << " = q0.y; x[2].v" << n
            << " = q1.x; x[3].v" << n << " = q1.y; }";
        }
      }
    o << " }\n";
    o << "  __device__ bool where_row(const Row& x, long long r) const { return "
      << (split ? (w2.is_int ? w2.e + " != 0" : "ws::truthy(" + w2.e + ")") : std::string("where(r)")) << "; }\n";
  }
  o << "  __device__ double key(int k, long long r) const { switch (k) {";
  for (size_t k = 0; k < keys.size(); ++k) o << "\n    case " << k << ": return " << as_d(keys[k]) << ";";
  o << " } return 0; }\n";
  o << "  __device__ long long ikey(int k, long long r) const { switch (k) {";
  for (size_t k = 0; k < keys.size(); ++k) o << "\n    case " << k << ": return " << as_i(keys[k]) << ";";
  o << " } return 0; }\n";
  o << "  __device__ double agg(int a, long long r) const { switch (a) {";
  for (size_t a = 0; a < aggs.size(); ++a) o << "\n    case " << a << ": return " << as_d(aggs[a]) << ";";
  o << " } return 0; }\n";
  o << "  __device__ long long iagg(int a, long long r) const { switch (a) {";
  for (size_t a = 0; a < aggs.size(); ++a) {
    // A nullable int column: its raw value, NULL sentinel included (the kernel skips it).
    std::string e = as_i(aggs[a]);
    if (plan.agg_null[a]) e = tr.raw_int(env.code[plan.aggs[a].start].arg);
    o << "\n    case " << a << ": return " << e << ";";
  }
  o << " } return 0; }\n};\n";
  if (plan.agg_null[0] || plan.nkeys > WS_MAX_KEYS || plan.naggs > WS_MAX_AGGS)
    throw Error("codegen: plan out of range", WS_ERR_INTERNAL);
  for (int a = 0; a < plan.naggs; ++a)
    if (plan.agg_null[a] && !(plan.aggs[a].len == 1 && env.code[plan.aggs[a].start].op == WS_OP_COL))
      throw Error("codegen: nullable aggregate argument is not a column", WS_ERR_INTERNAL);
  // (The columns named by agg_null arguments were added by the translator above.)
  return o.str();
}

struct PlanBuild {
  std::string src, key, path;
  std::vector<std::string> opts;
};
PlanBuild plan_build(const ws_jit_program& p) {
  PlanBuild b;
  b.src = std::string(kWsJitPrelude) + "\n" + p.source + kKernelMacro[p.fn_kind] + "(QEval, " + kKernelName[p.fn_kind] + ")\n";
  b.opts = {"--offload-arch=" + arch_name(), "-O3", "-std=c++17", "-munsafe-fp-atomics", "-Xclang", "-target-feature",
            "-Xclang", "-packed-fp32-ops", "-DNDEBUG"};
  if (!g_disk_dir.empty()) {
    int maj = 0, min = 0;
    hiprtcVersion(&maj, &min);
    b.key = "hiprtc " + std::to_string(maj) + "." + std::to_string(min) + "\n";
    for (auto& o : b.opts) b.key += o + "\n";
    b.key += b.src;
    b.path = disk_path(b.key);
  }
  return b;
}

void load_image(ws_jit_program& p, const std::string& image) {
  WS_CK(hipSetDevice(0));
  hipModule_t mod;
  WS_CK(hipModuleLoadData(&mod, image.data()));
  p.module = mod;
  hipFunction_t f;
  WS_CK(hipModuleGetFunction(&f, mod, kKernelName[p.fn_kind]));
  p.fn = f;
  if (p.fn_kind == WS_JF_DENSE) WS_CK(hipModuleGetFunction(&p.fn_batch, mod, (std::string(kKernelName[p.fn_kind]) + "_batch").c_str()));
  if (p.fn_kind == WS_JF_FLAG || p.fn_kind == WS_JF_FLAGM) {
    WS_CK(hipModuleGetFunction(&p.fn_chunk, mod, (std::string(kKernelName[p.fn_kind]) + "_chunk").c_str()));
    WS_CK(hipModuleGetFunction(&p.fn_staged, mod, (std::string(kKernelName[p.fn_kind]) + "_staged").c_str()));
    WS_CK(hipFuncGetAttribute(&p.staged_static_lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, p.fn_staged));
  }
  if (p.fn_kind == WS_JF_GROUPED) {
    const std::string base = kKernelName[p.fn_kind];
    WS_CK(hipModuleGetFunction(&p.fn_part[0], mod, (base + "_pcount").c_str()));
    WS_CK(hipModuleGetFunction(&p.fn_part[1], mod, (base + "_pscatter").c_str()));
  }
}

bool try_disk(ws_jit_program& p) {
  if (g_disk_dir.empty()) return false;
  const auto t0 = std::chrono::steady_clock::now();
  PlanBuild b = plan_build(p);
  std::string image;
  if (!disk_load(b.path, b.key, image)) return false;
  try {
    load_image(p, image);
  } catch (const std::exception&) {
    return false;  // e.g. a stale file for another runtime: recompile
  }
  p.from_disk = true;
  p.compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return true;
}

// Compiles the image only (no device needed); returns false and logs on failure.
bool compile_image(const ws_jit_program& p, const PlanBuild& b, std::string& image, std::string& log) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, b.src.c_str(), "warpstore_query.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) return false;
  std::vector<const char*> opts;
  for (auto& o : b.opts) opts.push_back(o.c_str());
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  size_t n = 0;
  hiprtcGetProgramLogSize(prog, &n);
  log.assign(n, '\0');
  if (n) hiprtcGetProgramLog(prog, log.data());
  if (rc != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return false;
  }
  hiprtcGetCodeSize(prog, &n);
  image.assign(n, '\0');
  hiprtcGetCode(prog, image.data());
  hiprtcDestroyProgram(&prog);
  (void)p;
  return true;
}

bool compile(ws_jit_program& p) {
  PlanBuild b = plan_build(p);
  if (const char* d = getenv("WARPSTORE_DUMP")) {
    FILE* f = fopen((std::string(d) + "/plan" + std::to_string(p.id) + ".hip").c_str(), "w");
    if (f) fwrite(b.src.data(), 1, b.src.size(), f), fclose(f);
  }
  std::string image, log;
  const auto t0 = std::chrono::steady_clock::now();
  if (!compile_image(p, b, image, log)) {
    fprintf(stderr, "[warpstore_amd] codegen failed, using the interpreter:\n%s\n", log.c_str());
    return false;
  }
  const double tc = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (!b.path.empty()) disk_store(b.path, b.key, image);
  load_image(p, image);
  if (getenv("WARPSTORE_TRACE"))
    fprintf(stderr, "[codegen] plan #%d %s: hiprtc %.1f ms, %zu source bytes\n", p.id, kKernelName[p.fn_kind], tc, b.src.size());
  return true;
}

void compile_and_publish(const std::shared_ptr<ws_jit_program>& p) {
  const auto t0 = std::chrono::steady_clock::now();
  bool ok = false;
  try {
    ok = compile(*p);
  } catch (const std::exception& e) {
    fprintf(stderr, "[warpstore_amd] codegen failed: %s\n", e.what());
  }
  p->compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  p->state.store(ok ? 1 : 2, std::memory_order_release);
}

int mode_now() { return t_mode >= 0 ? t_mode : g_default_mode.load(); }

}  // namespace

void* jit_function(const ws_jit_program* p, int fn) {
  if (!p || p->state.load(std::memory_order_acquire) != 1) return nullptr;
  if (fn == kJitPartCount || fn == kJitPartScatter)
    return p->fn_kind == WS_JF_GROUPED ? (void*)p->fn_part[fn == kJitPartScatter ? 1 : 0] : nullptr;
  if (fn == kJitBatchDense) return p->fn_kind == WS_JF_DENSE ? (void*)p->fn_batch : nullptr;
  if (fn == kJitSelectChunk) return p->fn_kind == WS_JF_FLAG || p->fn_kind == WS_JF_FLAGM ? (void*)p->fn_chunk : nullptr;
  if (fn == kJitSelectStaged) return p->fn_kind == WS_JF_FLAG || p->fn_kind == WS_JF_FLAGM ? (void*)p->fn_staged : nullptr;
  const bool sel = (fn == WS_JF_FLAG || fn == WS_JF_FLAGM) && (p->fn_kind == WS_JF_FLAG || p->fn_kind == WS_JF_FLAGM);
  return (p->fn_kind == fn || sel) ? (void*)p->fn : nullptr;
}
int jit_rows(const ws_jit_program* p) { return p ? p->rows : 1; }
int jit_staged_static_lds(const ws_jit_program* p) { return p ? p->staged_static_lds : 0; }
const char* jit_prof_name(const ws_jit_program* p, const char* base) {
  if (!p || p->kind.empty()) return base;
  static std::mutex mu;
  static std::set<std::string> names;  // kept for the process: profile slots hold the pointer
  std::lock_guard<std::mutex> g(mu);
  return names.insert(std::string(base) + "[" + p->kind + "]").first->c_str();
}
void jit_launch_raw(void* fn, unsigned grid, unsigned block, size_t shmem, void** args, const char* prof_name) {
  ProfScope ps(prof_name);
  WS_CK(hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, block, 1, 1, (unsigned)shmem, stream(), args, nullptr));
}

// Generates the evaluator and the full source for a plan without compiling (tests, tooling).
std::string jit_source_for(const ws_expr_env& env, int ncols, const ws_agg_plan& plan, int fn) {
  int rows = 1;
  return std::string(kKernelMacro[fn]) + "\n" + gen_evaluator(env, ncols, plan, &rows);
}

}  // namespace ws

using namespace ws;

extern "C" {

int ws_jit_enabled(void) { return mode_now() != WS_JIT_OFF; }
int ws_jit_mode(void) { return mode_now(); }
void ws_jit_set_mode(int m) { t_mode = m; }
void ws_jit_set_default_mode(int m) { g_default_mode = m; }
void ws_jit_set_cache_dir(const char* dir) {
  if (!dir || !*dir) {
    g_disk_dir.clear();
    return;
  }
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  g_disk_dir = ec ? "" : dir;
  if (ec) fprintf(stderr, "[warpstore_amd] plan cache directory %s unusable (%s)\n", dir, ec.message().c_str());
}

int ws_prog_is_int(const ws_expr_env* env, ws_prog p, int* out) {
  return guard([&] {
    Translator tr{*env, WS_MAX_COLS, {}, {}, {}};
    *out = tr.run(p).is_int ? 1 : 0;
  });
}

int ws_jit_get(const ws_expr_env* env, int ncols, const ws_agg_plan* plan, int fn, const char* kind,
               const ws_jit_program** out) {
  return guard([&] {
    *out = nullptr;
    const int mode = mode_now();
    if (mode == WS_JIT_OFF) return;
    if (fn < 0 || fn > WS_JF_FLAGM) throw Error("jit: bad kernel kind", WS_ERR_INTERNAL);
    const std::string shape = shape_key(*env, ncols, *plan, fn);
    std::shared_ptr<ws_jit_program> p;
    bool fresh = false;
    {
      std::lock_guard<std::mutex> g(g_mu);
      auto it = g_by_shape.find(shape);
      if (it != g_by_shape.end()) p = it->second;
    }
    if (!p) {
      int rows = 1;
      const std::string ev = gen_evaluator(*env, ncols, *plan, &rows);
      const std::string key = std::string(kKernelMacro[fn]) + "\n" + ev;
      {
        std::lock_guard<std::mutex> g(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
          p = it->second;
          g_by_shape.emplace(shape, p);
        } else {
          p = std::make_shared<ws_jit_program>();
          p->id = (int)g_by_id.size() + 1;
          p->fn_kind = fn;
          p->kind = kind ? kind : "";
          p->sql = t_sql;
          p->source = ev;
          p->rows = rows;
          g_by_id.push_back(p);
          g_cache.emplace(key, p);
          g_by_shape.emplace(shape, p);
          fresh = true;
        }
      }
    }
    ensure_init();
    if (fresh && try_disk(*p)) {
      p->state.store(1, std::memory_order_release);
      ++t_info.from_disk;
    } else if (fresh) {
      if (mode == WS_JIT_SYNC) {
        compile_and_publish(p);
        if (p->state.load() == 1) {
          ++t_info.compiled;
          t_info.compile_ms += p->compile_ms;
        }
      } else {
        std::thread([p] { compile_and_publish(p); }).detach();  // interpret first, compile in background
      }
    } else if (mode == WS_JIT_SYNC) {
      while (p->state.load(std::memory_order_acquire) == 0) std::this_thread::yield();  // compiled by another thread
    }
    const int st = p->state.load(std::memory_order_acquire);
    if (st != 1) {
      if (st == 0) ++t_info.background;
      return;
    }
    if (!fresh) ++t_info.hits;
    std::lock_guard<std::mutex> g(g_mu);
    if (std::find(t_used.begin(), t_used.end(), p.get()) == t_used.end()) {
      t_used.push_back(p.get());
      if (t_info.nids < 16) t_info.ids[t_info.nids++] = p->id;
    }
    *out = p.get();
  });
}

int ws_jit_program_id(const ws_jit_program* p) { return p ? p->id : 0; }

int ws_jit_launch(const ws_jit_program* p, int fn, unsigned grid, unsigned block, size_t shmem, void** args) {
  return guard([&] {
    void* f = jit_function(p, fn);
    if (!f) throw Error("jit: kernel not compiled for this kind", WS_ERR_INTERNAL);
    jit_launch_raw(f, grid, block, shmem, args, "jit_kernel");
  });
}

void ws_jit_begin_query(const char* sql) {
  t_sql.clear();
  for (const char* c = sql ? sql : ""; *c; ++c) {
    const bool sp = *c == ' ' || *c == '\n' || *c == '\t' || *c == '\r';
    if (sp && (t_sql.empty() || t_sql.back() == ' ')) continue;
    t_sql += sp ? ' ' : *c;
  }
  while (!t_sql.empty() && t_sql.back() == ' ') t_sql.pop_back();
  t_used.clear();
  t_info = ws_jit_query_info{};
}

void ws_jit_end_query(double query_ms, ws_jit_query_info* info) {
  std::lock_guard<std::mutex> g(g_mu);
  for (ws_jit_program* p : t_used) {
    ++p->execs;
    p->total_ms += query_ms;
    p->last_ms = query_ms;
  }
  t_used.clear();
  if (info) *info = t_info;
}

int ws_jit_plan_cache(ws_plan_cache_row* rows, int cap) {
  std::lock_guard<std::mutex> g(g_mu);
  int i = 0;
  for (auto& p : g_by_id) {
    if (i < cap) {
      ws_plan_cache_row& r = rows[i];
      const int st = p->state.load();
      std::string kind = st == 1 ? p->kind : p->kind + (st == 0 ? " (compiling)" : " (compile failed)");
      snprintf(r.kind, sizeof r.kind, "%s", kind.c_str());
      snprintf(r.origin, sizeof r.origin, "%s", st != 1 ? "" : p->from_disk ? "disk" : "compiled");
      snprintf(r.sql, sizeof r.sql, "%s", p->sql.c_str());
      r.id = p->id;
      r.compile_ms = p->compile_ms;
      r.execs = p->execs;
      r.avg_ms = p->execs ? p->total_ms / p->execs : 0;
      r.last_ms = p->last_ms;
      r.source_bytes = p->source.size();
    }
    ++i;
  }
  return i;
}

int ws_jit_source_of(int id, char* buf, size_t cap) {
  std::lock_guard<std::mutex> g(g_mu);
  for (auto& p : g_by_id)
    if (p->id == id) {
      if (buf && cap) snprintf(buf, cap, "%s", p->source.c_str());
      return (int)p->source.size();
    }
  return 0;
}

void ws_jit_clear_cache(void) {
  // Modules stay loaded: another thread may still be running one of them.
  std::lock_guard<std::mutex> g(g_mu);
  g_by_id.clear();
  g_cache.clear();
  g_by_shape.clear();
}

// Host-only check that a plan's generated source compiles for gfx950 (no device needed): returns
// 0 and the code object size, or the hipRTC log in buf.
int ws_jit_compile_check(const ws_expr_env* env, int ncols, const ws_agg_plan* plan, int fn, const char* arch, char* buf,
                         size_t cap, size_t* code_bytes) {
  return guard([&] {
    ws_jit_program p;
    p.fn_kind = fn;
    int rows = 1;
    p.source = gen_evaluator(*env, ncols, *plan, &rows);
    PlanBuild b;
    b.src = std::string(kWsJitPrelude) + "\n" + p.source + kKernelMacro[fn] + "(QEval, " + kKernelName[fn] + ")\n";
    b.opts = {std::string("--offload-arch=") + (arch ? arch : "gfx950"), "-O3", "-std=c++17", "-munsafe-fp-atomics",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-DNDEBUG"};
    std::string image, log;
    const bool ok = compile_image(p, b, image, log);
    if (buf && cap) snprintf(buf, cap, "%s", ok ? p.source.c_str() : log.c_str());
    if (!ok) throw Error("hipRTC compile failed: " + log.substr(0, 2000), WS_ERR_INTERNAL);
    if (code_bytes) *code_bytes = image.size();
  });
}

}  // extern "C"
