// Engine-side drivers of the measured query shapes, written against the drop-in gdb:: operator
// surface (include/warpstore_amd/gpu.h, jit.h) exactly the way the engine's planner calls it:
//   - storage: GPU-resident segments of 2^20 rows per column (reference storage.h:33), optional
//     frame-of-reference encoding per segment column (storage.cu:1138-1170);
//   - bytecode compile of WHERE / keys / aggregate arguments into an ExprEnv (engine_core.h:1485);
//   - aggregate slot layout, dense-vs-hash sizing and the overflow retry (engine_select.cpp:647-724,
//     1130-1235, 1493-1613);
//   - COUNT(DISTINCT) in two GPU phases and post-aggregation ORDER BY .. LIMIT on the GPU
//     (engine_select.cpp:1239-1373, 1403-1492);
//   - materialize with a join-filter bitmap, hash join and the join output gathers
//     (engine_core.h:4345-4611, engine_select.cpp:1015-1041).
// The host side above the C ABI: it owns no kernels and calls only the public surface.
#define WARPSTORE_AMD_STANDALONE 1
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "warpstore_amd/gpu.h"
#include "warpstore_amd/jit.h"

extern "C" void wsg_c4_shapes(uint64_t seed, int64_t nq, int64_t* rows, int64_t* distinct, int64_t* theta);
extern "C" void wsg_c4_data(uint64_t seed, int64_t q, int64_t rows, int64_t distinct, int32_t* key, int64_t* val);
extern "C" int wsq_table_load(int ncols, const int* types, const int* encode, int64_t n, const void* const* host, void** out);
extern "C" const char* wsq_last_error(void);

namespace {

using namespace gdb;
constexpr int64_t kSegRows = 1 << 20;

// ------------------------------------------------------------------ storage

struct SegCol {
  void* dev = nullptr;  // raw column, or header + codes when enc != 0
  uint8_t enc = 0;
};
struct DevColumn {
  int t = WS_T_I32;  // type code of the column's views
  std::vector<SegCol> segs;
  int64_t mn = INT64_MAX, mx = INT64_MIN;  // min / max of integer columns (colRange)
  ws_col_stats stats{};                    // merged segment statistics (computed at load)
  bool unique = false;  // integer column without repeated values (a key), found at load when its span < 2^27
};
struct DevTable {
  std::vector<int64_t> seg_rows;
  std::vector<DevColumn> cols;
  int64_t rows = 0;
  ~DevTable() {
    for (auto& c : cols)
      for (auto& s : c.segs) devFree(s.dev);
  }
  ColView view(int c, size_t seg) const {
    ColView v{};
    v.p = cols[c].segs[seg].dev;
    v.t = (uint8_t)cols[c].t;
    v.enc = cols[c].segs[seg].enc;
    v.off = 0;
    return v;
  }
};

int width_of(int t) { return (t == WS_T_I32 || t == WS_T_I32N) ? 4 : 8; }

// ------------------------------------------------------------------ bytecode compiler (subset)

struct Comp {
  ExprEnv env;
  int ncode = 0, nconst = 0, ncols = 0;
  Comp() { std::memset(&env, 0, sizeof env); }
  int slot(int t, uint8_t enc_hint = 0) {
    const int s = ncols++;
    env.cols[s].t = (uint8_t)t;
    env.cols[s].enc = enc_hint;
    return s;
  }
  int const_int(int64_t v) {
    double d;
    std::memcpy(&d, &v, 8);
    for (int i = 0; i < nconst; ++i)
      if (std::memcmp(&env.consts[i], &d, 8) == 0) return i;
    env.consts[nconst] = d;
    return nconst++;
  }
  int const_dbl(double v) {
    for (int i = 0; i < nconst; ++i)
      if (env.consts[i] == v) return i;
    env.consts[nconst] = v;
    return nconst++;
  }
  struct I {
    uint8_t op, is_int;
    uint16_t arg;
  };
  Prog prog(std::initializer_list<I> code) {
    Prog p;
    p.start = (int16_t)ncode;
    for (const I& i : code) env.code[ncode++] = ws_instr{i.op, i.is_int, i.arg};
    p.len = (int16_t)(ncode - p.start);
    return p;
  }
  I col(int s) { return I{WS_OP_COL, 0, (uint16_t)s}; }
  I ilit(int64_t v) { return I{WS_OP_CONST, 1, (uint16_t)const_int(v)}; }
  static I op(int o) { return I{(uint8_t)o, 0, 0}; }
};

// Per batch: the slots of e (the plan's environment, copied once per loop of batches) bound to
// segment seg's columns (bindView, engine_core.h:4073).
void bind_segment(ExprEnv& e, const Comp& c, const DevTable& t, const std::vector<int>& slot_col, size_t seg) {
  for (size_t s = 0; s < slot_col.size(); ++s) {
    ColView v = t.view(slot_col[s], seg);
    v.t = c.env.cols[s].t;
    e.cols[s] = v;
  }
}

int dense_kind(int op, bool is_int) {
  return op == WS_AGG_COUNT ? WS_DK_COUNT
         : op == WS_AGG_SUM ? (is_int ? WS_DK_SUM_I : WS_DK_SUM_D)
         : op == WS_AGG_MIN ? (is_int ? WS_DK_MIN_I : WS_DK_MIN_D)
                            : (is_int ? WS_DK_MAX_I : WS_DK_MAX_D);
}

// Dense spec decision for integer keys with known ranges (engine_select.cpp:1169-1229).
bool plan_dense(const AggPlan& plan, const std::vector<std::pair<int64_t, int64_t>>& ranges, int64_t rows_in, DenseSpec& ds) {
  std::memset(&ds, 0, sizeof ds);
  int64_t G = 1;
  for (int k = 0; k < plan.nkeys; ++k) {
    const uint64_t span = (uint64_t)(ranges[k].second - ranges[k].first);
    if (span >= (1ull << 26)) return false;
    const int64_t range = (int64_t)span + 1;
    if (G > (1ll << 26) / range) return false;
    ds.kmin[k] = ranges[k].first;
    ds.range[k] = range;
    ds.stride[k] = G;
    G *= range;
  }
  int off = 0;
  bool int_only = true;
  for (int a = 0; a < plan.naggs; ++a) {
    ds.kind[a] = (uint8_t)dense_kind(plan.ops[a], plan.agg_int[a] != 0);
    ds.slot_off[a] = off;
    const int64_t bytes = G * (ds.kind[a] == WS_DK_COUNT ? 4 : 8);
    off += (int)std::min<int64_t>((bytes + 7) / 8 * 8, 1 << 30);
    int_only &= ds.kind[a] == WS_DK_COUNT || ds.kind[a] == WS_DK_SUM_I || ds.kind[a] == WS_DK_MIN_I || ds.kind[a] == WS_DK_MAX_I;
  }
  if (G * plan.naggs <= 88) {
    ds.priv = 1;
    ds.G = G;
  } else if (off <= 48 * 1024 && G < 64 && !int_only) {
    return false;  // hash aggregation
  } else if (off <= 48 * 1024) {
    ds.copies = (int)std::min<int64_t>(8, (48 * 1024) / off);
    ds.copy_bytes = off;
    ds.G = G;
  } else if (G <= std::max<int64_t>(4 * rows_in, 1 << 16) && G * (8 + 8 * plan.nkeys + 8 * plan.naggs) <= (int64_t)512 << 20) {
    ds.copies = 0;
    ds.G = G;
  } else {
    return false;
  }
  return true;
}

int64_t pow2_at_least(int64_t v, int64_t lo) {
  int64_t c = lo;
  while (c < v) c <<= 1;
  return c;
}

// ------------------------------------------------------------------ C1: filter + dense GROUP BY

// SELECT k, COUNT(*), SUM(v), MIN(v), MAX(v), AVG(p) FROM t WHERE k < 500 AND v > 1000 GROUP BY k
// Table columns: 0 k (INT), 1 v (BIGINT), 2 p (DOUBLE). Result rows: k, count, sum, min, max as
// int64 and avg as double.
int64_t query_c1(const DevTable& t, int64_t* out_i, double* out_d, int64_t cap_rows, bool dense_on) {
  Comp c;
  const int sk = c.slot(WS_T_I32), sv = c.slot(WS_T_I64), sp = c.slot(WS_T_F64);
  const std::vector<int> slot_col = {0, 1, 2};
  AggPlan plan;
  std::memset(&plan, 0, sizeof plan);
  plan.has_where = 1;
  plan.where = c.prog({c.col(sk), c.ilit(500), Comp::op(WS_OP_LT), c.col(sv), c.ilit(1000), Comp::op(WS_OP_GT),
                       Comp::op(WS_OP_AND)});
  plan.nkeys = 1;
  plan.keys[0] = c.prog({c.col(sk)});
  plan.key_int[0] = (uint8_t)progIsInt(c.env, plan.keys[0]);
  const uint8_t ops[] = {WS_AGG_COUNT, WS_AGG_SUM, WS_AGG_SUM, WS_AGG_MIN, WS_AGG_MAX, WS_AGG_SUM};
  plan.naggs = 6;
  for (int a = 0; a < 6; ++a) plan.ops[a] = ops[a];
  for (int a = 1; a <= 4; ++a) plan.aggs[a] = c.prog({c.col(sv)});
  plan.aggs[5] = c.prog({c.col(sp)});
  plan.agg_part[1] = 1, plan.agg_part[2] = 2;
  plan.agg_int[0] = 1;
  for (int a = 1; a < 6; ++a) {
    plan.agg_int[a] = (uint8_t)progIsInt(c.env, plan.aggs[a]);
    plan.agg_bits[a] = (uint8_t)(plan.agg_int[a] && a <= 4);  // BIGINT column: exact int64 bits
  }
  DenseSpec ds;
  const bool dense = dense_on && plan_dense(plan, {{t.cols[0].mn, t.cols[0].mx}}, t.rows, ds);
  plan.dense_priv = dense ? (ds.priv ? 1 : 2) : 0;
  const double est = std::min<double>((double)(t.cols[0].mx - t.cols[0].mn + 1), (double)std::max<int64_t>(t.rows, 1));
  int64_t cap = 1024;
  const int64_t max_cap = (1ll << 30) / (8 + 8 * plan.nkeys + 8 * plan.naggs);
  while (cap < est * 2 && cap * 2 <= max_cap) cap <<= 1;
  const JitProgram* jit = jitGet(c.env, c.ncols, plan, dense ? JF_DENSE : JF_GROUPED, "aggregate");
  std::vector<double> keys, accs;
  int64_t g = 0;
  for (;;) {
    GpuAggregator agg(plan, cap, dense ? &ds : nullptr);
    ExprEnv e = c.env;
    for (size_t s = 0; s < t.seg_rows.size(); ++s) {
      bind_segment(e, c, t, slot_col, s);
      agg.addBatch(e, c.ncols, t.seg_rows[s], nullptr);
    }
    agg.run(c.env, jit);
    if (agg.overflowed()) {
      cap = pow2_at_least(std::max<int64_t>(cap * 4, agg.groupCount() * 4), 1);
      continue;
    }
    g = agg.fetch(keys, accs);
    break;
  }
  int64_t n = std::min(g, cap_rows);
  for (int64_t i = 0; i < n; ++i) {
    const double* a = &accs[i * 6];
    int64_t kbits, lo, hi, mn, mx;
    std::memcpy(&kbits, &keys[i], 8);
    std::memcpy(&lo, &a[1], 8);
    std::memcpy(&hi, &a[2], 8);
    std::memcpy(&mn, &a[3], 8);
    std::memcpy(&mx, &a[4], 8);
    const __int128 sum = (__int128)hi * ((__int128)1 << 32) + lo;  // engine_core.h:1972-1983
    out_i[i * 5 + 0] = kbits;
    out_i[i * 5 + 1] = (int64_t)a[0];
    out_i[i * 5 + 2] = (int64_t)sum;
    out_i[i * 5 + 3] = mn;
    out_i[i * 5 + 4] = mx;
    out_d[i] = a[5] / a[0];
  }
  return g;
}

// ------------------------------------------------------------------ learned sizing hook

// The sizing model the drivers consult (null: the engine's range rule alone). Every aggregation
// the drivers run carries its features, so its exact group count is logged for training.
std::atomic<ws_sizer*> g_sizer{nullptr};

// Fraction of an integer column's range below `lt` (uniform within min / max).
double frac_below(const DevColumn& c, int64_t lt) {
  if (c.mn > c.mx) return 1.0;
  return std::min(1.0, std::max(0.0, (double)(lt - c.mn) / (double)(c.mx - c.mn + 1)));
}
// Features of a GROUP BY over columns `keys` of t (or over `rows` input rows); expected groups
// from the model when one is set (0 otherwise).
int64_t size_groupby(const std::vector<const DevColumn*>& keys, int64_t rows, double sel, float* f) {
  std::vector<ws_col_stats> st;
  for (const DevColumn* k : keys) st.push_back(k->stats);
  wsd::ck(ws_sizing_features(st.data(), (int)st.size(), nullptr, rows, sel, f));
  ws_sizer* m = g_sizer.load();
  int64_t g = 0;
  if (m) wsd::ck(ws_sizer_groups(m, f, &g));
  return g;
}

// ------------------------------------------------------------------ autotuner

// Physical-plan knobs (ws_tune_set) chosen per operator shape by measuring the shape's own
// workload: coordinate descent from the current values, each knob in turn trying each candidate
// with the others fixed, a value kept when the median of `reps` timed runs beats the best by more
// than 2 %; at most two sweeps. The choice is recorded under the shape's name and applied before
// every later run of that shape (results never depend on the knobs).
struct KnobSpace {
  const char* name;
  std::vector<int64_t> values;
};
struct TuneResult {
  std::vector<std::pair<std::string, int64_t>> chosen;
  double default_ms = 0, best_ms = 0;
  int trials = 0;
};
std::mutex g_tuned_mu;
std::map<std::string, std::vector<std::pair<std::string, int64_t>>> g_tuned;

void apply_tuned(const std::string& shape) {
  std::vector<std::pair<std::string, int64_t>> kv;
  {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    auto it = g_tuned.find(shape);
    if (it == g_tuned.end()) return;
    kv = it->second;
  }
  for (auto& e : kv) wsd::ck(ws_tune_set(e.first.c_str(), e.second));
}

TuneResult autotune(const std::string& shape, const std::vector<KnobSpace>& space, int reps, const std::function<void()>& work) {
  auto timed = [&] {
    std::vector<double> t;
    for (int r = 0; r < std::max(reps, 1); ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      work();
      t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
  };
  TuneResult res;
  std::vector<int64_t> cur(space.size());
  for (size_t i = 0; i < space.size(); ++i) wsd::ck(ws_tune_get(space[i].name, &cur[i]));
  work();  // compiles and warms what the shape needs
  double best = timed();
  res.default_ms = best;
  res.trials = 1;
  for (int sweep = 0; sweep < 2; ++sweep) {
    bool changed = false;
    for (size_t i = 0; i < space.size(); ++i)
      for (int64_t v : space[i].values) {
        if (v == cur[i]) continue;
        wsd::ck(ws_tune_set(space[i].name, v));
        const double t = timed();
        ++res.trials;
        if (t < best * 0.98) {
          best = t;
          cur[i] = v;
          changed = true;
        } else {
          wsd::ck(ws_tune_set(space[i].name, cur[i]));
        }
      }
    if (!changed) break;
  }
  for (size_t i = 0; i < space.size(); ++i) res.chosen.push_back({space[i].name, cur[i]});
  res.best_ms = best;
  std::lock_guard<std::mutex> g(g_tuned_mu);
  g_tuned[shape] = res.chosen;
  return res;
}
const char* const kC4Shape = "dense GROUP BY batch (1 key, 5 slots)";
const char* const kC2Shape = "partitioned GROUP BY (web query, both phases)";

// ------------------------------------------------------------------ C2: COUNT DISTINCT + top 10

// SELECT user_id, COUNT(*) c, SUM(dur), COUNT(DISTINCT url) FROM h WHERE event_date < 19723+300 AND
// region < 224 GROUP BY user_id ORDER BY c DESC LIMIT 10.
// Table columns: 0 user_id (BIGINT), 1 region (INT, may be encoded), 2 event_date (INT), 3 url
// (INT codes), 4 dur (INT). Result rows: user_id, count, sum(dur), count distinct url.
struct C2Stats {
  int64_t pairs = 0, groups = 0, cap1 = 0, retries = 0;
};
int64_t query_c2(const DevTable& t, int64_t* out, int64_t limit, C2Stats* st, int64_t date_lt = 19723 + 300,
                 int64_t region_lt = 224) {
  apply_tuned(kC2Shape);
  Comp c;
  // slots in the engine's compile order: WHERE, keys, aggregates, then the DISTINCT argument
  const int s_date = c.slot(WS_T_I32), s_region = c.slot(WS_T_I32, t.cols[1].segs[0].enc);
  const int s_user = c.slot(WS_T_I64), s_dur = c.slot(WS_T_I32), s_url = c.slot(WS_T_I32);
  const std::vector<int> slot_col = {2, 1, 0, 4, 3};
  AggPlan plan;
  std::memset(&plan, 0, sizeof plan);
  plan.has_where = 1;
  plan.where = c.prog({c.col(s_date), c.ilit(date_lt), Comp::op(WS_OP_LT), c.col(s_region), c.ilit(region_lt),
                       Comp::op(WS_OP_LT), Comp::op(WS_OP_AND)});
  plan.nkeys = 1;
  plan.keys[0] = c.prog({c.col(s_user)});
  plan.key_int[0] = (uint8_t)progIsInt(c.env, plan.keys[0]);
  plan.naggs = 2;
  plan.ops[0] = WS_AGG_COUNT;
  plan.ops[1] = WS_AGG_SUM;
  plan.aggs[1] = c.prog({c.col(s_dur)});
  plan.agg_int[0] = 1;
  plan.agg_int[1] = (uint8_t)progIsInt(c.env, plan.aggs[1]);  // INT column: no split, no bits
  const Prog dist = c.prog({c.col(s_url)});
  const bool dist_int = progIsInt(c.env, dist);
  // phase 1: GROUP BY (user_id, url)
  AggPlan plan1 = plan;
  plan1.nkeys = 2;
  plan1.keys[1] = dist;
  plan1.key_int[1] = (uint8_t)dist_int;
  plan1.dense_priv = 0;
  const JitProgram* jit1 = jitGet(c.env, c.ncols, plan1, JF_GROUPED, "aggregate (distinct pairs)");
  int64_t cap1 = pow2_at_least(std::min<int64_t>(std::max<int64_t>(t.rows, 1), 1 << 22), 1024);
  float f1[WS_SIZER_INPUTS];
  const double sel = frac_below(t.cols[2], date_lt) * frac_below(t.cols[1], region_lt);
  const int64_t pred1 = size_groupby({&t.cols[0], &t.cols[3]}, t.rows, sel, f1);
  if (pred1 > 0) cap1 = pow2_at_least(pred1 * 5 / 2, 1024);  // 2x headroom + 25 % for the error
  DevVec<double> dk, da;
  int64_t g1 = 0;
  for (;;) {
    GpuAggregator a1(plan1, cap1, nullptr);
    ws_agg_set_features(a1.handle(), f1);
    if (pred1 > 0) wsd::ck(ws_agg_set_strategy(a1.handle(), WS_STRAT_AUTO, pred1 + pred1 / 4));
    ExprEnv e = c.env;
    for (size_t s = 0; s < t.seg_rows.size(); ++s) {
      bind_segment(e, c, t, slot_col, s);
      a1.addBatch(e, c.ncols, t.seg_rows[s], nullptr);
    }
    a1.run(c.env, jit1);
    if (a1.overflowed()) {
      cap1 = pow2_at_least(std::max<int64_t>(cap1 * 4, a1.groupCount() * 4), 1);
      if (st) ++st->retries;
      continue;
    }
    g1 = a1.compact(dk, da);
    break;
  }
  // phase 2: GROUP BY user_id over the distinct pairs (interpreted, as the engine runs it)
  const int nk = plan.nkeys, na = plan1.naggs;
  ExprEnv env2;
  std::memset(&env2, 0, sizeof env2);
  int nc = 0;
  for (int k = 0; k <= nk; ++k) {
    const bool in = k < nk ? plan.key_int[k] != 0 : dist_int;
    env2.cols[k].p = dk.p + (int64_t)k * g1;
    env2.cols[k].t = (uint8_t)(in ? WS_T_I64 : WS_T_F64);
    env2.code[nc++] = ws_instr{WS_OP_COL, (uint8_t)in, (uint16_t)k};
  }
  AggPlan plan2;
  std::memset(&plan2, 0, sizeof plan2);
  plan2.nkeys = nk;
  for (int k = 0; k < nk; ++k) {
    plan2.keys[k] = Prog{(int16_t)k, 1};
    plan2.key_int[k] = plan.key_int[k];
  }
  for (int a = 0; a < na; ++a) {
    env2.cols[nk + 1 + a].p = da.p + (int64_t)a * g1;
    env2.cols[nk + 1 + a].t = (uint8_t)(plan1.agg_bits[a] ? WS_T_I64 : WS_T_F64);
    plan2.aggs[a] = Prog{(int16_t)nc, 1};
    env2.code[nc++] = ws_instr{WS_OP_COL, 0, (uint16_t)(nk + 1 + a)};
    plan2.ops[a] = plan1.ops[a] == WS_AGG_COUNT ? WS_AGG_SUM : plan1.ops[a];
    plan2.agg_int[a] = plan2.agg_bits[a] = plan1.agg_bits[a];
  }
  if (dist_int) {
    plan2.aggs[na] = Prog{(int16_t)nc, 1};
    env2.code[nc++] = ws_instr{WS_OP_CONST, 0, 0};
    env2.consts[0] = 1.0;
  } else {
    plan2.aggs[na] = Prog{(int16_t)nc, 3};
    env2.code[nc++] = ws_instr{WS_OP_COL, 0, (uint16_t)nk};
    env2.code[nc++] = ws_instr{WS_OP_ISNULL, 0, 0};
    env2.code[nc++] = ws_instr{WS_OP_NOT, 0, 0};
  }
  plan2.ops[na] = WS_AGG_SUM;
  plan2.naggs = na + 1;
  // The engine interprets this phase; a compiled plan gives the same groups (the plan cache keeps
  // it across queries: the environment's column pointers are bound per run, not compiled in).
  const JitProgram* jit2 = jitGet(env2, nk + 1 + na, plan2, JF_GROUPED, "aggregate (distinct users)");
  float f2[WS_SIZER_INPUTS];
  const int64_t pred2 = size_groupby({&t.cols[0]}, g1, 1.0, f2);
  int64_t cap2 = pow2_at_least(2 * g1, 1024);
  if (pred2 > 0) cap2 = pow2_at_least(std::min<int64_t>(g1, pred2 + pred2 / 4) * 2, 1024);
  // (without a model cap2 >= 2 * g1 cannot overflow; a model's under-estimate retries bigger)
  std::unique_ptr<GpuAggregator> agg2;
  int64_t expect2 = pred2 > 0 ? std::min<int64_t>(g1, pred2 + pred2 / 4) : 0;
  for (;;) {
    agg2.reset(new GpuAggregator(plan2, cap2, nullptr));
    ws_agg_set_features(agg2->handle(), f2);
    if (expect2 > 0) wsd::ck(ws_agg_set_strategy(agg2->handle(), WS_STRAT_AUTO, expect2));
    agg2->addBatch(env2, nk + 1 + na, g1, nullptr);
    agg2->run(env2, jit2);
    if (!agg2->overflowed()) break;
    cap2 = pow2_at_least(std::max<int64_t>(cap2 * 4, agg2->groupCount() * 4), 1);
    expect2 = std::max<int64_t>(expect2 * 4, agg2->groupCount() * 2);
    if (st) ++st->retries;
  }
  // post-aggregation: ORDER BY c DESC LIMIT n on the GPU over the compacted groups
  DevVec<double> gk, ga;
  const int64_t g = agg2->compact(gk, ga);
  ExprEnv penv;
  std::memset(&penv, 0, sizeof penv);
  std::vector<DevVec<double>> conv;
  for (int k = 0; k < nk; ++k) {
    if (plan2.key_int[k]) {
      conv.emplace_back(std::max<int64_t>(g, 1));
      gpuBitsToDouble(gk.p + (int64_t)k * g, g, conv.back().p);
      penv.cols[k].p = conv.back().p;
    } else {
      penv.cols[k].p = gk.p + (int64_t)k * g;
    }
    penv.cols[k].t = WS_T_F64;
  }
  for (int a = 0; a < plan2.naggs; ++a) {
    penv.cols[nk + a].p = ga.p + (int64_t)a * g;
    penv.cols[nk + a].t = WS_T_F64;
  }
  penv.code[0] = ws_instr{WS_OP_COL, 0, (uint16_t)(nk + 0)};  // c = COUNT(*) = slot 0
  const Prog order{0, 1};
  const int64_t mc = std::min<int64_t>(g, limit);
  DevVec<double> okey(std::max<int64_t>(g, 1));
  gpuEvalAt(penv, order, nullptr, g, okey.p);
  DevVec<int32_t> perm = gpuTopKPerm({{okey.p, true}}, g, mc);
  const int W = nk + plan2.naggs;
  DevVec<double> tmp(std::max<int64_t>(mc * W, 1));
  for (int col = 0; col < W; ++col)
    gpuGather(col < nk ? gk.p + (int64_t)col * g : ga.p + (int64_t)(col - nk) * g, PType::F64, perm.p, mc,
              tmp.p + (int64_t)col * mc);
  const std::vector<double> all = tmp.download(mc * W);
  for (int64_t i = 0; i < mc; ++i) {
    int64_t user;
    std::memcpy(&user, &all[i], 8);
    out[i * 4 + 0] = user;
    out[i * 4 + 1] = (int64_t)all[(size_t)(1 * mc + i)];
    out[i * 4 + 2] = std::llround(all[(size_t)(2 * mc + i)]);
    out[i * 4 + 3] = (int64_t)all[(size_t)(3 * mc + i)];
  }
  if (st) st->pairs = g1, st->groups = g, st->cap1 = cap1;
  return mc;
}

// ------------------------------------------------------------------ C3: star join

// A materialized relation: device columns of the selected rows.
struct RelCol {
  DevVec<uint8_t> data;
  int t;
};

// materialize(table, WHERE) of the needed columns (engine_core.h:4345-4487): one fused filter
// launch over all segments, then one gather per needed column.
int64_t materialize(const DevTable& t, Comp& c, bool has_where, Prog where, const std::vector<int>& slot_col,
                    const std::vector<int>& need, std::vector<RelCol>& out, const char* kind = "filter") {
  const JitProgram* jit = nullptr;
  if (has_where) {
    AggPlan fp;
    std::memset(&fp, 0, sizeof fp);
    fp.has_where = 1;
    fp.where = where;
    jit = jitGet(c.env, c.ncols, fp, JF_FLAGM, kind);
  }
  std::vector<int64_t> bn;
  std::vector<const uint32_t*> bdel;
  std::vector<ColView> bcols;
  ExprEnv e = c.env;
  for (size_t s = 0; s < t.seg_rows.size(); ++s) {
    bn.push_back(t.seg_rows[s]);
    bdel.push_back(nullptr);
    bind_segment(e, c, t, slot_col, s);
    for (int k = 0; k < c.ncols; ++k) bcols.push_back(e.cols[k]);
  }
  DevVec<int32_t> sel;
  std::vector<int64_t> per;
  const int64_t m = gpuSelectMulti(c.env, has_where, where, bn, bdel, bcols, c.ncols, sel, per, jit);
  std::vector<int64_t> boff(bn.size(), 0);
  for (size_t b = 1; b < bn.size(); ++b) boff[b] = boff[b - 1] + bn[b - 1];
  out.clear();
  for (int col : need) {
    RelCol rc;
    rc.t = t.cols[col].t;
    const int w = width_of(rc.t);
    rc.data = DevVec<uint8_t>(std::max<int64_t>(m, 1) * w);
    std::vector<ColView> src;
    for (size_t s = 0; s < bn.size(); ++s) src.push_back(t.view(col, s));
    const PType pt = w == 4 ? PType::I32 : PType::I64;
    if (m > 0) gpuGatherMulti(src, boff, pt, w, sel.p, m, rc.data.p);
    out.push_back(std::move(rc));
  }
  return m;
}

// SELECT f.flag, COUNT(*), SUM(f.price), SUM(f.qty) FROM fact f JOIN dim d ON f.fk = d.key
// WHERE d.attr < 50 GROUP BY f.flag.
// fact columns: 0 fk (INT), 1 qty (BIGINT), 2 price (DECIMAL(15,2) as int64), 3 flag (INT);
// dim columns: 0 key (INT), 1 attr (INT). Result rows: flag, count, raw price sum (scaled by
// 100), qty sum.
struct C3Stats {
  int64_t dim_rows = 0, fact_rows = 0, pairs = 0, predicted_pairs = 0;
};
// The join output model the drivers consult (null: the reference's count, scan, read-back sizing).
std::atomic<ws_sizer*> g_join_sizer{nullptr};

// Set to false to keep the hash join when the build keys are distinct (tests of both plans, and
// the learned join sizing, which sizes that join).
std::atomic<bool> g_join_rewrite{true};

// The aggregation of C3 over rows given by `bind` (columns flag, price, qty in slots 0-2), with an
// optional WHERE over slot 3 (fk - lo IN set 0): rows (flag, count, raw price sum, qty sum).
struct C3Agg {
  Comp c;
  AggPlan plan;
  int f_flag, f_price, f_qty, f_fk = -1;
  std::function<void()> after_launch;  // called once the aggregation is launched, before its results are read
  C3Agg(const uint32_t* bits, int64_t lo, int64_t nbits) {
    f_flag = c.slot(WS_T_I32);
    f_price = c.slot(WS_T_DECIMAL);
    f_qty = c.slot(WS_T_I64);
    c.env.cols[f_price].scale = 2;
    std::memset(&plan, 0, sizeof plan);
    if (bits) {
      f_fk = c.slot(WS_T_I32);
      c.env.sets[0] = bits;
      c.env.set_size[0] = (int32_t)nbits;
      plan.has_where = 1;
      plan.where = c.prog({c.col(f_fk), c.ilit(lo), Comp::op(WS_OP_SUB), Comp::I{WS_OP_INSET, 0, 0}});
    }
    plan.nkeys = 1;
    plan.keys[0] = c.prog({c.col(f_flag)});
    plan.key_int[0] = (uint8_t)progIsInt(c.env, plan.keys[0]);
    plan.naggs = 4;
    plan.ops[0] = WS_AGG_COUNT;
    plan.ops[1] = plan.ops[2] = plan.ops[3] = WS_AGG_SUM;
    plan.aggs[1] = c.prog({Comp::I{WS_OP_COLRAW, 1, (uint16_t)f_price}});
    plan.aggs[2] = plan.aggs[3] = c.prog({c.col(f_qty)});
    plan.agg_part[2] = 1, plan.agg_part[3] = 2;
    plan.agg_int[0] = 1;
    for (int a = 1; a < 4; ++a) plan.agg_int[a] = (uint8_t)progIsInt(c.env, plan.aggs[a]);
    plan.agg_bits[2] = plan.agg_bits[3] = 1;  // BIGINT column SUM with GROUP BY: split halves
  }
  // Runs over batches added by `add` (called with the aggregator); returns the group count.
  template <class AddBatches>
  int64_t run(const DevColumn& flag_col, int64_t rows, const char* kind, AddBatches add, int64_t* out) {
    DenseSpec ds;
    const bool dense = plan_dense(plan, {{flag_col.mn, flag_col.mx}}, rows, ds);
    plan.dense_priv = dense ? (ds.priv ? 1 : 2) : 0;
    const JitProgram* jit = jitGet(c.env, c.ncols, plan, dense ? JF_DENSE : JF_GROUPED, kind);
    int64_t cap = pow2_at_least(2 * 25, 1024);
    std::vector<double> keys, accs;
    int64_t g = 0;
    for (;;) {
      GpuAggregator agg(plan, cap, dense ? &ds : nullptr);
      add(agg);
      agg.run(c.env, jit);
      if (after_launch) after_launch();
      if (agg.overflowed()) {
        cap = pow2_at_least(std::max<int64_t>(cap * 4, agg.groupCount() * 4), 1);
        continue;
      }
      g = agg.fetch(keys, accs);
      break;
    }
    for (int64_t i = 0; i < g; ++i) {
      int64_t kb, lo2, hi2;
      std::memcpy(&kb, &keys[i], 8);
      std::memcpy(&lo2, &accs[i * 4 + 2], 8);
      std::memcpy(&hi2, &accs[i * 4 + 3], 8);
      out[i * 4 + 0] = kb;
      out[i * 4 + 1] = (int64_t)accs[i * 4 + 0];
      out[i * 4 + 2] = (int64_t)accs[i * 4 + 1];  // exact: per-group raw sums < 2^53
      out[i * 4 + 3] = (int64_t)((__int128)hi2 * ((__int128)1 << 32) + lo2);
    }
    return g;
  }
};

int64_t query_c3(const DevTable& fact, const DevTable& dim, int64_t* out, C3Stats* st, int64_t attr_lt = 50) {
  const int64_t lo = dim.cols[0].mn, nbits = dim.cols[0].mx - dim.cols[0].mn + 1;
  // A key column on the dimension side: each fact row that passes the bitmap of the filtered keys
  // joins exactly one dimension row, and no dimension column is read after the join, so the join
  // is the bitmap filter itself. The bitmap is set in one scan of the dimension (no materialized
  // build side) and the fact table aggregated in one pass with the bitmap test as its WHERE (the
  // set staged in LDS): no selection vector, gathers or join, and no read-back before the results.
  if (g_join_rewrite.load() && dim.cols[0].unique && dim.rows > 0 && nbits <= ((int64_t)1 << 30)) {
    Comp cd;
    const int d_key = cd.slot(WS_T_I32), d_attr = cd.slot(WS_T_I32);
    const Prog dwhere = cd.prog({cd.col(d_attr), cd.ilit(attr_lt), Comp::op(WS_OP_LT)});
    const Prog dkey = cd.prog({cd.col(d_key)});
    std::vector<int64_t> bn;
    std::vector<const uint32_t*> bdel;
    std::vector<ColView> bcols;
    ExprEnv e = cd.env;
    for (size_t sg = 0; sg < dim.seg_rows.size(); ++sg) {
      bn.push_back(dim.seg_rows[sg]);
      bdel.push_back(nullptr);
      bind_segment(e, cd, dim, {0, 1}, sg);
      for (int k = 0; k < cd.ncols; ++k) bcols.push_back(e.cols[k]);
    }
    AggPlan kp;  // the selection plan of the filter, with the key as keys[0] (compiled key bitmap kernel)
    std::memset(&kp, 0, sizeof kp);
    kp.has_where = 1;
    kp.where = dwhere;
    kp.nkeys = 1;
    kp.keys[0] = dkey;
    kp.key_int[0] = 1;
    const JitProgram* kjit = jitGet(cd.env, cd.ncols, kp, JF_FLAGM, "key bitmap (dimension filter)");
    DevVec<uint32_t> bits((nbits + 31) / 32);
    DevVec<uint64_t> nsel(1);
    wsd::ck(ws_key_bitmap_where(wsd::env(cd.env), 1, wsd::prog(dwhere), wsd::prog(dkey), (int)bn.size(), bn.data(), bdel.data(),
                                reinterpret_cast<const ws_colview*>(bcols.data()), cd.ncols, lo, nbits, bits.p, nsel.p,
                                wsd::handle(kjit)));
    // the count comes back with the aggregation's read-back: its copy is queued after the
    // aggregation's launches, off the path to the fact scan
    static thread_local uint64_t* t_nsel = nullptr;
    if (!t_nsel) wsd::ck(ws_pinned_alloc(sizeof(uint64_t), (void**)&t_nsel));
    C3Agg A(bits.p, lo, nbits);
    A.after_launch = [&] { wsd::ck(ws_d2h_async(t_nsel, nsel.p, sizeof(uint64_t))); };
    const std::vector<int> slot_col = {3, 2, 1, 0};  // flag, price, qty, fk
    const int64_t g = A.run(fact.cols[3], fact.rows, "aggregate fact (join as filter)", [&](GpuAggregator& agg) {
      ExprEnv e = A.c.env;
      for (size_t sg = 0; sg < fact.seg_rows.size(); ++sg) {
        bind_segment(e, A.c, fact, slot_col, sg);
        agg.addBatch(e, A.c.ncols, fact.seg_rows[sg], nullptr);
      }
    }, out);
    if (st) {
      int64_t pairs = 0;
      for (int64_t i = 0; i < g; ++i) pairs += out[i * 4 + 1];
      st->dim_rows = (int64_t)*t_nsel;
      st->fact_rows = st->pairs = pairs;
      st->predicted_pairs = 0;
    }
    return g;
  }
  // dim: materialize(dim, [attr < attr_lt], {key})
  Comp cd;
  const int d_attr = cd.slot(WS_T_I32);
  const Prog dwhere = cd.prog({cd.col(d_attr), cd.ilit(attr_lt), Comp::op(WS_OP_LT)});
  std::vector<RelCol> drel;
  const int64_t nd = materialize(dim, cd, true, dwhere, {1}, {0}, drel, "filter (dimension)");
  // join filter: the dim keys were reduced below half their table, fact is large: exact bitmap
  // of the keys over the dim key range (engine_select.cpp:1015-1041)
  DevVec<uint32_t> bits((nbits + 31) / 32);
  ColView kv{};
  kv.p = drel[0].data.p;
  kv.t = WS_T_I32;
  gpuKeyBitmap(kv, nd, lo, nbits, bits.p);
  // Distinct filtered keys on a dimension without a key column (as many set bits as filtered
  // rows): the join is the bitmap filter, as above, over the materialized build side.
  if (g_join_rewrite.load() && nd > 0 && gpuBitmapCount(bits.p, (int64_t)bits.n) == nd) {
    C3Agg A(bits.p, lo, nbits);
    const std::vector<int> slot_col = {3, 2, 1, 0};  // flag, price, qty, fk
    const int64_t g = A.run(fact.cols[3], fact.rows, "aggregate fact (join as filter)", [&](GpuAggregator& agg) {
      ExprEnv e = A.c.env;
      for (size_t sg = 0; sg < fact.seg_rows.size(); ++sg) {
        bind_segment(e, A.c, fact, slot_col, sg);
        agg.addBatch(e, A.c.ncols, fact.seg_rows[sg], nullptr);
      }
    }, out);
    int64_t pairs = 0;
    for (int64_t i = 0; i < g; ++i) pairs += out[i * 4 + 1];
    if (st) st->dim_rows = nd, st->fact_rows = pairs, st->pairs = pairs, st->predicted_pairs = 0;
    return g;
  }
  // fact: materialize(fact, [INSET(fk - lo, bitmap)], {fk, price, qty, flag})
  Comp cf;
  const int f_fk = cf.slot(WS_T_I32);
  cf.env.sets[0] = bits.p;
  cf.env.set_size[0] = (int32_t)nbits;
  const Prog fwhere = cf.prog({cf.col(f_fk), cf.ilit(lo), Comp::op(WS_OP_SUB), Comp::I{WS_OP_INSET, 0, 0}});
  std::vector<RelCol> frel;
  const int64_t nf = materialize(fact, cf, true, fwhere, {0}, {0, 2, 1, 3}, frel, "filter (fact by join keys)");
  // hash join: build on the smaller side (dim), probe the filtered fact rows
  DevVec<int32_t> ib, ip;
  ColView bk{}, pk{};
  bk.p = drel[0].data.p, bk.t = WS_T_I32;
  pk.p = frel[0].data.p, pk.t = WS_T_I32;
  // output size: predicted by the join model when one is installed (pair buffers sized ahead, no
  // read-back between count and write), else counted first; the exact size is logged either way
  float jf[WS_SIZER_INPUTS];
  wsd::ck(ws_join_features(&dim.cols[0].stats, nd, dim.rows, &fact.cols[0].stats, nf, jf));
  int64_t pred = 0;
  if (ws_sizer* model = g_join_sizer.load()) wsd::ck(ws_sizer_groups(model, jf, &pred));
  int64_t m = 0;
  if (pred > 0) {
    int32_t *b = nullptr, *p = nullptr;
    wsd::ck(ws_hash_join_sized(wsd::view(bk), nd, wsd::view(pk), nf, pred + pred / 8 + 1024, &b, &p, &m));
    ib = DevVec<int32_t>::adopt(b, b ? m : 0);
    ip = DevVec<int32_t>::adopt(p, p ? m : 0);
  } else {
    m = gpuHashJoin(bk, nd, pk, nf, ib, ip);
  }
  ws_join_log_add(jf, m);
  // joined relation: gather every needed column of both sides by pair index
  std::vector<DevVec<uint8_t>> jcols;
  auto gather_side = [&](const RelCol& rc, const int32_t* idx) {
    const int w = width_of(rc.t);
    jcols.emplace_back(std::max<int64_t>(m, 1) * w);
    if (m) gpuGather(rc.data.p, w == 4 ? PType::I32 : PType::I64, idx, m, jcols.back().p);
  };
  gather_side(drel[0], ib.p);
  for (auto& rc : frel) gather_side(rc, ip.p);
  // aggregate the joined relation: keys flag; slots COUNT(*), SUM(#RAW price), SUM(qty) lo / hi
  C3Agg A(nullptr, 0, 0);
  const int64_t g = A.run(fact.cols[3], m, "aggregate join result", [&](GpuAggregator& agg) {
    ExprEnv e = A.c.env;
    e.cols[A.f_flag].p = jcols[4].p;   // flag
    e.cols[A.f_price].p = jcols[2].p;  // price
    e.cols[A.f_qty].p = jcols[3].p;    // qty
    agg.addBatch(e, A.c.ncols, m, nullptr);
  }, out);
  if (st) st->dim_rows = nd, st->fact_rows = nf, st->pairs = m, st->predicted_pairs = pred;
  return g;
}

thread_local std::string t_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const SqlError& e) {
    t_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    t_err = e.what();
    return 1105;
  }
}

// ------------------------------------------------------------------ C4: many small aggregations

// SELECT key, COUNT(*), SUM(val), MIN(val), MAX(val) FROM seg_q WHERE val > theta_q GROUP BY key
// over its own table per query (columns 0 key INT, 1 val BIGINT). Every query has the same plan
// shape (theta is a constant of the environment), so one compiled plan serves all of them.
struct C4Query {
  const DevTable* t = nullptr;
  int64_t theta = 0;
};
struct C4Plan {
  Comp c;
  AggPlan plan;
  DenseSpec ds;
  bool dense = false;
  int64_t cap = 1024;
};
const std::vector<int> kC4Slots = {0, 1};

// batched: the plan of the batched forms, which keep every query dense (a dense table for a query
// with many possible keys and few rows costs memory, not correctness: dense and hash give the same
// groups); the per-query path keeps the engine's rule.
C4Plan c4_plan(const C4Query& q, bool batched = false) {
  C4Plan p;
  const int sk = p.c.slot(WS_T_I32), sv = p.c.slot(WS_T_I64);
  std::memset(&p.plan, 0, sizeof p.plan);
  p.plan.has_where = 1;
  p.plan.where = p.c.prog({p.c.col(sv), p.c.ilit(q.theta), Comp::op(WS_OP_GT)});
  p.plan.nkeys = 1;
  p.plan.keys[0] = p.c.prog({p.c.col(sk)});
  p.plan.key_int[0] = 1;
  const uint8_t ops[] = {WS_AGG_COUNT, WS_AGG_SUM, WS_AGG_SUM, WS_AGG_MIN, WS_AGG_MAX};
  p.plan.naggs = 5;
  for (int a = 0; a < 5; ++a) p.plan.ops[a] = ops[a];
  p.plan.agg_int[0] = 1;
  for (int a = 1; a < 5; ++a) {
    p.plan.aggs[a] = p.c.prog({p.c.col(sv)});
    p.plan.agg_int[a] = p.plan.agg_bits[a] = 1;
  }
  p.plan.agg_part[1] = 1, p.plan.agg_part[2] = 2;
  const DevColumn& k = q.t->cols[0];
  p.dense = plan_dense(p.plan, {{k.mn, k.mx}}, batched ? ((int64_t)1 << 40) : q.t->rows, p.ds);
  p.plan.dense_priv = p.dense ? (p.ds.priv ? 1 : 2) : 0;
  const double est = std::min<double>((double)(k.mx - k.mn + 1), (double)std::max<int64_t>(q.t->rows, 1));
  while (p.cap < est * 2) p.cap <<= 1;
  return p;
}

// Order-independent digest of one query's result rows (key, count, sum, min, max): equal digests
// mean equal row sets.
uint64_t c4_digest(const double* keys, const double* accs, int64_t g) {
  uint64_t d = 0x12345678ull;
  for (int64_t i = 0; i < g; ++i) {
    int64_t key, lo, hi, mn, mx;
    std::memcpy(&key, &keys[i], 8);
    std::memcpy(&lo, &accs[i * 5 + 1], 8);
    std::memcpy(&hi, &accs[i * 5 + 2], 8);
    std::memcpy(&mn, &accs[i * 5 + 3], 8);
    std::memcpy(&mx, &accs[i * 5 + 4], 8);
    const int64_t cnt = (int64_t)accs[i * 5];
    const int64_t sum = hi * (int64_t)4294967296ll + lo;
    uint64_t h = 0x9e3779b97f4a7c15ull;
    for (int64_t v : {key, cnt, sum, mn, mx}) {
      h ^= (uint64_t)v;
      h *= 0xbf58476d1ce4e5b9ull;
      h ^= h >> 31;
    }
    d += h;  // sum of row hashes: independent of row order
  }
  return d + (uint64_t)g;
}

// One query alone through the plain per-query path (the engine's aggregation driver).
uint64_t c4_alone(const C4Query& q, int64_t* groups, bool digest = true) {
  C4Plan p = c4_plan(q);
  const JitProgram* jit = jitGet(p.c.env, p.c.ncols, p.plan, p.dense ? JF_DENSE : JF_GROUPED, "aggregate");
  std::vector<double> keys, accs;
  int64_t cap = p.cap, g = 0;
  for (;;) {
    GpuAggregator agg(p.plan, cap, p.dense ? &p.ds : nullptr);
    ExprEnv e = p.c.env;
    for (size_t s = 0; s < q.t->seg_rows.size(); ++s) {
      bind_segment(e, p.c, *q.t, kC4Slots, s);
      agg.addBatch(e, p.c.ncols, q.t->seg_rows[s], nullptr);
    }
    agg.run(p.c.env, jit);
    if (agg.overflowed()) {
      cap = pow2_at_least(std::max<int64_t>(cap * 4, agg.groupCount() * 4), 1);
      continue;
    }
    g = agg.fetch(keys, accs);
    break;
  }
  if (groups) *groups = g;
  return digest ? c4_digest(keys.data(), accs.data(), g) : 0;
}

// Page-locked result buffers of the batched fetch (per thread, grown as needed): the groups of a
// whole batch (millions for the 256-query workload) cross the link at its full speed.
struct PinnedResults {
  double* keys = nullptr;
  double* accs = nullptr;
  int64_t cap = 0;
  void reserve(int64_t groups, int nk, int na) {
    if (groups <= cap) return;
    release();
    int64_t c = 1 << 16;
    while (c < groups) c <<= 1;
    void* k = nullptr;
    void* a = nullptr;
    wsd::ck(ws_pinned_alloc(sizeof(double) * (size_t)(c * std::max(nk, 1)), &k));
    const int rc = ws_pinned_alloc(sizeof(double) * (size_t)(c * std::max(na, 1)), &a);
    if (rc) {
      ws_pinned_free(k);
      wsd::ck(rc);
    }
    keys = (double*)k, accs = (double*)a, cap = c;
  }
  void release() {
    ws_pinned_free(keys);
    ws_pinned_free(accs);
    keys = accs = nullptr;
    cap = 0;
  }
  ~PinnedResults() { release(); }
};
thread_local PinnedResults t_results;

// Queries of one plan shape together: every dense query (any group count: a tile whose query has
// more groups than its LDS table folds into the query's accumulators directly) in one batched
// launch and one batched fetch; a query that is not dense runs alone. digests may be null (the
// results are fetched but not digested).
void c4_batched(const std::vector<const C4Query*>& qs, std::vector<uint64_t>* digests, std::vector<int64_t>& groups) {
  const size_t n = qs.size();
  if (digests) digests->assign(n, 0);
  groups.assign(n, 0);
  if (!n) return;
  std::vector<C4Plan> plans;
  plans.reserve(n);
  std::vector<size_t> dense_idx;
  for (size_t i = 0; i < n; ++i) {
    plans.push_back(c4_plan(*qs[i], true));
    if (plans.back().dense) {
      dense_idx.push_back(i);
    } else {
      const uint64_t d = c4_alone(*qs[i], &groups[i]);
      if (digests) (*digests)[i] = d;
    }
  }
  if (dense_idx.empty()) return;
  apply_tuned(kC4Shape);
  const C4Plan& p0 = plans[dense_idx[0]];
  const JitProgram* jit = jitGet(p0.c.env, p0.c.ncols, p0.plan, JF_DENSE, "aggregate (batched)");
  std::vector<std::unique_ptr<GpuAggregator>> aggs;
  std::vector<ws_agg*> handles;
  std::vector<const ws_expr_env*> envs;
  int64_t cap = 0;
  for (size_t i : dense_idx) {
    const C4Plan& p = plans[i];
    aggs.emplace_back(new GpuAggregator(p.plan, p.cap, &p.ds));
    ExprEnv e = p.c.env;
    for (size_t s = 0; s < qs[i]->t->seg_rows.size(); ++s) {
      bind_segment(e, p.c, *qs[i]->t, kC4Slots, s);
      aggs.back()->addBatch(e, p.c.ncols, qs[i]->t->seg_rows[s], nullptr);
    }
    handles.push_back(aggs.back()->handle());
    envs.push_back(reinterpret_cast<const ws_expr_env*>(&p.c.env));
    cap += p.ds.G;
  }
  const int m = (int)handles.size();
  wsd::ck(ws_agg_run_batch(m, handles.data(), envs.data(), jit ? jit->handle : nullptr));
  std::vector<int64_t> counts((size_t)m);
  int64_t total = 0;
  t_results.reserve(std::max<int64_t>(cap, 1), 1, 5);
  wsd::ck(ws_agg_fetch_many(m, handles.data(), counts.data(), &total, t_results.cap, t_results.keys, t_results.accs));
  int64_t off = 0;
  for (int j = 0; j < m; ++j) {
    const size_t i = dense_idx[(size_t)j];
    const int64_t g = counts[(size_t)j];
    if (digests) (*digests)[i] = c4_digest(t_results.keys + off, t_results.accs + off * 5, g);
    groups[i] = g;
    off += g;
  }
}

// A C4 set prepared once (plans, aggregators, the batch's launches recorded in a graph) and replayed:
// the repeated-dashboard form of the batched workload.
struct C4Prepared {
  std::vector<C4Plan> plans;
  std::vector<std::unique_ptr<GpuAggregator>> aggs;
  ws_agg_batch_plan* plan = nullptr;
  int64_t cap = 0;
  ~C4Prepared() { ws_agg_batch_release(plan); }
};
C4Prepared* c4_prepare(const std::vector<const C4Query*>& qs) {
  auto P = std::make_unique<C4Prepared>();
  apply_tuned(kC4Shape);
  for (const C4Query* q : qs) {
    P->plans.push_back(c4_plan(*q, true));
    if (!P->plans.back().dense) throw SqlError("prepared batch: every query must be dense", 1105);
  }
  const C4Plan& p0 = P->plans[0];
  const JitProgram* jit = jitGet(p0.c.env, p0.c.ncols, p0.plan, JF_DENSE, "aggregate (batched)");
  std::vector<ws_agg*> handles;
  std::vector<const ws_expr_env*> envs;
  for (size_t i = 0; i < qs.size(); ++i) {
    const C4Plan& p = P->plans[i];
    P->aggs.emplace_back(new GpuAggregator(p.plan, p.cap, &p.ds));
    ExprEnv e = p.c.env;
    for (size_t s = 0; s < qs[i]->t->seg_rows.size(); ++s) {
      bind_segment(e, p.c, *qs[i]->t, kC4Slots, s);
      P->aggs.back()->addBatch(e, p.c.ncols, qs[i]->t->seg_rows[s], nullptr);
    }
    handles.push_back(P->aggs.back()->handle());
    envs.push_back(reinterpret_cast<const ws_expr_env*>(&p.c.env));
    P->cap += p.ds.G;
  }
  wsd::ck(ws_agg_batch_prepare((int)handles.size(), handles.data(), envs.data(), jit ? jit->handle : nullptr, &P->plan));
  return P.release();
}
void c4_replay(C4Prepared& P, std::vector<uint64_t>* digests, std::vector<int64_t>& groups) {
  const size_t n = P.plans.size();
  std::vector<int64_t> counts(n);
  int64_t total = 0;
  t_results.reserve(std::max<int64_t>(P.cap, 1), 1, 5);
  wsd::ck(ws_agg_batch_replay(P.plan, counts.data(), &total, t_results.cap, t_results.keys, t_results.accs));
  if (digests) digests->assign(n, 0);
  groups.assign(n, 0);
  int64_t off = 0;
  for (size_t i = 0; i < n; ++i) {
    if (digests) (*digests)[i] = c4_digest(t_results.keys + off, t_results.accs + off * 5, counts[i]);
    groups[i] = counts[i];
    off += counts[i];
  }
}

// Steps of a prepared set as a pipeline over D prepared copies of it (their own aggregators,
// graphs and streams): step s launches copy s % D, then finishes step s - (D - 1), whose groups are
// copied to the host while the D - 1 later steps' kernels run (D = 3 keeps two batches' kernels
// queued during each copy). digests / groups: of the last step.
void c4_pipeline(const std::vector<C4Prepared*>& P, int steps, std::vector<uint64_t>* digests, std::vector<int64_t>& groups) {
  const int D = (int)P.size();
  if (D < 1) throw SqlError("pipelined batch: no prepared copies", 1105);
  const size_t n = P[0]->plans.size();
  int64_t cap = 1;
  for (const C4Prepared* p : P) {
    if (p->plans.size() != n) throw SqlError("pipelined batch: the prepared copies differ in size", 1105);
    cap = std::max(cap, p->cap);
  }
  std::vector<int64_t> counts(n, 0);
  t_results.reserve(cap, 1, 5);
  std::vector<bool> out((size_t)D, false);  // launched, not yet finished
  auto finish = [&](int k) {
    int64_t total = 0;
    out[(size_t)k] = false;
    wsd::ck(ws_agg_batch_finish(P[(size_t)k]->plan, counts.data(), &total, t_results.cap, t_results.keys, t_results.accs));
  };
  try {
    for (int s = 0; s < steps; ++s) {
      wsd::ck(ws_agg_batch_launch(P[(size_t)(s % D)]->plan));
      out[(size_t)(s % D)] = true;
      if (s >= D - 1) finish((s - (D - 1)) % D);
    }
    for (int s = std::max(0, steps - (D - 1)); s < steps; ++s) finish(s % D);
  } catch (...) {
    for (int k = 0; k < D; ++k)
      if (out[(size_t)k]) {
        int64_t total = 0;
        (void)ws_agg_batch_finish(P[(size_t)k]->plan, counts.data(), &total, t_results.cap, t_results.keys, t_results.accs);
      }
    throw;
  }
  if (digests) digests->assign(n, 0);
  groups.assign(n, 0);
  int64_t off = 0;
  for (size_t i = 0; i < n && steps > 0; ++i) {
    if (digests) (*digests)[i] = c4_digest(t_results.keys + off, t_results.accs + off * 5, counts[i]);
    groups[i] = counts[i];
    off += counts[i];
  }
}

// steps batches of the queries, each built from scratch (plans, aggregators, descriptor blob and
// graph: c4_prepare) and pipelined `depth` deep: batch s is launched before batch s - (depth - 1)
// is finished and released, so one batch's build, result copy and release overlap the kernels of
// the batches after it. digests / groups: of the last batch.
void c4_fresh_pipeline(const std::vector<const C4Query*>& qs, int steps, int depth, std::vector<uint64_t>* digests,
                       std::vector<int64_t>& groups) {
  if (depth < 1) throw SqlError("pipelined batch: depth below 1", 1105);
  const size_t n = qs.size();
  std::vector<int64_t> counts(n, 0);
  std::deque<std::unique_ptr<C4Prepared>> inflight;  // launched, not yet finished
  auto finish_front = [&] {
    std::unique_ptr<C4Prepared> P = std::move(inflight.front());
    inflight.pop_front();
    t_results.reserve(std::max<int64_t>(P->cap, 1), 1, 5);
    int64_t total = 0;
    wsd::ck(ws_agg_batch_finish(P->plan, counts.data(), &total, t_results.cap, t_results.keys, t_results.accs));
  };
  try {
    for (int s = 0; s < steps; ++s) {
      std::unique_ptr<C4Prepared> P(c4_prepare(qs));
      wsd::ck(ws_agg_batch_launch(P->plan));
      inflight.push_back(std::move(P));
      if ((int)inflight.size() >= depth) finish_front();
    }
    while (!inflight.empty()) finish_front();
  } catch (...) {
    while (!inflight.empty()) {  // the plans' destructors wait for their streams
      int64_t total = 0;
      (void)ws_agg_batch_finish(inflight.front()->plan, counts.data(), &total, t_results.cap, t_results.keys, t_results.accs);
      inflight.pop_front();
    }
    throw;
  }
  if (digests) digests->assign(n, 0);
  groups.assign(n, 0);
  int64_t off = 0;
  for (size_t i = 0; i < n && steps > 0; ++i) {
    if (digests) (*digests)[i] = c4_digest(t_results.keys + off, t_results.accs + off * 5, counts[i]);
    groups[i] = counts[i];
    off += counts[i];
  }
}

// ------------------------------------------------------------------ serving loop (open-loop stream)

// A request of the mixed stream: a small aggregate over one of the pool's segments, the web
// query, or the star join.
enum { kReqSmall = 0, kReqWeb = 1, kReqJoin = 2 };
struct Request {
  int type = kReqSmall;
  int64_t pick = 0;       // small: pool index
  int64_t theta = 0;      // small: filter constant
  double arrive = 0;      // seconds from the start
  double done = 0;
  uint64_t digest = 0;
};
struct ServeSet {
  std::vector<const DevTable*> pool;  // small segments (columns 0 key INT, 1 val BIGINT)
  const DevTable* web = nullptr;
  const DevTable* fact = nullptr;
  const DevTable* dim = nullptr;
};

double u01(uint64_t seed, uint64_t i) {
  uint64_t x = seed * 0x9e3779b97f4a7c15ull + i * 0xd1b54a32d192ed03ull + 1;
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// The stream: Poisson arrivals at `rate` per second, 70 % small aggregates, 25 % web queries,
// 5 % star joins (when the set has them).
std::vector<Request> make_stream(const ServeSet& ss, double rate, int64_t count, uint64_t seed) {
  std::vector<Request> rs((size_t)count);
  double t = 0;
  for (int64_t i = 0; i < count; ++i) {
    Request& r = rs[(size_t)i];
    t += -std::log(1.0 - u01(seed, 4 * i)) / rate;
    r.arrive = t;
    const double x = u01(seed, 4 * i + 1);
    r.type = x < 0.70 || !ss.web ? kReqSmall : (x < 0.95 || !ss.fact ? kReqWeb : kReqJoin);
    r.pick = (int64_t)(u01(seed, 4 * i + 2) * (double)ss.pool.size());
    r.theta = 100000 + (int64_t)(u01(seed, 4 * i + 3) * 800000.0);
  }
  return rs;
}

uint64_t run_heavy(const ServeSet& ss, const Request& r) {
  if (r.type == kReqWeb) {
    int64_t out[40];
    C2Stats st;
    const int64_t n = query_c2(*ss.web, out, 10, &st);
    uint64_t d = (uint64_t)st.groups * 1000003ull + (uint64_t)st.pairs;
    for (int64_t i = 0; i < n; ++i) d = d * 31 + (uint64_t)out[i * 4 + 1];  // the top-10 counts (ties may reorder keys)
    return d;
  }
  int64_t out[4 * 64];
  C3Stats st;
  const int64_t g = query_c3(*ss.fact, *ss.dim, out, &st);
  uint64_t d = (uint64_t)g;
  for (int64_t i = 0; i < g * 4; ++i) d = d * 1099511628211ull + (uint64_t)out[i];
  return d;
}

// Device memory a request's temporaries take, estimated before it runs: from the installed sizing
// models when there are (predicted groups and join pairs), else from the plain rules' upper bounds
// (every filtered row a group or a pair). Used by the serving loop's HBM-budget admission.
int64_t need_of(const ServeSet& ss, const Request& r) {
  if (r.type == kReqSmall) {
    const DevTable& t = *ss.pool[(size_t)r.pick];
    return (t.cols[0].mx - t.cols[0].mn + 1) * 48 + (64 << 10);  // dense table + fetch buffers
  }
  if (r.type == kReqWeb) {
    const DevTable& t = *ss.web;
    const double sel = frac_below(t.cols[2], 19723 + 300) * frac_below(t.cols[1], 224);
    float f[WS_SIZER_INPUTS];
    const int64_t pred = size_groupby({&t.cols[0], &t.cols[3]}, t.rows, sel, f);
    const int64_t pass = (int64_t)((double)t.rows * sel) + 1;
    const int64_t groups = pred > 0 ? std::min<int64_t>(pred + pred / 4, pass) : pass;
    return pass * 24 + groups * 96;  // phase-1 records + both phases' groups and tables
  }
  const DevTable &f = *ss.fact, &d = *ss.dim;
  const double sel = frac_below(d.cols[1], 50);
  const int64_t nd = (int64_t)((double)d.rows * sel) + 1, nf = (int64_t)((double)f.rows * sel) + 1;
  int64_t pairs = f.rows;  // plain bound: every fact row matches
  if (ws_sizer* model = g_join_sizer.load()) {
    float jf[WS_SIZER_INPUTS];
    wsd::ck(ws_join_features(&d.cols[0].stats, nd, d.rows, &f.cols[0].stats, nf, jf));
    int64_t pred = 0;
    wsd::ck(ws_sizer_groups(model, jf, &pred));
    pairs = std::min<int64_t>(f.rows, pred + pred / 4);
  }
  return d.rows * 8 + pairs * 72;  // dimension relation + fact relation, join pairs and gathers
}

// HBM-budget admission: a request starts when its estimated temporaries fit the budget next to
// those of the requests running (one that alone exceeds the budget runs by itself).
struct Admission {
  int64_t budget = 0;  // 0: no budget
  int64_t reserved = 0, peak = 0, waits = 0;
  std::mutex mu;
  std::condition_variable cv;
  int64_t acquire(int64_t need) {
    if (budget <= 0) return 0;
    need = std::min(need, budget);
    std::unique_lock<std::mutex> g(mu);
    if (reserved + need > budget) ++waits;
    cv.wait(g, [&] { return reserved + need <= budget; });
    reserved += need;
    peak = std::max(peak, reserved);
    return need;
  }
  void release(int64_t held) {
    if (held <= 0) return;
    std::lock_guard<std::mutex> g(mu);
    reserved -= held;
    cv.notify_all();
  }
};

uint64_t run_alone(const ServeSet& ss, const Request& r) {
  if (r.type != kReqSmall) return run_heavy(ss, r);
  C4Query q{ss.pool[(size_t)r.pick], r.theta};
  return c4_alone(q, nullptr);
}

}  // namespace

namespace {

// Mode 0 (the reference's behaviour, engine.cpp:1262-1337): up to `workers` requests at a time,
// each run alone through the per-query path on its own thread and stream, in arrival order.
// Mode 1 (continuous batching): one thread serves the small aggregates, each time taking every
// small request that has arrived (up to `max_batch`) into one batched launch and fetch; `workers`
// threads serve the web / join queries at the same time, so their work overlaps on the device.
// Requests that arrive while others run join the next batch. Latency = completion - arrival.
void serve(const ServeSet& ss, std::vector<Request>& rs, int mode, int workers, int max_batch, Admission& adm) {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<size_t> small_q, heavy_q, all_q;
  bool closing = false;
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  std::vector<std::thread> th;
  std::exception_ptr err;
  auto guard_thread = [&](auto fn) {
    return [&, fn] {
      try {
        fn();
      } catch (...) {
        std::lock_guard<std::mutex> g(mu);
        if (!err) err = std::current_exception();
        closing = true;
        cv.notify_all();
      }
    };
  };
  if (mode == 0) {
    for (int w = 0; w < workers; ++w)
      th.emplace_back(guard_thread([&] {
        for (;;) {
          size_t i;
          {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] { return closing || !all_q.empty(); });
            if (all_q.empty()) return;
            i = all_q.front();
            all_q.pop_front();
          }
          const int64_t held = adm.acquire(need_of(ss, rs[i]));
          rs[i].digest = run_alone(ss, rs[i]);
          adm.release(held);
          rs[i].done = now();
        }
      }));
  } else {
    th.emplace_back(guard_thread([&] {
      for (;;) {
        std::vector<size_t> batch;
        {
          std::unique_lock<std::mutex> g(mu);
          cv.wait(g, [&] { return closing || !small_q.empty(); });
          if (small_q.empty()) return;
          while (!small_q.empty() && (int)batch.size() < max_batch) {
            batch.push_back(small_q.front());
            small_q.pop_front();
          }
        }
        std::vector<C4Query> qs(batch.size());
        std::vector<const C4Query*> qp(batch.size());
        for (size_t j = 0; j < batch.size(); ++j) {
          qs[j] = C4Query{ss.pool[(size_t)rs[batch[j]].pick], rs[batch[j]].theta};
          qp[j] = &qs[j];
        }
        std::vector<uint64_t> dg;
        std::vector<int64_t> gr;
        int64_t need = 0;
        for (size_t j : batch) need += need_of(ss, rs[j]);
        const int64_t held = adm.acquire(need);
        c4_batched(qp, &dg, gr);
        adm.release(held);
        const double d = now();
        for (size_t j = 0; j < batch.size(); ++j) rs[batch[j]].digest = dg[j], rs[batch[j]].done = d;
      }
    }));
    for (int w = 0; w < workers; ++w)
      th.emplace_back(guard_thread([&] {
        for (;;) {
          size_t i;
          {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] { return closing || !heavy_q.empty(); });
            if (heavy_q.empty()) return;
            i = heavy_q.front();
            heavy_q.pop_front();
          }
          const int64_t held = adm.acquire(need_of(ss, rs[i]));
          rs[i].digest = run_heavy(ss, rs[i]);
          adm.release(held);
          rs[i].done = now();
        }
      }));
  }
  // arrivals (open loop: requests are released at their arrival times whatever the backlog)
  for (size_t i = 0; i < rs.size(); ++i) {
    const double wait = rs[i].arrive - now();
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    std::lock_guard<std::mutex> g(mu);
    if (closing) break;
    if (mode == 0) all_q.push_back(i);
    else if (rs[i].type == kReqSmall) small_q.push_back(i);
    else heavy_q.push_back(i);
    cv.notify_all();
  }
  {
    std::lock_guard<std::mutex> g(mu);
    closing = true;
    cv.notify_all();
  }
  for (auto& t : th) t.join();
  if (err) std::rethrow_exception(err);
}

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const size_t k = (size_t)std::min<double>((double)v.size() - 1, std::floor(p * (double)(v.size() - 1) + 0.5));
  return v[k];
}

}  // namespace

extern "C" {

const char* wsq_last_error(void) { return t_err.c_str(); }

// A GPU-resident table: ncols columns of types[] (WS_T_*), n rows from host arrays, cut into
// 2^20-row segments; encode[c] != 0 stores column c frame-of-reference encoded per segment.
int wsq_table_load(int ncols, const int* types, const int* encode, int64_t n, const void* const* host, void** out) {
  return guard([&] {
    auto t = std::make_unique<DevTable>();
    t->rows = n;
    for (int64_t r = 0; r < n; r += kSegRows) t->seg_rows.push_back(std::min<int64_t>(kSegRows, n - r));
    t->cols.resize(ncols);
    for (int c = 0; c < ncols; ++c) {
      DevColumn& col = t->cols[c];
      col.t = types[c];
      const int w = width_of(col.t);
      int64_t r0 = 0;
      for (int64_t rows : t->seg_rows) {
        const uint8_t* src = (const uint8_t*)host[c] + r0 * w;
        for (int64_t i = 0; i < rows; ++i) {
          if (col.t == WS_T_F64) break;
          const int64_t v = w == 4 ? ((const int32_t*)src)[i] : ((const int64_t*)src)[i];
          col.mn = std::min(col.mn, v);
          col.mx = std::max(col.mx, v);
        }
        SegCol sc;
        if (encode && encode[c]) {
          std::vector<uint8_t> enc(ws_encoded_bytes_max(rows));
          int e = 0;
          size_t bytes = 0;
          wsd::ck(ws_encode_column(src, col.t == WS_T_F64 ? WS_P_F64 : w == 8 ? WS_P_I64 : WS_P_I32, rows, enc.data(), enc.size(),
                                   &e, &bytes));
          if (!e) throw SqlError("column does not fit an encoding", 1105);
          sc.enc = (uint8_t)e;
          sc.dev = devAlloc(bytes);
          h2d(sc.dev, enc.data(), bytes);
        } else {
          sc.dev = devAlloc((size_t)(rows * w));
          h2d(sc.dev, src, (size_t)(rows * w));
        }
        col.segs.push_back(sc);
        {  // segment statistics, once at load (the sizing model's inputs)
          ws_colview v{};
          v.p = sc.dev;
          v.t = (uint8_t)col.t;
          v.enc = sc.enc;
          ws_col_stats st;
          wsd::ck(ws_column_stats(&v, rows, nullptr, &st));
          if (col.segs.size() == 1) col.stats = st;
          else ws_col_stats_merge(&col.stats, &st);
        }
        r0 += rows;
      }
      // distinct values (a key column): one pass over a bitmap of the value span
      if (col.t != WS_T_F64 && n > 0 && (uint64_t)(col.mx - col.mn) < (1ull << 27)) {
        std::vector<uint64_t> seen((size_t)((col.mx - col.mn) / 64 + 1), 0);
        col.unique = true;
        for (int64_t i = 0; i < n && col.unique; ++i) {
          const int64_t v = w == 4 ? ((const int32_t*)host[c])[i] : ((const int64_t*)host[c])[i];
          const uint64_t b = (uint64_t)(v - col.mn);
          if (seen[b >> 6] >> (b & 63) & 1) col.unique = false;
          seen[b >> 6] |= 1ull << (b & 63);
        }
      }
    }
    *out = t.release();
  });
}
void wsq_table_free(void* t) { delete (DevTable*)t; }

int wsq_c1(void* t, int dense_on, int64_t* out_i, double* out_d, int64_t cap_rows, int64_t* groups) {
  return guard([&] { *groups = query_c1(*(DevTable*)t, out_i, out_d, cap_rows, dense_on != 0); });
}
// stats: pairs, groups, cap1, retries
int wsq_c2(void* t, int64_t limit, int64_t* out, int64_t* nrows, int64_t* stats) {
  return guard([&] {
    C2Stats st;
    *nrows = query_c2(*(DevTable*)t, out, limit, &st);
    if (stats) stats[0] = st.pairs, stats[1] = st.groups, stats[2] = st.cap1, stats[3] = st.retries;
  });
}
// The same query with WHERE event_date < date_lt AND region < region_lt.
int wsq_c2_where(void* t, int64_t date_lt, int64_t region_lt, int64_t limit, int64_t* out, int64_t* nrows, int64_t* stats) {
  return guard([&] {
    C2Stats st;
    *nrows = query_c2(*(DevTable*)t, out, limit, &st, date_lt, region_lt);
    if (stats) stats[0] = st.pairs, stats[1] = st.groups, stats[2] = st.cap1, stats[3] = st.retries;
  });
}
// stats: dim rows, fact rows after the join filter, join pairs, predicted pairs (0: no model).
// The dimension filter is attr < attr_lt (the measured query: 50).
int wsq_c3(void* fact, void* dim, int64_t attr_lt, int64_t* out, int64_t* groups, int64_t* stats) {
  return guard([&] {
    C3Stats st;
    *groups = query_c3(*(DevTable*)fact, *(DevTable*)dim, out, &st, attr_lt);
    if (stats) stats[0] = st.dim_rows, stats[1] = st.fact_rows, stats[2] = st.pairs, stats[3] = st.predicted_pairs;
  });
}
// The join output model (null: count first, as the reference). Not owned.
void wsq_set_join_sizer(ws_sizer* m) { g_join_sizer.store(m); }

// 1 (default): a star join whose filtered build keys are distinct and whose build columns are not
// read after the join runs as the bitmap filter fused into the aggregation; 0 keeps the hash join.
void wsq_set_join_rewrite(int on) { g_join_rewrite.store(on != 0); }

// The sizing model the drivers consult (null to go back to the range rule alone). Not owned.
void wsq_set_sizer(ws_sizer* m) { g_sizer.store(m); }

// C4: the generated segment set of the many-small-aggregations workload (seed; nq queries with
// rows lognormal between 4096 and 4M, distinct keys 10^U(1,5), theta uniform in [1e5, 9e5)).
int wsq_c4_load(uint64_t seed, int64_t nq, int64_t rows_override, void** out, int64_t* total_rows) {
  return guard([&] {
    auto set = new std::vector<std::pair<DevTable*, int64_t>>();
    std::vector<int64_t> rows((size_t)nq), distinct((size_t)nq), theta((size_t)nq);
    wsg_c4_shapes(seed, nq, rows.data(), distinct.data(), theta.data());
    int64_t tot = 0;
    try {
      for (int64_t q = 0; q < nq; ++q) {
        const int64_t n = rows_override > 0 ? rows_override : rows[(size_t)q];
        std::vector<int32_t> key((size_t)n);
        std::vector<int64_t> val((size_t)n);
        wsg_c4_data(seed, q, n, distinct[(size_t)q], key.data(), val.data());
        const int types[2] = {WS_T_I32, WS_T_I64};
        const void* host[2] = {key.data(), val.data()};
        void* t = nullptr;
        const int rc = wsq_table_load(2, types, nullptr, n, host, &t);
        if (rc) throw SqlError(wsq_last_error(), rc);
        set->push_back({(DevTable*)t, theta[(size_t)q]});
        tot += n;
      }
    } catch (...) {
      for (auto& e : *set) delete e.first;
      delete set;
      throw;
    }
    *out = set;
    if (total_rows) *total_rows = tot;
  });
}
void wsq_c4_free(void* set) {
  auto* v = (std::vector<std::pair<DevTable*, int64_t>>*)set;
  for (auto& e : *v) delete e.first;
  delete v;
}
// Runs every query of the set: batched (one launch + one fetch) or one by one (plain path).
// digests[q] / groups[q] per query; digests may be null (results fetched, not digested: what a
// timing run measures).
int wsq_c4_run(void* set, int batched, uint64_t* digests, int64_t* groups) {
  return guard([&] {
    auto* v = (std::vector<std::pair<DevTable*, int64_t>>*)set;
    std::vector<C4Query> qs;
    for (auto& e : *v) qs.push_back(C4Query{e.first, e.second});
    if (batched) {
      std::vector<const C4Query*> qp;
      for (auto& q : qs) qp.push_back(&q);
      std::vector<uint64_t> d;
      std::vector<int64_t> g;
      c4_batched(qp, digests ? &d : nullptr, g);
      for (size_t i = 0; i < qs.size(); ++i) {
        if (digests) digests[i] = d[i];
        groups[i] = g[i];
      }
    } else {
      for (size_t i = 0; i < qs.size(); ++i) {
        const uint64_t d = c4_alone(qs[i], &groups[i], digests != nullptr);
        if (digests) digests[i] = d;
      }
    }
  });
}

// The set prepared once for replays (see c4_prepare); wsq_c4_replay runs every query again with the
// recorded launches; wsq_c4_release frees it (the set must outlive it).
int wsq_c4_prepare(void* set, void** out) {
  return guard([&] {
    auto* v = (std::vector<std::pair<DevTable*, int64_t>>*)set;
    std::vector<C4Query> qs;
    for (auto& e : *v) qs.push_back(C4Query{e.first, e.second});
    std::vector<const C4Query*> qp;
    for (auto& q : qs) qp.push_back(&q);
    *out = c4_prepare(qp);
  });
}
int wsq_c4_replay(void* prepared, uint64_t* digests, int64_t* groups) {
  return guard([&] {
    auto* P = (C4Prepared*)prepared;
    std::vector<uint64_t> d;
    std::vector<int64_t> g;
    c4_replay(*P, digests ? &d : nullptr, g);
    for (size_t i = 0; i < g.size(); ++i) {
      if (digests) digests[i] = d[i];
      groups[i] = g[i];
    }
  });
}
// steps batches of the set, pipelined over ncopies prepared copies of it (see c4_pipeline); digests
// and groups of the last step.
int wsq_c4_pipeline(void* const* copies, int ncopies, int steps, uint64_t* digests, int64_t* groups) {
  return guard([&] {
    std::vector<uint64_t> d;
    std::vector<int64_t> g;
    std::vector<C4Prepared*> P;
    for (int k = 0; k < ncopies; ++k) P.push_back((C4Prepared*)copies[k]);
    c4_pipeline(P, steps, digests ? &d : nullptr, g);
    for (size_t i = 0; i < g.size(); ++i) {
      if (digests) digests[i] = d[i];
      groups[i] = g[i];
    }
  });
}
// steps batches of the set, each built from scratch and pipelined depth deep (see
// c4_fresh_pipeline); digests and groups of the last batch.
int wsq_c4_fresh_pipeline(void* set, int steps, int depth, uint64_t* digests, int64_t* groups) {
  return guard([&] {
    auto* v = (std::vector<std::pair<DevTable*, int64_t>>*)set;
    std::vector<C4Query> qs;
    for (auto& e : *v) qs.push_back(C4Query{e.first, e.second});
    std::vector<const C4Query*> qp;
    for (auto& q : qs) qp.push_back(&q);
    std::vector<uint64_t> d;
    std::vector<int64_t> g;
    c4_fresh_pipeline(qp, steps, depth, digests ? &d : nullptr, g);
    for (size_t i = 0; i < g.size(); ++i) {
      if (digests) digests[i] = d[i];
      groups[i] = g[i];
    }
  });
}
void* wsq_c4_plan_handle(void* prepared) { return ((C4Prepared*)prepared)->plan; }
void wsq_c4_release(void* prepared) { delete (C4Prepared*)prepared; }

// The autotuner (see autotune()) on the batched workload of a C4 set: knobs batch_tile_rows,
// batch_lds_slots, batch_blocks_per_cu. chosen[3] gets their values, ms[2] the median time of one
// batched run with the starting and the chosen values; *trials the configurations timed.
int wsq_c4_autotune(void* set, int reps, int64_t* chosen, double* ms, int* trials) {
  return guard([&] {
    auto* v = (std::vector<std::pair<DevTable*, int64_t>>*)set;
    std::vector<C4Query> qs;
    for (auto& e : *v) qs.push_back(C4Query{e.first, e.second});
    std::vector<const C4Query*> qp;
    for (auto& q : qs) qp.push_back(&q);
    {
      std::lock_guard<std::mutex> g(g_tuned_mu);
      g_tuned.erase(kC4Shape);  // tune from the current knob values
    }
    const std::vector<KnobSpace> space = {{"batch_tile_rows", {1024, 2048, 4096, 8192, 16384, 32768}},
                                          {"batch_lds_slots", {256, 512, 1024, 2048, 4096}},
                                          {"batch_blocks_per_cu", {1, 2, 4, 8, 16}}};
    std::vector<int64_t> g;
    const TuneResult r = autotune(kC4Shape, space, reps, [&] { c4_batched(qp, nullptr, g); });
    for (size_t i = 0; i < r.chosen.size(); ++i) chosen[i] = r.chosen[i].second;
    ms[0] = r.default_ms, ms[1] = r.best_ms;
    *trials = r.trials;
  });
}
// The autotuner on the web query over table t: knobs part_chunks_per_cu, part_scatter_lds_kib,
// part_merge_lds_kib (chosen[3]), as wsq_c4_autotune.
int wsq_c2_autotune(void* t, int reps, int64_t* chosen, double* ms, int* trials) {
  return guard([&] {
    {
      std::lock_guard<std::mutex> g(g_tuned_mu);
      g_tuned.erase(kC2Shape);
    }
    const std::vector<KnobSpace> space = {{"part_chunks_per_cu", {1, 2, 4}},
                                          {"part_scatter_lds_kib", {32, 48, 64, 96, 120}},
                                          {"part_merge_lds_kib", {32, 48, 60, 96, 120}}};
    int64_t out[40];
    const TuneResult r = autotune(kC2Shape, space, reps, [&] { query_c2(*(DevTable*)t, out, 10, nullptr); });
    for (size_t i = 0; i < r.chosen.size(); ++i) chosen[i] = r.chosen[i].second;
    ms[0] = r.default_ms, ms[1] = r.best_ms;
    *trials = r.trials;
  });
}

// The serving loop over a mixed open-loop stream (see serve()). pool: a C4 set whose tables serve
// the small aggregates; web / fact / dim may be null (the mix then shifts to what exists).
// stats: [0] completed, [1] wall seconds, [2] queries/s, [3] p50 ms, [4] p99 ms, [5] small p50,
// [6] small p99, [7] heavy p50, [8] heavy p99 (ms). digests / types per request (may be null).
// budget_bytes > 0: HBM-budget admission of the requests' estimated temporaries. stats gets 11
// values: the 9 above, then [9] requests that waited for memory, [10] peak estimated bytes held.
int wsq_serve(void* pool, void* web, void* fact, void* dim, double rate, int64_t count, int mode, int workers, int max_batch,
              uint64_t seed, int64_t budget_bytes, double* stats, uint64_t* digests, int32_t* types) {
  return guard([&] {
    auto* v = (std::vector<std::pair<DevTable*, int64_t>>*)pool;
    ServeSet ss;
    for (auto& e : *v) ss.pool.push_back(e.first);
    ss.web = (const DevTable*)web;
    ss.fact = (const DevTable*)fact;
    ss.dim = (const DevTable*)dim;
    std::vector<Request> rs = make_stream(ss, rate, count, seed);
    Admission adm;
    adm.budget = budget_bytes;
    serve(ss, rs, mode, workers, max_batch, adm);
    std::vector<double> all, sm, hv;
    double end = 0;
    for (size_t i = 0; i < rs.size(); ++i) {
      const double ms = (rs[i].done - rs[i].arrive) * 1e3;
      all.push_back(ms);
      (rs[i].type == kReqSmall ? sm : hv).push_back(ms);
      end = std::max(end, rs[i].done);
      if (digests) digests[i] = rs[i].digest;
      if (types) types[i] = rs[i].type;
    }
    const double wall = end - (rs.empty() ? 0 : rs[0].arrive);
    const double out[11] = {(double)rs.size(), wall, wall > 0 ? (double)rs.size() / wall : 0, pct(all, 0.5), pct(all, 0.99),
                            pct(sm, 0.5), pct(sm, 0.99), pct(hv, 0.5), pct(hv, 0.99), (double)adm.waits, (double)adm.peak};
    std::memcpy(stats, out, sizeof out);
  });
}
// One request of the stream (same seed, rate and index) run alone through the plain path.
int wsq_serve_alone(void* pool, void* web, void* fact, void* dim, double rate, int64_t count, uint64_t seed, int64_t index,
                    uint64_t* digest) {
  return guard([&] {
    auto* v = (std::vector<std::pair<DevTable*, int64_t>>*)pool;
    ServeSet ss;
    for (auto& e : *v) ss.pool.push_back(e.first);
    ss.web = (const DevTable*)web;
    ss.fact = (const DevTable*)fact;
    ss.dim = (const DevTable*)dim;
    std::vector<Request> rs = make_stream(ss, rate, count, seed);
    *digest = run_alone(ss, rs[(size_t)index]);
  });
}

}  // extern "C"
