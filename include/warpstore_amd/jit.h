// Drop-in C++ surface of the per-query plan compiler over the warpstore_amd C ABI
// (reference src/jit.h:14-75). Plans are compiled with hipRTC for the device's gfx950 target.
#pragma once
#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "gpu.h"

namespace gdb {

enum JitFn { JF_NOKEYS = 0, JF_GROUPED = 1, JF_DENSE = 2, JF_FLAG = 3, JF_FLAGM = 4 };

// A compiled plan as the engine sees it (the id it reports in EXPLAIN / PROFILE).
struct JitProgram {
  int id = 0;
  const ws_jit_program* handle = nullptr;
};

namespace wsd {
inline const ws_jit_program* handle(const JitProgram* p) { return p ? p->handle : nullptr; }
// One JitProgram per library plan, alive until the process ends (plans are never unloaded).
inline const JitProgram* wrap(const ws_jit_program* h) {
  static std::mutex mu;
  static std::unordered_map<const ws_jit_program*, std::unique_ptr<JitProgram>> all;
  if (!h) return nullptr;
  std::lock_guard<std::mutex> g(mu);
  auto& slot = all[h];
  if (!slot) {
    slot.reset(new JitProgram);
    slot->id = ws_jit_program_id(h);
    slot->handle = h;
  }
  return slot.get();
}
}  // namespace wsd

enum JitMode { JIT_OFF = 0, JIT_ON = 1, JIT_SYNC = 2 };
inline bool jitEnabled() { return ws_jit_enabled() != 0; }
inline JitMode jitMode() { return (JitMode)ws_jit_mode(); }
inline void jitSetMode(JitMode m) { ws_jit_set_mode((int)m); }
inline void jitSetDefaultMode(JitMode m) { ws_jit_set_default_mode((int)m); }
inline void jitSetCacheDir(const std::string& dir) { ws_jit_set_cache_dir(dir.c_str()); }

inline const JitProgram* jitGet(const ExprEnv& env, int ncols, const AggPlan& plan, JitFn fn, const std::string& kind) {
  const ws_jit_program* h = nullptr;
  wsd::ck(ws_jit_get(wsd::env(env), ncols, wsd::plan(plan), (int)fn, kind.c_str(), &h));
  return wsd::wrap(h);
}
inline void jitLaunch(const JitProgram* p, JitFn f, unsigned grid, unsigned block, size_t shmem, void** args) {
  wsd::ck(ws_jit_launch(wsd::handle(p), (int)f, grid, block, shmem, args));
}
inline bool progIsInt(const ExprEnv& env, Prog p) {
  int r = 0;
  wsd::ck(ws_prog_is_int(wsd::env(env), wsd::prog(p), &r));
  return r != 0;
}

struct JitQueryInfo {
  int compiled = 0, hits = 0, fromDisk = 0, background = 0;
  double compileMs = 0;
  std::vector<int> ids;
};
inline void jitBeginQuery(const std::string& sql) { ws_jit_begin_query(sql.c_str()); }
inline JitQueryInfo jitEndQuery(double queryMs) {
  ws_jit_query_info i{};
  ws_jit_end_query(queryMs, &i);
  JitQueryInfo o;
  o.compiled = i.compiled, o.hits = i.hits, o.fromDisk = i.from_disk, o.background = i.background;
  o.compileMs = i.compile_ms;
  o.ids.assign(i.ids, i.ids + i.nids);
  return o;
}

struct PlanCacheRow {
  int id;
  std::string kind;
  std::string origin;
  double compileMs;
  int64_t execs;
  double avgMs, lastMs;
  size_t sourceBytes;
  std::string sql;
};
inline std::vector<PlanCacheRow> jitPlanCache() {
  std::vector<ws_plan_cache_row> rows(64);
  int n = ws_jit_plan_cache(rows.data(), (int)rows.size());
  if (n > (int)rows.size()) {
    rows.resize((size_t)n);
    n = ws_jit_plan_cache(rows.data(), n);
  }
  std::vector<PlanCacheRow> out;
  for (int i = 0; i < n; ++i)
    out.push_back({rows[i].id, rows[i].kind, rows[i].origin, rows[i].compile_ms, rows[i].execs, rows[i].avg_ms,
                   rows[i].last_ms, rows[i].source_bytes, rows[i].sql});
  return out;
}
inline std::string jitSourceOf(int id) {
  const int n = ws_jit_source_of(id, nullptr, 0);
  std::string s((size_t)n + 1, '\0');
  ws_jit_source_of(id, s.data(), s.size());
  s.resize((size_t)n);
  return s;
}
inline void jitClearCache() { ws_jit_clear_cache(); }

}  // namespace gdb
