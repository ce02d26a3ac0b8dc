// Drop-in C++ operator surface over the warpstore_amd C ABI.
//
// Reproduces the declarations of the engine's src/gpu.h (reference gpu.h:12-201) as inline
// functions that call libwarpstore_amd, rethrowing failures as SqlError with the same codes
// (1041 out of memory, 1105 internal, 1235 needs a compiled plan). An engine built against this
// header instead of its CUDA gpu.cu keeps its call sites unchanged.
//
// Plan types: inside the engine, include this after the engine's own expr.h (its gdb::ExprEnv,
// AggPlan, DenseSpec, ColView, Prog, PType and SqlError are byte-compatible with the ws_* structs,
// which the static_asserts below check). Standalone builds (this repository's drivers and tests)
// define WARPSTORE_AMD_STANDALONE and get those names as aliases of the ws_* types.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../warpstore_amd.h"

#ifdef WARPSTORE_AMD_STANDALONE
namespace gdb {
using ExprEnv = ws_expr_env;
using AggPlan = ws_agg_plan;
using DenseSpec = ws_dense_spec;
using ColView = ws_colview;
using Prog = ws_prog;
using BatchSet = ws_batch_set;
using Instr = ws_instr;
enum class PType : uint8_t { I32 = 0, I64 = 1, F64 = 2, VEC = 5 };
inline int psize(PType p) { return p == PType::I32 ? 4 : 8; }
struct SqlError : std::runtime_error {
  int code;
  SqlError(const std::string& m, int c = 1064) : std::runtime_error(m), code(c) {}
};
}  // namespace gdb
#endif

namespace gdb {

static_assert(sizeof(ExprEnv) == sizeof(ws_expr_env) && sizeof(AggPlan) == sizeof(ws_agg_plan) &&
                  sizeof(DenseSpec) == sizeof(ws_dense_spec) && sizeof(ColView) == sizeof(ws_colview) &&
                  sizeof(Prog) == sizeof(ws_prog),
              "engine plan structs must match the ws_* wire layout");

namespace wsd {  // shim internals
inline void ck(int rc) {
  if (rc != WS_OK) throw SqlError(ws_last_error(), rc);
}
inline const ws_expr_env* env(const ExprEnv& e) { return reinterpret_cast<const ws_expr_env*>(&e); }
inline const ws_agg_plan* plan(const AggPlan& p) { return reinterpret_cast<const ws_agg_plan*>(&p); }
inline ws_prog prog(Prog p) {
  ws_prog q;
  static_assert(sizeof q == sizeof p, "Prog");
  std::memcpy(&q, &p, sizeof q);
  return q;
}
inline ws_colview view(const ColView& v) {
  ws_colview w;
  std::memcpy(&w, &v, sizeof w);
  return w;
}
}  // namespace wsd

// ------------------------------------------------------------------ device, memory (gpu.h:12-56)
inline void gpuInit() { wsd::ck(ws_init()); }
inline int gpuSmCount() { return ws_sm_count(); }
inline int gpuArch() { return ws_arch(); }
inline std::string gpuDescribe() {
  char b[256];
  wsd::ck(ws_describe(b, sizeof b));
  return b;
}
inline void* devAlloc(size_t bytes) {
  void* p = nullptr;
  wsd::ck(ws_dev_alloc(bytes, &p));
  return p;
}
inline void devFree(void* p) { ws_dev_free(p); }
inline void gpuSetVramLimit(int64_t bytes) { ws_set_vram_limit(bytes); }
inline int64_t gpuVramUsed() { return ws_vram_used(); }
inline double gpuVramFill() { return ws_vram_fill(); }
inline void* placeColumn(void* host, size_t bytes, bool& mapped, bool hostOnly = false) {
  int m = 0;
  void* p = nullptr;
  wsd::ck(ws_place_column(host, bytes, hostOnly ? 1 : 0, &m, &p));
  mapped = m != 0;
  return p;
}
inline void releaseColumn(void* dev, void* host, size_t bytes, bool mapped) { ws_release_column(dev, host, bytes, mapped); }
inline void* gpuPatchColumn(const void* oldDev, size_t bytes, int width, const int32_t* rows, const void* vals, int64_t m) {
  return ws_patch_column(oldDev, bytes, width, rows, vals, m);
}
inline void* vramAcquire(const void* host, size_t bytes) { return ws_vram_acquire(host, bytes); }
inline void vramRelease(void* dev, size_t bytes) { ws_vram_release(dev, bytes); }
inline void* hostMap(void* host, size_t bytes) { return ws_host_map(host, bytes); }
inline void hostUnmap(void* host) { ws_host_unmap(host); }
inline int64_t gpuVramLimit() { return ws_vram_limit(); }
inline int64_t gpuVramFree() { return ws_vram_free(); }
inline void* tmpAlloc(size_t bytes) {
  void* p = nullptr;
  wsd::ck(ws_tmp_alloc(bytes, &p));
  return p;
}
inline void gpuTrimPool() { ws_trim_pool(); }
inline void tmpFree(void* p) { ws_tmp_free(p); }
inline void h2d(void* dst, const void* src, size_t bytes) { wsd::ck(ws_h2d(dst, src, bytes)); }
inline void d2h(void* dst, const void* src, size_t bytes) { wsd::ck(ws_d2h(dst, src, bytes)); }
inline void d2d(void* dst, const void* src, size_t bytes) { wsd::ck(ws_d2d(dst, src, bytes)); }
inline void gpuSync() { wsd::ck(ws_sync()); }

// RAII temp device array from the library's caching allocator (gpu.h:59-79 semantics).
template <class T>
struct DevVec {
  T* p = nullptr;
  int64_t n = 0;
  DevVec() = default;
  explicit DevVec(int64_t count) : n(count) { p = count ? (T*)tmpAlloc(sizeof(T) * (size_t)count) : nullptr; }
  DevVec(const DevVec&) = delete;
  DevVec& operator=(const DevVec&) = delete;
  DevVec(DevVec&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
  DevVec& operator=(DevVec&& o) noexcept {
    if (this != &o) reset(), p = o.p, n = o.n, o.p = nullptr, o.n = 0;
    return *this;
  }
  ~DevVec() { reset(); }
  void reset() {
    if (p) tmpFree(p);
    p = nullptr, n = 0;
  }
  // Adopts a pool buffer returned by the library.
  static DevVec adopt(T* q, int64_t count) {
    DevVec v;
    v.p = q, v.n = count;
    return v;
  }
  std::vector<T> download(int64_t count = -1) const {
    std::vector<T> v(count < 0 ? n : count);
    if (!v.empty()) d2h(v.data(), p, sizeof(T) * v.size());
    return v;
  }
};

struct JitProgram;  // jit.h

namespace wsd {
inline const ws_jit_program* handle(const JitProgram* p);
}

// ------------------------------------------------------------------ aggregation (gpu.h:87-130)
class GpuAggregator {
 public:
  GpuAggregator(const AggPlan& plan, int64_t capacity, const DenseSpec* dense = nullptr) : plan_(plan) {
    wsd::ck(ws_agg_create(wsd::plan(plan), capacity, reinterpret_cast<const ws_dense_spec*>(dense), &h_));
  }
  ~GpuAggregator() { ws_agg_destroy(h_); }
  GpuAggregator(const GpuAggregator&) = delete;
  GpuAggregator& operator=(const GpuAggregator&) = delete;
  void addBatch(const ExprEnv& env, int ncols, int64_t n, const uint32_t* del) {
    wsd::ck(ws_agg_add_batch(h_, wsd::env(env), ncols, n, del));
  }
  void run(const ExprEnv& env, const JitProgram* jit) { wsd::ck(ws_agg_run(h_, wsd::env(env), wsd::handle(jit))); }
  void runWindow(const ExprEnv& env, const JitProgram* jit) {
    wsd::ck(ws_agg_run_window(h_, wsd::env(env), wsd::handle(jit)));
  }
  // The C ABI handle (for the batched form, ws_agg_run_batch / ws_agg_fetch_many; no reference
  // counterpart).
  ws_agg* handle() const { return h_; }
  int batches() const { return ws_agg_batches(h_); }
  int windows() const { return ws_agg_windows(h_); }
  int64_t rowsIn() const { return ws_agg_rows_in(h_); }
  bool dense() const { return ws_agg_dense(h_) != 0; }
  bool overflowed() {
    int o = 0;
    wsd::ck(ws_agg_overflowed(h_, &o));
    return o != 0;
  }
  // Returns number of groups; keys is groups*nkeys, accs is groups*naggs.
  int64_t fetch(std::vector<double>& keys, std::vector<double>& accs) {
    int64_t g = 0;
    wsd::ck(ws_agg_fetch_begin(h_, &g));
    const int nk = wsd::plan(plan_)->nkeys, na = wsd::plan(plan_)->naggs;
    keys.assign((size_t)(nk ? g * nk : 0), 0.0);
    accs.assign((size_t)(g * na), 0.0);
    wsd::ck(ws_agg_fetch_copy(h_, keys.data(), accs.data()));
    hi_.assign((size_t)na, 0);
    ws_agg_sum_high(h_, hi_.data());
    return g;
  }
  std::vector<uint8_t> sumBitsSlots() const {
    std::vector<uint8_t> f((size_t)wsd::plan(plan_)->naggs);
    ws_agg_sum_bits_slots(h_, f.data());
    return f;
  }
  const std::vector<int64_t>& sumHigh() const { return hi_; }
  int64_t compact(DevVec<double>& keys, DevVec<double>& accs) {
    double *k = nullptr, *a = nullptr;
    int64_t g = 0;
    wsd::ck(ws_agg_compact(h_, &k, &a, &g));
    const int nk = wsd::plan(plan_)->nkeys, na = wsd::plan(plan_)->naggs;
    keys = DevVec<double>::adopt(k, std::max<int64_t>(g * nk, 1));
    accs = DevVec<double>::adopt(a, std::max<int64_t>(g * na, 1));
    return g;
  }
  int64_t capacity() const { return ws_agg_capacity(h_); }
  int64_t groupCount() {
    int64_t c = 0;
    wsd::ck(ws_agg_group_count(h_, &c));
    return c;
  }

 private:
  AggPlan plan_;
  ws_agg* h_ = nullptr;
  std::vector<int64_t> hi_;
};

// ------------------------------------------------------------------ selection / gathers (gpu.h:133-195)
inline int64_t gpuSelect(const ExprEnv& env, bool hasWhere, Prog where, int64_t n, const uint32_t* del, int32_t* outIdx,
                         const JitProgram* jit = nullptr) {
  int64_t c = 0;
  wsd::ck(ws_select(wsd::env(env), hasWhere, wsd::prog(where), n, del, outIdx, wsd::handle(jit), &c));
  return c;
}
inline int64_t gpuSelectMulti(const ExprEnv& env, bool hasWhere, Prog where, const std::vector<int64_t>& bn,
                              const std::vector<const uint32_t*>& bdel, const std::vector<ColView>& bcols, int ncols,
                              DevVec<int32_t>& outIdx, std::vector<int64_t>& perBatch, const JitProgram* jit = nullptr) {
  perBatch.assign(bn.size(), 0);
  int32_t* idx = nullptr;
  int64_t c = 0;
  int64_t total = 0;
  for (int64_t x : bn) total += x;
  wsd::ck(ws_select_multi(wsd::env(env), hasWhere, wsd::prog(where), (int)bn.size(), bn.data(), bdel.data(),
                          reinterpret_cast<const ws_colview*>(bcols.data()), ncols, &idx, perBatch.data(), wsd::handle(jit),
                          &c));
  outIdx = DevVec<int32_t>::adopt(idx, idx ? total : 0);
  return c;
}
inline void gpuGatherMulti(const std::vector<ColView>& src, const std::vector<int64_t>& boff, PType t, int esize,
                           const int32_t* idx, int64_t m, void* dst) {
  wsd::ck(ws_gather_multi(reinterpret_cast<const ws_colview*>(src.data()), boff.data(), (int)src.size(), (int)t, esize, idx,
                          m, dst));
}
inline void gpuRemapCodes(const int32_t* in, int64_t n, const int32_t* map, int64_t mapN, int32_t* out) {
  wsd::ck(ws_remap_codes(in, n, map, mapN, out));
}
inline void gpuBitsToDouble(const double* src, int64_t n, double* dst) { wsd::ck(ws_bits_to_double(src, n, dst)); }
inline void gpuEvalAtInt(const ExprEnv& env, Prog prog, const int32_t* idx, int64_t m, int64_t* out) {
  wsd::ck(ws_eval_at_int(wsd::env(env), wsd::prog(prog), idx, m, out));
}
inline std::vector<int64_t> gpuDistinctInts(const ExprEnv& env, Prog prog, const int32_t* idx, int64_t m, int64_t base,
                                            int64_t span) {
  std::vector<int64_t> out(1);
  int64_t n = 0;
  wsd::ck(ws_distinct_ints(wsd::env(env), wsd::prog(prog), idx, m, base, span, out.data(), 0, &n));
  out.assign((size_t)n, 0);
  if (n) wsd::ck(ws_distinct_ints(wsd::env(env), wsd::prog(prog), idx, m, base, span, out.data(), n, &n));
  return out;
}
struct TempColSpec {
  int8_t src = 0;
  int idx = 0, hi = -1, cnt = -1;
  uint8_t kind = 0;
  uint8_t keyInt = 0, sumBits = 0;
  uint8_t out = 2;
  double div = 1;
  double mul = 1;
};
inline uint32_t* gpuBitsFromColumn(const void* col, int out, int64_t n, int64_t base, int64_t nbits, bool* sawNull) {
  int s = 0;
  uint32_t* bits = nullptr;
  wsd::ck(ws_bits_from_column(col, out, n, base, nbits, &s, &bits));
  *sawNull = s != 0;
  return bits;
}
inline void* gpuTempColumn(const double* keys, int nk, const double* accs, int na, int64_t g, const TempColSpec& s, double& mn,
                           double& mx) {
  ws_temp_col_spec w{s.src, s.idx, s.hi, s.cnt, s.kind, s.keyInt, s.sumBits, s.out, s.div, s.mul};
  void* out = nullptr;
  wsd::ck(ws_temp_column(keys, nk, accs, na, g, &w, &mn, &mx, &out));
  return out;
}
inline void gpuEvalAt(const ExprEnv& env, Prog prog, const int32_t* idx, int64_t m, double* out) {
  wsd::ck(ws_eval_at(wsd::env(env), wsd::prog(prog), idx, m, out));
}
inline void gpuGather(const void* src, PType t, const int32_t* idx, int64_t m, void* dst) {
  wsd::ck(ws_gather(src, (int)t, idx, m, dst));
}
inline void gpuGatherEnc(const ColView& v, PType t, const int32_t* idx, int64_t m, void* dst) {
  const ws_colview w = wsd::view(v);
  wsd::ck(ws_gather_enc(&w, (int)t, idx, m, dst));
}
inline void gpuKeyBitmap(ColView key, int64_t n, int64_t lo, int64_t nbits, uint32_t* bits) {
  wsd::ck(ws_key_bitmap(wsd::view(key), n, lo, nbits, bits));
}
inline void gpuSetBits(uint32_t* bits, const int32_t* idx, int64_t m) { wsd::ck(ws_set_bits(bits, idx, m)); }
inline int64_t gpuBitmapCount(const uint32_t* bits, int64_t words) {  // (no reference counterpart)
  int64_t c = 0;
  wsd::ck(ws_bitmap_count(bits, words, &c));
  return c;
}
inline void gpuGatherNull(const void* src, PType t, const int32_t* idx, int64_t m, void* dst) {
  wsd::ck(ws_gather_null(src, (int)t, idx, m, dst));
}
inline void gpuGatherBytes(const void* src, int esize, const int32_t* idx, int64_t m, void* dst, bool nulls) {
  wsd::ck(ws_gather_bytes(src, esize, idx, m, dst, nulls ? 1 : 0));
}

// ------------------------------------------------------------------ joins (gpu.h:178-191)
inline int64_t gpuHashJoinParts(ColView buildKey, const std::vector<int64_t>& bOff, ColView probeKey,
                                const std::vector<int64_t>& pOff, DevVec<int32_t>& outBuild, DevVec<int32_t>& outProbe,
                                std::vector<int64_t>* outOff) {
  int32_t *b = nullptr, *p = nullptr;
  int64_t m = 0;
  std::vector<int64_t> off(pOff.size(), 0);
  wsd::ck(ws_hash_join_parts(wsd::view(buildKey), bOff.data(), (int)bOff.size() - 1, wsd::view(probeKey), pOff.data(),
                             (int)pOff.size() - 1, &b, &p, off.data(), &m));
  outBuild = DevVec<int32_t>::adopt(b, b ? m : 0);
  outProbe = DevVec<int32_t>::adopt(p, p ? m : 0);
  if (outOff) *outOff = off;
  return m;
}
inline int64_t gpuHashJoin(ColView buildKey, int64_t nb, ColView probeKey, int64_t np, DevVec<int32_t>& outBuild,
                           DevVec<int32_t>& outProbe) {
  return gpuHashJoinParts(buildKey, {0, nb}, probeKey, {0, np}, outBuild, outProbe, nullptr);
}
inline int64_t gpuHashJoinOuter(ColView buildKey, int64_t nb, ColView probeKey, int64_t np, DevVec<int32_t>& outBuild,
                                DevVec<int32_t>& outProbe, bool full) {
  int32_t *b = nullptr, *p = nullptr;
  int64_t m = 0;
  wsd::ck(ws_hash_join_outer(wsd::view(buildKey), nb, wsd::view(probeKey), np, &b, &p, full ? 1 : 0, &m));
  outBuild = DevVec<int32_t>::adopt(b, std::max<int64_t>(m, 1));
  outProbe = DevVec<int32_t>::adopt(p, std::max<int64_t>(m, 1));
  return m;
}
inline int64_t gpuCrossJoin(int64_t na, int64_t nb, DevVec<int32_t>& outA, DevVec<int32_t>& outB) {
  int32_t *a = nullptr, *b = nullptr;
  int64_t m = 0;
  wsd::ck(ws_cross_join(na, nb, &a, &b, &m));
  outA = DevVec<int32_t>::adopt(a, std::max<int64_t>(m, 1));
  outB = DevVec<int32_t>::adopt(b, std::max<int64_t>(m, 1));
  return m;
}

// ------------------------------------------------------------------ sorting (gpu.h:197-201)
inline DevVec<int32_t> gpuSortPerm(const std::vector<std::pair<const double*, bool>>& keys, int64_t m) {
  std::vector<const double*> k;
  std::vector<uint8_t> desc;
  for (auto& x : keys) k.push_back(x.first), desc.push_back(x.second ? 1 : 0);
  int32_t* perm = nullptr;
  wsd::ck(ws_sort_perm(k.data(), desc.data(), (int)k.size(), m, &perm));
  return DevVec<int32_t>::adopt(perm, m);
}
// First min(k, m) entries of gpuSortPerm(keys, m) (ORDER BY ... LIMIT k).
inline DevVec<int32_t> gpuTopKPerm(const std::vector<std::pair<const double*, bool>>& keys, int64_t m, int64_t k) {
  std::vector<const double*> kp;
  std::vector<uint8_t> desc;
  for (auto& x : keys) kp.push_back(x.first), desc.push_back(x.second ? 1 : 0);
  int32_t* perm = nullptr;
  wsd::ck(ws_topk_perm(kp.data(), desc.data(), (int)kp.size(), m, k, &perm));
  return DevVec<int32_t>::adopt(perm, std::min(k, m));
}
inline void gpuMapRank(double* keys, const int32_t* rank, int64_t m) { wsd::ck(ws_map_rank(keys, rank, m)); }
inline std::vector<int32_t> gpuSortPermHost(const void* hostCol, PType t, int64_t n) {
  std::vector<int32_t> out((size_t)n);
  wsd::ck(ws_sort_perm_host(hostCol, (int)t, n, out.data()));
  return out;
}

}  // namespace gdb
