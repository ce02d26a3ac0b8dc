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

#include <fcntl.h>
#include <unistd.h>

#include <cstddef>

#include "ws_codegen.h"
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
  hipFunction_t fn_keybits = nullptr;             // selection plans: key bitmap of the passing rows (keys[0])
  hipFunction_t fn_staged = nullptr;              // selection and dense plans: the kernel with the sets in LDS
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

const char* kKernelMacro[] = {"WS_KERNEL_NOKEYS", "WS_KERNEL_GROUPED", "WS_KERNEL_DENSE_JIT", "WS_KERNEL_SELECT_JIT",
                              "WS_KERNEL_SELECT_JIT"};
const char* kKernelName[] = {"ws_q_nokeys", "ws_q_grouped", "ws_q_dense", "ws_q_select", "ws_q_select"};

// ---------------------------------------------------------------- persistent plan cache
//
// One file per compiled plan, named by a 128-bit digest of its key (hipRTC version, options,
// target and full source): a 64-byte header, the key (checked byte for byte on load, so a digest
// collision can only miss), then the gfx950 code object. The header carries digests of itself and
// of the image; a file that fails either check is discarded and the plan recompiled. Files are
// written to a private name and renamed into place, so readers never see a partial file.

constexpr char kPlanMagic[8] = {'W', 'S', 'G', 'F', 'X', 'P', 'L', 'N'};
constexpr uint32_t kPlanVersion = 2;
constexpr uint64_t kMaxImageBytes = 256ull << 20;

struct PlanHeader {
  char magic[8];
  uint32_t version, header_bytes;
  uint64_t key_digest[2];
  uint64_t key_bytes, image_bytes, image_digest;
  uint64_t header_digest;  // digest of the 56 bytes before it
};
static_assert(sizeof(PlanHeader) == 64, "plan file header is 64 bytes");

// 64-bit digest of a byte string: 8-byte little-endian words folded by multiply-rotate, the tail
// zero-padded, the length mixed in, then the splitmix64 finalizer.
uint64_t digest64(const void* data, size_t n, uint64_t seed) {
  const uint8_t* p = (const uint8_t*)data;
  uint64_t h = seed ^ (n * 0x9e3779b97f4a7c15ull);
  for (size_t i = 0; i < n; i += 8) {
    uint64_t w = 0;
    std::memcpy(&w, p + i, std::min<size_t>(8, n - i));
    h ^= w * 0xc2b2ae3d27d4eb4full;
    h = ((h << 29) | (h >> 35)) * 0x165667b19e3779f9ull;
  }
  h ^= h >> 30;
  h *= 0xbf58476d1ce4e5b9ull;
  h ^= h >> 27;
  h *= 0x94d049bb133111ebull;
  return h ^ (h >> 31);
}
std::string plan_file(const std::string& key) {
  char name[48];
  snprintf(name, sizeof name, "%016llx%016llx.wsplan", (unsigned long long)digest64(key.data(), key.size(), 1),
           (unsigned long long)digest64(key.data(), key.size(), 2));
  return g_disk_dir + "/" + name;
}
bool read_all(int fd, void* dst, size_t n) {
  uint8_t* p = (uint8_t*)dst;
  while (n) {
    const ssize_t r = ::read(fd, p, n);
    if (r <= 0) return false;
    p += r, n -= (size_t)r;
  }
  return true;
}
bool write_all(int fd, const void* src, size_t n) {
  const uint8_t* p = (const uint8_t*)src;
  while (n) {
    const ssize_t r = ::write(fd, p, n);
    if (r <= 0) return false;
    p += r, n -= (size_t)r;
  }
  return true;
}
// Loads the image stored under `key`; false when absent, for another key, or damaged.
bool plan_file_load(const std::string& path, const std::string& key, std::string& image) {
  const int fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd < 0) return false;
  struct Closer {
    int fd;
    ~Closer() { ::close(fd); }
  } closer{fd};
  PlanHeader h;
  if (!read_all(fd, &h, sizeof h) || std::memcmp(h.magic, kPlanMagic, 8) != 0 || h.version != kPlanVersion) return false;
  const bool header_ok = h.header_bytes == sizeof h && h.header_digest == digest64(&h, offsetof(PlanHeader, header_digest), 3) &&
                         h.image_bytes <= kMaxImageBytes;
  std::string k;
  if (header_ok && h.key_bytes == key.size()) {
    k.resize(key.size());
    if (!read_all(fd, k.data(), k.size()) || k != key) return false;  // another key under this name
    image.resize(h.image_bytes);
    if (read_all(fd, image.data(), image.size()) && digest64(image.data(), image.size(), 4) == h.image_digest) return true;
  } else if (header_ok) {
    return false;
  }
  fprintf(stderr, "[warpstore_amd] discarding damaged plan file %s; the plan is compiled again\n", path.c_str());
  return false;
}
void plan_file_store(const std::string& path, const std::string& key, const std::string& image) {
  PlanHeader h{};
  std::memcpy(h.magic, kPlanMagic, 8);
  h.version = kPlanVersion;
  h.header_bytes = sizeof h;
  h.key_digest[0] = digest64(key.data(), key.size(), 1);
  h.key_digest[1] = digest64(key.data(), key.size(), 2);
  h.key_bytes = key.size();
  h.image_bytes = image.size();
  h.image_digest = digest64(image.data(), image.size(), 4);
  h.header_digest = digest64(&h, offsetof(PlanHeader, header_digest), 3);
  static std::atomic<uint64_t> seq{0};
  const std::string tmp = path + ".part" + std::to_string((unsigned long long)::getpid()) + "." + std::to_string(seq++);
  const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0) return;
  const bool ok = write_all(fd, &h, sizeof h) && write_all(fd, key.data(), key.size()) && write_all(fd, image.data(), image.size());
  ::close(fd);
  if (!ok || ::rename(tmp.c_str(), path.c_str()) != 0) ::unlink(tmp.c_str());
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
  // extra hipRTC options, whitespace-separated (part of the disk cache key like every option)
  if (const char* extra = getenv("WARPSTORE_HIPRTC_OPTS")) {
    std::istringstream in(extra);
    for (std::string o; in >> o;) b.opts.push_back(o);
  }
  if (!g_disk_dir.empty()) {
    int maj = 0, min = 0;
    hiprtcVersion(&maj, &min);
    b.key = "hiprtc " + std::to_string(maj) + "." + std::to_string(min) + "\n";
    for (auto& o : b.opts) b.key += o + "\n";
    b.key += b.src;
    b.path = plan_file(b.key);
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
  if (p.fn_kind == WS_JF_DENSE) {
    WS_CK(hipModuleGetFunction(&p.fn_batch, mod, (std::string(kKernelName[p.fn_kind]) + "_batch").c_str()));
    WS_CK(hipModuleGetFunction(&p.fn_staged, mod, (std::string(kKernelName[p.fn_kind]) + "_staged").c_str()));
    WS_CK(hipFuncGetAttribute(&p.staged_static_lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, p.fn_staged));
  }
  if (p.fn_kind == WS_JF_FLAG || p.fn_kind == WS_JF_FLAGM) {
    WS_CK(hipModuleGetFunction(&p.fn_chunk, mod, (std::string(kKernelName[p.fn_kind]) + "_chunk").c_str()));
    WS_CK(hipModuleGetFunction(&p.fn_keybits, mod, (std::string(kKernelName[p.fn_kind]) + "_keybits").c_str()));
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
  if (!plan_file_load(b.path, b.key, image)) return false;
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
  if (!b.path.empty()) plan_file_store(b.path, b.key, image);
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
  if (fn == kJitDenseStaged) return p->fn_kind == WS_JF_DENSE ? (void*)p->fn_staged : nullptr;
  if (fn == kJitKeyBits) return p->fn_kind == WS_JF_FLAG || p->fn_kind == WS_JF_FLAGM ? (void*)p->fn_keybits : nullptr;
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
  return std::string(kKernelMacro[fn]) + "\n" + codegen::evaluator_source(env, ncols, plan, &rows);
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
    *out = codegen::program_is_int(*env, p) ? 1 : 0;
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
      const std::string ev = codegen::evaluator_source(*env, ncols, *plan, &rows);
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
    p.source = codegen::evaluator_source(*env, ncols, *plan, &rows);
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
