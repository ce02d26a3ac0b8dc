// Device setup, memory, staging, profiling and the C ABI entry points for them
// (reference src/gpu.cu:32-186 covers the same ground for CUDA).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ws_internal.h"

namespace ws {

// ------------------------------------------------------------------ errors

static thread_local std::string t_last_error;
thread_local hipStream_t t_stream_override = nullptr;
void set_last_error(const std::string& msg) { t_last_error = msg; }

void hip_fail(hipError_t e, const char* what, const char* file, int line) {
  (void)hipGetLastError();  // clear the sticky error so the next call starts clean
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) throw Error("GPU out of memory", WS_ERR_OOM);
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error: %s at %s (%s:%d)", hipGetErrorString(e), what, file, line);
  throw Error(buf, WS_ERR_INTERNAL);
}

// ------------------------------------------------------------------ device

static std::once_flag g_init_once;
static int g_cus = 256;
static std::string g_arch = "gfx950";
static bool g_init_failed = false;
static std::string g_init_error;

static void init_device() {
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) {
    g_init_failed = true;
    g_init_error = "no HIP device available";
    (void)hipGetLastError();
    return;
  }
  e = hipSetDevice(0);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, 0);
  if (e != hipSuccess) {
    g_init_failed = true;
    g_init_error = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
    (void)hipGetLastError();
    return;
  }
  g_cus = prop.multiProcessorCount;
  g_arch = prop.gcnArchName;
  (void)hipGetLastError();
}

void ensure_init() {
  std::call_once(g_init_once, init_device);
  if (g_init_failed) throw Error(g_init_error, WS_ERR_INTERNAL);
}
int cu_count() {
  ensure_init();
  return g_cus;
}
const std::string& arch_name() {
  ensure_init();
  return g_arch;
}

// ------------------------------------------------------------------ memory

static std::atomic<int64_t> g_vram_used{0};
static int64_t g_vram_limit = INT64_MAX;

// Query temporaries come from a caching allocator. A freed block goes into the freeing thread's
// cache, keyed by size class, and is handed out again only to that thread: all of a thread's work
// is issued in order on its per-thread stream, so the block's earlier uses precede its next use
// without any synchronization or driver call. hipMalloc runs only on a cache miss (warm-up), and
// a miss that finds no free device memory releases the cached blocks and retries once. Blocks of
// threads that exited, and blocks freed by a thread other than the one that took them (its stream
// may still be using them), are reused by others after a device synchronization.
namespace {
size_t size_class(size_t b) {
  if (b <= 512) return 512;
  if (b <= ((size_t)1 << 20)) {
    size_t c = 1024;
    while (c < b) c <<= 1;
    return c;
  }
  const int lg = 63 - __builtin_clzll((unsigned long long)b);
  const size_t step = (size_t)1 << (lg - 3);  // 8 classes per octave: at most 12.5 % slack
  return (b + step - 1) / step * step;
}
std::mutex g_mem_mu;
struct LiveBlock {
  size_t cls;      // size class
  uint64_t owner;  // thread that took it last (t_id); a free by another thread is a cross-thread free
};
std::unordered_map<void*, LiveBlock> g_live;                // every block this allocator made
std::atomic<uint64_t> g_thread_seq{0};
thread_local const uint64_t t_id = ++g_thread_seq;         // never reused, unlike std::thread::id
std::unordered_map<size_t, std::vector<void*>> g_orphans;  // cached blocks of exited threads
std::atomic<int64_t> g_orphan_count{0};
struct BlockCache {
  std::unordered_map<size_t, std::vector<void*>> free;
  ~BlockCache() {  // thread exit: host bookkeeping only, no runtime calls
    std::lock_guard<std::mutex> g(g_mem_mu);
    for (auto& kv : free)
      for (void* p : kv.second) {
        g_orphans[kv.first].push_back(p);
        ++g_orphan_count;
      }
  }
};
thread_local BlockCache t_cache;

void release_cached() {  // frees this thread's cached blocks and the orphans
  (void)hipDeviceSynchronize();
  std::vector<void*> v;
  {
    std::lock_guard<std::mutex> g(g_mem_mu);
    for (auto& kv : t_cache.free)
      for (void* p : kv.second) v.push_back(p);
    t_cache.free.clear();
    for (auto& kv : g_orphans)
      for (void* p : kv.second) v.push_back(p);
    g_orphans.clear();
    g_orphan_count = 0;
    for (void* p : v) g_live.erase(p);
  }
  for (void* p : v) (void)hipFree(p);
  (void)hipGetLastError();
}
}  // namespace

void trim_pool() { release_cached(); }

void* tmp_alloc(size_t bytes) {
  ensure_init();
  if (!bytes) return nullptr;
  const size_t c = size_class(bytes);
  auto it = t_cache.free.find(c);
  if (it != t_cache.free.end() && !it->second.empty()) {
    void* p = it->second.back();  // cached blocks are this thread's: owner already t_id
    it->second.pop_back();
    return p;
  }
  if (g_orphan_count.load(std::memory_order_relaxed) > 0) {
    void* p = nullptr;
    {
      std::lock_guard<std::mutex> g(g_mem_mu);
      auto o = g_orphans.find(c);
      if (o != g_orphans.end() && !o->second.empty()) {
        p = o->second.back();
        o->second.pop_back();
        --g_orphan_count;
        g_live[p].owner = t_id;
      }
    }
    if (p) {
      WS_CK(hipDeviceSynchronize());  // the previous owner's last work on the block has ended
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, c) != hipSuccess) {
    (void)hipGetLastError();
    release_cached();
    WS_CK(hipMalloc(&p, c));
  }
  {
    std::lock_guard<std::mutex> g(g_mem_mu);
    g_live[p] = LiveBlock{c, t_id};
  }
  return p;
}
void tmp_free(void* p) {
  if (!p) return;
  size_t c = 0;
  {
    std::lock_guard<std::mutex> g(g_mem_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) return;  // not a block of this allocator
    c = it->second.cls;
    if (it->second.owner != t_id) {  // taken by another thread: its stream may still use the block
      g_orphans[c].push_back(p);
      ++g_orphan_count;
      return;
    }
  }
  t_cache.free[c].push_back(p);
}

// Pinned staging for small host-to-device uploads (launch descriptors, plans): a per-thread ring
// of buffers, each with an event marking when the copy out of it has run. A slot is reused after
// its event completes, which is normally long past since the ring holds kStageSlots uploads.
constexpr int kStageSlots = 16;
struct StageSlot {
  uint8_t* buf = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};
struct Stage {
  StageSlot slot[kStageSlots];
  int next = 0;
  int cur = -1;
  ~Stage() {
    // Thread exit: a buffer may still feed an in-flight copy; wait for it before freeing.
    for (StageSlot& s : slot) {
      if (s.ev) {
        if (s.pending) (void)hipEventSynchronize(s.ev);
        (void)hipEventDestroy(s.ev);
      }
      if (s.buf) (void)hipHostFree(s.buf);
    }
  }
};
static thread_local Stage t_stage;

// Initial capacity of each ring buffer, all allocated at a thread's first upload (a page-locked
// allocation costs tens of microseconds: none should land inside a query once the ring is warm).
constexpr size_t kStageInitBytes = 128 << 10;
uint8_t* pinned_stage(size_t bytes) {
  ensure_init();
  Stage& st = t_stage;
  if (!st.slot[0].ev)
    for (StageSlot& s : st.slot) {
      WS_CK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
      s.cap = kStageInitBytes;
      WS_CK(hipHostMalloc((void**)&s.buf, s.cap, hipHostMallocDefault));
    }
  StageSlot& s = st.slot[st.next];
  st.cur = st.next;
  st.next = (st.next + 1) % kStageSlots;
  if (s.pending) {
    WS_CK(hipEventSynchronize(s.ev));
    s.pending = false;
  }
  if (bytes > s.cap) {
    if (s.buf) WS_CK(hipHostFree(s.buf));
    s.buf = nullptr;
    s.cap = std::max<size_t>(bytes, kStageInitBytes);
    WS_CK(hipHostMalloc((void**)&s.buf, s.cap, hipHostMallocDefault));
  }
  return s.buf;
}
void pinned_stage_done() {
  StageSlot& s = t_stage.slot[t_stage.cur];
  WS_CK(hipEventRecord(s.ev, stream()));
  s.pending = true;
}
void h2d_sync(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  WS_CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
  WS_CK(hipStreamSynchronize(stream()));
}
// Pinned per-thread landing buffer for small device-to-host reads: a copy into pageable memory
// goes through the runtime's own staging and costs an extra wait.
constexpr size_t kLandBytes = 64 << 10;
struct Landing {
  uint8_t* buf = nullptr;
  uint32_t* dev = nullptr;  // the same memory as a kernel sees it
  ~Landing() {
    if (buf) (void)hipHostFree(buf);  // every read through it ended with a stream sync
  }
};
static thread_local Landing t_land;
static Landing& landing() {
  ensure_init();
  if (!t_land.buf) {
    WS_CK(hipHostMalloc((void**)&t_land.buf, kLandBytes, hipHostMallocDefault));
    WS_CK(hipHostGetDevicePointer((void**)&t_land.dev, t_land.buf, 0));
  }
  return t_land;
}
void d2h_sync(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  if (bytes <= kLandBytes) {
    Landing& l = landing();
    WS_CK(hipMemcpyAsync(l.buf, src, bytes, hipMemcpyDeviceToHost, stream()));
    WS_CK(hipStreamSynchronize(stream()));
    std::memcpy(dst, l.buf, bytes);
    return;
  }
  WS_CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
  WS_CK(hipStreamSynchronize(stream()));
}

// One kernel copies every piece into the landing buffer, then one stream sync: a single round trip
// for the handful of counters and flags a query step reads back.
struct D2HList {
  const uint32_t* src[kD2HMaxPieces];
  uint32_t words[kD2HMaxPieces];
  uint32_t off[kD2HMaxPieces];
  int n;
};
__global__ void d2h_gather_k(D2HList l, uint32_t* dst) {
  for (int i = 0; i < l.n; ++i)
    for (uint32_t w = threadIdx.x; w < l.words[i]; w += blockDim.x) dst[l.off[i] + w] = l.src[i][w];
}
void d2h_gather_sync(const D2HPiece* pieces, int n) {
  if (n <= 0) return;
  if (n > kD2HMaxPieces) throw Error("d2h_gather_sync: too many pieces", WS_ERR_INTERNAL);
  D2HList l{};
  uint32_t off = 0;
  for (int i = 0; i < n; ++i) {
    if (pieces[i].bytes % 4 || (uintptr_t)pieces[i].src % 4)
      throw Error("d2h_gather_sync: pieces must be 4-byte multiples", WS_ERR_INTERNAL);
    l.src[i] = (const uint32_t*)pieces[i].src;
    l.words[i] = (uint32_t)(pieces[i].bytes / 4);
    l.off[i] = off;
    off += l.words[i];
  }
  if ((size_t)off * 4 > kLandBytes) throw Error("d2h_gather_sync: pieces too large", WS_ERR_INTERNAL);
  l.n = n;
  Landing& land = landing();
  hipLaunchKernelGGL(d2h_gather_k, dim3(1), dim3(256), 0, stream(), l, land.dev);
  check_launch("d2h_gather");
  WS_CK(hipStreamSynchronize(stream()));
  for (int i = 0; i < n; ++i) std::memcpy(pieces[i].dst, land.buf + (size_t)l.off[i] * 4, pieces[i].bytes);
}
void h2d_async_staged(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  uint8_t* st = pinned_stage(bytes);
  std::memcpy(st, src, bytes);
  WS_CK(hipMemcpyAsync(dst, st, bytes, hipMemcpyHostToDevice, stream()));
  pinned_stage_done();
}
uint8_t* upload_blob(const Blob& b) {
  uint8_t* d = (uint8_t*)tmp_alloc(std::max<size_t>(b.host.size(), 16));
  try {
    h2d_async_staged(d, b.host.data(), b.host.size());
  } catch (...) {
    tmp_free(d);
    throw;
  }
  return d;
}

// ------------------------------------------------------------------ profiling

struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};
static thread_local bool t_prof = false;
static thread_local std::string t_prof_only;  // record only this kernel name ("" = all)
static thread_local std::vector<ProfRec> t_prof_recs;
static thread_local std::vector<hipEvent_t> t_prof_free;

bool prof_enabled() { return t_prof; }
static hipEvent_t prof_event() {
  if (!t_prof_free.empty()) {
    hipEvent_t e = t_prof_free.back();
    t_prof_free.pop_back();
    return e;
  }
  hipEvent_t e;
  WS_CK(hipEventCreate(&e));
  return e;
}
ProfScope::ProfScope(const char* kernel) {
  if (!t_prof) return;
  if (!t_prof_only.empty() && t_prof_only != kernel) return;
  ProfRec r{kernel, prof_event(), prof_event()};
  WS_CK(hipEventRecord(r.a, stream()));
  slot = (int)t_prof_recs.size();
  t_prof_recs.push_back(r);
}
ProfScope::~ProfScope() {
  if (slot >= 0) (void)hipEventRecord(t_prof_recs[slot].b, stream());
}

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) hip_fail(e, what, __FILE__, __LINE__);
}

// ------------------------------------------------------------------ fills

template <class T>
__global__ void fill_k(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
void fill_u64(u64* p, int64_t n, u64 v) {
  if (n <= 0) return;
  ProfScope ps("fill");
  hipLaunchKernelGGL(fill_k<u64>, dim3(grid_for(n)), dim3(256), 0, stream(), p, n, v);
  check_launch("fill_u64");
}
void fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  if (n <= 0) return;
  ProfScope ps("fill");
  hipLaunchKernelGGL(fill_k<uint32_t>, dim3(grid_for(n)), dim3(256), 0, stream(), p, n, v);
  check_launch("fill_u32");
}
void fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  if (n <= 0) return;
  ProfScope ps("fill");
  hipLaunchKernelGGL(fill_k<uint8_t>, dim3(grid_for(n)), dim3(256), 0, stream(), p, n, v);
  check_launch("fill_u8");
}
void fill_i32(int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return;
  ProfScope ps("fill");
  hipLaunchKernelGGL(fill_k<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream(), p, n, v);
  check_launch("fill_i32");
}

// Plain streaming copy (16-byte vectors), used to calibrate byte counters against a known size.
__global__ void copy_k(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace ws

using namespace ws;

// ------------------------------------------------------------------ C ABI: errors, device, memory

extern "C" {

const char* ws_last_error(void) { return t_last_error.c_str(); }

int ws_init(void) { return guard([] { ensure_init(); }); }
int ws_sm_count(void) {
  try {
    return cu_count();
  } catch (...) {
    return 0;
  }
}
int ws_arch(void) {
  try {
    const std::string& a = arch_name();  // "gfx950..." -> 950
    int v = 0;
    for (size_t i = 3; i < a.size() && isdigit((unsigned char)a[i]); ++i) v = v * 10 + (a[i] - '0');
    return v;
  } catch (...) {
    return 0;
  }
}
int ws_describe(char* buf, size_t cap) {
  return guard([&] {
    ensure_init();
    hipDeviceProp_t prop;
    WS_CK(hipGetDeviceProperties(&prop, 0));
    size_t fr = 0, tot = 0;
    WS_CK(hipMemGetInfo(&fr, &tot));
    snprintf(buf, cap, "%s (%s, %d CUs, %.1f GiB free / %.1f GiB)", prop.name, prop.gcnArchName, prop.multiProcessorCount,
             fr / 1073741824.0, tot / 1073741824.0);
  });
}

int ws_dev_alloc(size_t bytes, void** out) {
  return guard([&] {
    ensure_init();
    *out = nullptr;
    if (bytes) WS_CK(hipMalloc(out, bytes));
  });
}
void ws_dev_free(void* p) {
  if (p) (void)hipFree(p);
}
int ws_tmp_alloc(size_t bytes, void** out) { return guard([&] { *out = tmp_alloc(bytes); }); }
int ws_pinned_alloc(size_t bytes, void** out) {
  return guard([&] {
    ensure_init();
    *out = nullptr;
    if (bytes) WS_CK(hipHostMalloc(out, bytes, hipHostMallocDefault));
  });
}
void ws_pinned_free(void* p) {
  if (p) (void)hipHostFree(p);
}
void ws_trim_pool(void) { trim_pool(); }
void ws_tmp_free(void* p) { tmp_free(p); }
int ws_h2d(void* dst, const void* src, size_t b) { return guard([&] { h2d_sync(dst, src, b); }); }
int ws_d2h(void* dst, const void* src, size_t b) { return guard([&] { d2h_sync(dst, src, b); }); }
int ws_d2h_async(void* dst, const void* src, size_t b) {
  return guard([&] {
    if (b) WS_CK(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, stream()));
  });
}
int ws_d2d(void* dst, const void* src, size_t b) {
  return guard([&] {
    if (!b) return;
    WS_CK(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, stream()));
    WS_CK(hipStreamSynchronize(stream()));
  });
}
int ws_sync(void) { return guard([] { WS_CK(hipStreamSynchronize(stream())); }); }

// VRAM tiering: columns always live in device memory here; the host tier (mapped host memory)
// is not provided, so a column that does not fit reports out of memory.
void ws_set_vram_limit(int64_t bytes) { g_vram_limit = bytes; }
int64_t ws_vram_used(void) { return g_vram_used; }
double ws_vram_fill(void) {
  if (g_vram_limit != INT64_MAX) return (double)g_vram_used / (double)g_vram_limit;
  size_t fr = 0, tot = 1;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return 1.0 - (double)fr / (double)tot;
}
int ws_place_column(void* host, size_t bytes, int host_only, int* mapped, void** out) {
  return guard([&] {
    *mapped = 0;
    *out = nullptr;
    if (!bytes) return;
    if (host_only || g_vram_used + (int64_t)bytes > g_vram_limit)
      throw Error("column does not fit the VRAM budget (no host tier)", WS_ERR_OOM);
    void* p = nullptr;
    WS_CK(hipMalloc(&p, bytes));
    g_vram_used += (int64_t)bytes;
    if (hipMemcpy(p, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(p);
      g_vram_used -= (int64_t)bytes;
      throw Error("column upload failed", WS_ERR_INTERNAL);
    }
    *out = p;
  });
}
void ws_release_column(void* dev, void* host, size_t bytes, int mapped) {
  (void)host;
  if (!dev || mapped) return;
  (void)hipFree(dev);
  g_vram_used -= (int64_t)bytes;
}
void* ws_vram_acquire(const void* host, size_t bytes) {
  if (!bytes || g_vram_used + (int64_t)bytes > g_vram_limit) return nullptr;
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipMemcpy(p, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(p);
    return nullptr;
  }
  g_vram_used += (int64_t)bytes;
  return p;
}
void ws_vram_release(void* dev, size_t bytes) {
  if (!dev) return;
  (void)hipFree(dev);
  g_vram_used -= (int64_t)bytes;
}
void* ws_host_map(void* host, size_t bytes) {
  (void)host, (void)bytes;
  return nullptr;  // no mapped host tier
}
void ws_host_unmap(void* host) { (void)host; }
int64_t ws_vram_limit(void) { return g_vram_limit; }
int64_t ws_vram_free(void) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return (int64_t)fr;
}
void* ws_patch_column(const void* old_dev, size_t bytes, int width, const int32_t* rows, const void* vals, int64_t m) {
  (void)old_dev, (void)bytes, (void)width, (void)rows, (void)vals, (void)m;
  return nullptr;  // DML stays on the engine's host path
}
void* ws_update_column(const void* old_col, const void* src, int64_t n, int w, int f64, const int32_t* rows, int64_t m,
                       ws_gpu_update* u) {
  (void)old_col, (void)src, (void)n, (void)w, (void)f64, (void)rows, (void)m;
  if (u) u->bad = 0;
  return nullptr;
}

int ws_copy_kernel(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) throw Error("copy_kernel: 16-byte multiples only", WS_ERR_INTERNAL);
    const int64_t n16 = (int64_t)(bytes / 16);
    if (!n16) return;
    ProfScope ps("copy_kernel");
    hipLaunchKernelGGL(copy_k, dim3(grid_for(n16)), dim3(256), 0, stream(), (const uint4*)src, (uint4*)dst, n16);
    check_launch("copy_kernel");
  });
}

// ------------------------------------------------------------------ C ABI: profiling

void ws_prof_enable(int on) {
  t_prof = on != 0;
  // events for the next 256 profiled launches made now, so that none is created inside a timed loop
  while (t_prof && t_prof_free.size() < 512) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {  // (no error out of a void entry point: profiling then creates as it goes)
      (void)hipGetLastError();
      break;
    }
    t_prof_free.push_back(e);
  }
}
void ws_prof_only(const char* kernel) { t_prof_only = kernel ? kernel : ""; }
void ws_prof_reset(void) {
  for (auto& r : t_prof_recs) {
    (void)hipEventSynchronize(r.b);
    t_prof_free.push_back(r.a);
    t_prof_free.push_back(r.b);
  }
  t_prof_recs.clear();
}
int ws_prof_collect(ws_prof_row* rows, int cap) {
  std::vector<ws_prof_row> agg;
  for (auto& r : t_prof_recs) {
    float ms = 0;
    if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    auto it = std::find_if(agg.begin(), agg.end(), [&](const ws_prof_row& x) { return strncmp(x.name, r.name, sizeof x.name - 1) == 0; });
    if (it == agg.end()) {
      ws_prof_row x{};
      snprintf(x.name, sizeof x.name, "%s", r.name);
      agg.push_back(x);
      it = agg.end() - 1;
    }
    it->launches += 1;
    it->total_ms += ms;
  }
  const int n = (int)std::min<size_t>(agg.size(), (size_t)std::max(cap, 0));
  for (int i = 0; i < n; ++i) rows[i] = agg[i];
  return (int)agg.size();
}

}  // extern "C"
