// Host-side internals shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "warpstore_amd.h"
#include "ws_kernels.h"

namespace ws {

// Errors cross the C ABI as status codes; inside the library they are exceptions carrying the code
// the engine's SqlError uses (1041 out of memory, 1105 internal, 1235 needs a compiled plan).
struct Error : std::runtime_error {
  int code;
  Error(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& msg);

[[noreturn]] void hip_fail(hipError_t e, const char* what, const char* file, int line);
#define WS_CK(x)                                            \
  do {                                                      \
    hipError_t e_ = (x);                                    \
    if (e_ != hipSuccess) ::ws::hip_fail(e_, #x, __FILE__, __LINE__); \
  } while (0)

// Runs f, converting exceptions into a status (and the thread's last-error message).
template <class F>
int guard(F&& f) {
  try {
    f();
    return WS_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host out of memory");
    return WS_ERR_INTERNAL;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return WS_ERR_INTERNAL;
  }
}

// ---- device
void ensure_init();
int cu_count();
// The current value of a tuning knob (ws_tune_set), by name; throws for an unknown name.
int64_t tune_knob(const char* name);
const std::string& arch_name();  // e.g. "gfx950:sramecc+:xnack-"
// The stream a thread's work goes to: its per-thread stream, or a capture stream while the thread
// records a graph (set through StreamOverride).
extern thread_local hipStream_t t_stream_override;
inline hipStream_t stream() { return t_stream_override ? t_stream_override : hipStreamPerThread; }
struct StreamOverride {
  hipStream_t saved;
  explicit StreamOverride(hipStream_t s) : saved(t_stream_override) { t_stream_override = s; }
  ~StreamOverride() { t_stream_override = saved; }
};

// Grid for a grid-stride kernel over n items.
inline unsigned grid_for(int64_t n, int block = 256, int per_cu = 8) {
  int64_t g = (n + block - 1) / block;
  const int64_t cap = (int64_t)cu_count() * per_cu;
  if (g > cap) g = cap;
  return (unsigned)(g < 1 ? 1 : g);
}

// ---- memory: caching allocator (tmp, reused in stream order per thread) and pinned staging
void* tmp_alloc(size_t bytes);
void tmp_free(void* p);
template <class T>
struct DevBuf {  // RAII tmp buffer
  T* p = nullptr;
  int64_t n = 0;
  DevBuf() = default;
  explicit DevBuf(int64_t count) : n(count) { p = count > 0 ? (T*)tmp_alloc(sizeof(T) * (size_t)count) : nullptr; }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) reset(), p = o.p, n = o.n, o.p = nullptr, o.n = 0;
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p) tmp_free(p);
    p = nullptr, n = 0;
  }
  T* release() {
    T* q = p;
    p = nullptr, n = 0;
    return q;
  }
};
// Thread-local ring of pinned buffers; a slot is reused once the copy out of it has run.
uint8_t* pinned_stage(size_t bytes);
void pinned_stage_done();  // records that copies issued from the stage are enqueued
void h2d_sync(void* dst, const void* src, size_t bytes);
void d2h_sync(void* dst, const void* src, size_t bytes);
// Several small device values (4-byte multiples, 64 KiB in all) read back in one round trip.
constexpr int kD2HMaxPieces = 8;
struct D2HPiece {
  void* dst;
  const void* src;
  size_t bytes;
};
void d2h_gather_sync(const D2HPiece* pieces, int n);
void h2d_async_staged(void* dst, const void* src, size_t bytes);  // via the pinned stage

// Byte blob assembled on the host and uploaded in one copy; offsets are resolved to device
// addresses after allocation.
struct Blob {
  std::vector<uint8_t> host;
  size_t put(const void* src, size_t bytes, size_t align = 16) {
    size_t off = (host.size() + align - 1) & ~(align - 1);
    host.resize(off + bytes);
    if (bytes && src) std::memcpy(host.data() + off, src, bytes);
    return off;
  }
  size_t reserve(size_t bytes, size_t align = 16) { return put(nullptr, bytes, align); }
  template <class T>
  T* at(size_t off) { return (T*)(host.data() + off); }
};
// Uploads the blob into a tmp buffer (returned; caller frees after the launches that read it).
uint8_t* upload_blob(const Blob& b);

// ---- kernel launch with optional profiling events
struct ProfScope {
  int slot = -1;
  explicit ProfScope(const char* kernel);
  ~ProfScope();
};
void check_launch(const char* what);
bool prof_enabled();

// ---- fill kernels (stream-ordered, capture-safe: no memset nodes)
void fill_u64(u64* p, int64_t n, u64 v);
void fill_u32(uint32_t* p, int64_t n, uint32_t v);
void fill_i32(int32_t* p, int64_t n, int32_t v);
void fill_u8(uint8_t* p, int64_t n, uint8_t v);

// ---- scan descriptors (ws_agg.hip)
struct ScanInput {  // one launch's batches, host side
  int nb = 0, ncols = 0;
  std::vector<int64_t> n;
  std::vector<const uint32_t*> del;
  std::vector<ws_colview> cols;  // nb * ncols
};

// Decoupled look-back compaction of slots [0, n) where pred holds (ws_compact.hip).
int64_t compact_occupied_hash(const u64* keys, int64_t n, int32_t* out_idx);
int64_t compact_dense(const double* acc, int naggs, int64_t G, int32_t* out_idx);
int64_t compact_flags(const uint8_t* flags, int64_t n, int32_t* out_idx);

// ---- JIT (ws_jit.cpp)
struct JitKernel;  // a compiled kernel ready to launch
// Returns the compiled function for kernel kind fn of program p, or nullptr.
void* jit_function(const ws_jit_program* p, int fn);  // fn: WS_JF_* or one of the kinds below
constexpr int kJitPartCount = 100, kJitPartScatter = 101;  // partition passes of a grouped plan
constexpr int kJitBatchDense = 102;                         // batched kernel of a dense plan
constexpr int kJitSelectChunk = 103;                        // chunked kernel of a selection plan
constexpr int kJitSelectStaged = 104;  // persistent chunked kernel of a selection plan (sets staged in LDS)
constexpr int kJitDenseStaged = 105;   // dense GROUP BY kernel of a dense plan with its sets staged in LDS
constexpr int kJitKeyBits = 106;       // key bitmap kernel of a selection plan (keys[0] of passing rows)
int jit_staged_static_lds(const ws_jit_program* p);  // static LDS bytes of either staged kernel
// ---- learned sizing (ws_sizing.hip): appends one (features, exact groups) sample to the log
void sizing_log(const float* features, int64_t groups);
void jit_launch_raw(void* fn, unsigned grid, unsigned block, size_t shmem, void** args, const char* prof_name);
// Profile name of a compiled plan's launch: base + "[" + the plan's kind + "]" (interned; base
// alone without a plan), so stages of one query that run the same kernel kind are told apart.
const char* jit_prof_name(const ws_jit_program* p, const char* base);
int jit_rows(const ws_jit_program* p);

}  // namespace ws
