// Kernel bodies of the fused scan pipeline, written for CDNA4 (gfx950): 64-lane wavefronts,
// 160 KiB of LDS per CU, HBM3E streaming.
//
// The bodies are templates over a row evaluator `Ev` (expressions of one query) and are built twice:
//   - ahead of time with `InterpEval`, which runs the engine's stack bytecode per row;
//   - at run time by hipRTC with a typed evaluator generated from the same bytecode (ws_jit.cpp).
// An evaluator supplies:
//   static constexpr int kKeys, kAggs, kRows;   // loop bounds, rows per thread per tile
//   static constexpr bool kPureKeys, kPureVals;  // keys / slot values have no side effect
//   static constexpr bool kPureWhere;             // neither has the WHERE (no overflow flag)
//   static constexpr bool kGlobalCols;            // bind() takes the batch's column views in global
//                                                 // memory (uniform loads: SGPR column bases) rather
//                                                 // than the block's LDS copy
//   Ev(const ScanDesc&);  void bind(const ws_colview* cols);   // per batch
//   int nkeys(), naggs(), op(a), agg_part(a), agg_null(a); bool has_where(), agg_int(a), agg_bits(a), key_int(k);
//   bool where(r); double key(k, r); i64 ikey(k, r); double agg(a, r); i64 iagg(a, r); uint8_t dkind(a);
// Semantics follow the reference kernels (src/kernels.cuh): NULL-skipping combines, exact int64 /
// int128 integer accumulation, the split BIGINT sum, and 64-bit hash group identity.
//
// Must stay compilable by hipRTC: no host headers.
#pragma once
#include "ws_vm.h"

namespace ws {

constexpr u64 kEmpty = ~0ull;
constexpr int kWave = 64;

// ------------------------------------------------------------------ launch descriptor

// One per launch, in device memory (uploaded together with the batch tables in one copy).
// Rows of all batches are cut into tiles of `tile_rows`; tiles never straddle batches, so a tile
// is (batch, first row, row count). Blocks walk tiles in increasing order.
struct ScanDesc {
  int32_t nb, ncols;
  int32_t tile_rows;
  int32_t mode;                  // dense: 0 global atomics, 1 per-thread private, 2 LDS copies
  i64 ntiles;
  const int64_t* n;              // [nb] rows per batch
  const uint32_t* const* del;    // [nb] deleted-row bitmaps (bit set = deleted), may be null
  const ws_colview* cols;        // [nb * ncols] column views per batch
  const int64_t* tile0;          // [nb + 1] first tile of each batch
  const int64_t* row0;           // [nb] first global row of each batch (select: output indices)
  const ws_expr_env* env;        // device copy of the query's expression environment
  const ws_agg_plan* plan;       // device copy of the aggregation plan
  const ws_dense_spec* ds;       // device copy of the dense spec (dense kernels)
  int32_t lds_slots;             // hash: LDS table slots (power of two); dense: copies
  int32_t copy_stride;           // dense LDS copies: bytes per copy
  int32_t slot_off[WS_MAX_AGGS]; // dense LDS copies: byte offset of each slot's array in a copy
  // Partitioned hash GROUP BY (the grouped plans' _pcount / _pscatter kernels).
  int32_t pbits;                 // log2 of the partition count (1..12); partition = top bits of the hash
  int32_t chunk_tiles;           // tiles per chunk; block c scans chunk c
  int32_t pre_slots;             // scatter: LDS pre-aggregation slots (power of two)
  int32_t rec_words;             // 1 + nkeys + naggs: words of one partial-group record
  uint32_t* part_cnt;            // [nchunks][P] surviving rows per chunk and partition
  const uint32_t* part_off;      // [nchunks][P] offset of the chunk's run inside its partition
  const uint32_t* part_start;    // [P + 1] first record of each partition
  u64* recs;                     // [rows][rec_words] records: hash, key values, slot values
  int* part_err;                 // set when a run would exceed its counted length
  int64_t* key_range;            // count pass: [chunk][2 * nkeys] min then max of the integer keys, or null
  const struct PackPlan* pack;   // scatter pass: the packed record layout chosen from them, or null
  int32_t set_words;             // persistent chunked selection: LDS words for staged set bitmaps
};

// Packed partial-group records: when every GROUP BY key is an integer (at most kPackKeys of them)
// and their ranges over the surviving rows fit 63 bits together, a record is [packed keys, slot
// values] instead of [hash, key values, slot values]: key k contributes (v - kmin[k]) << shift[k].
// The merge recomputes the hash from the unpacked keys, so groups are identified exactly as on the
// plain path. A packed word is below 2^63, so the free marker (all ones) stays distinct.
constexpr int kPackKeys = 4;
struct PackPlan {
  int32_t ok;  // 1: records are packed (1 + naggs words each)
  int32_t shift[kPackKeys];
  u64 mask[kPackKeys];  // (1 << width) - 1
  int64_t kmin[kPackKeys];
};

// Global hash table of one grouped aggregation (row-major accumulators per slot).
struct GTable {
  u64* keys;   // [cap] group hash, kEmpty = free
  double* kv;  // [cap * nkeys] key values (int keys as int64 bits)
  double* acc; // [cap * naggs] accumulators (agg_bits slots as int64 bits)
  u64 mask;    // cap - 1
  int* flags;  // [0] overflow, [1] dense key out of range
  u64* count;  // distinct groups inserted
};

// ------------------------------------------------------------------ small helpers

// Column access for compiled evaluators. A batch's column view is the same for every lane, so it
// is moved into scalar registers (readfirstlane per dword): column bases become SGPRs, loads use
// the global (not flat) instructions with a 32-bit lane offset, and the encoding switch is a
// scalar branch.
#define WS_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ ws_colview uniform_view(const ws_colview& v) {
  static_assert(sizeof(ws_colview) == 16, "column view layout");
  unsigned w[4];
  __builtin_memcpy(w, &v, 16);
#pragma unroll
  for (int i = 0; i < 4; ++i) w[i] = (unsigned)__builtin_amdgcn_readfirstlane((int)w[i]);
  ws_colview u;
  __builtin_memcpy(&u, w, 16);
  return u;
}
// Frame-of-reference integer column (encodings 1-3): codes from `data` (offset applied), value =
// base + code, the top code standing for the type's NULL.
// Builtin vector types for 16-byte loads through address-space-qualified pointers.
typedef int ws_i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned ws_u32x4 __attribute__((ext_vector_type(4)));
typedef long long ws_i64x2 __attribute__((ext_vector_type(2)));
typedef double ws_f64x2 __attribute__((ext_vector_type(2)));

struct EncView {
  const unsigned char WS_GLOBAL* data;
  i64 base, nul;
  int enc;
  int lg;        // log2 of the code width (0, 1, 2)
  unsigned top;  // the NULL code, also the code mask
};
__device__ __forceinline__ EncView enc_view(const ws_colview& v) {
  EncView e;
  e.enc = v.enc;
  e.nul = (v.t == WS_T_I32 || v.t == WS_T_I32N) ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
  e.base = v.enc ? *(const i64 WS_GLOBAL*)v.p : 0;
  e.data = (const unsigned char WS_GLOBAL*)v.p + WS_ENC_HEADER + (i64)v.off * enc_width(v.enc);
  e.lg = v.enc == 1 ? 0 : v.enc == 2 ? 1 : 2;
  e.top = v.enc == 1 ? 0xFFu : v.enc == 2 ? 0xFFFFu : 0xFFFFFFFFu;
  return e;
}
// One aligned 32-bit load for every code width, then a shift and mask: no branch on the encoding,
// so a row's evaluation stays one basic block. The word holding a code lies inside the buffer: the
// codes start 16 bytes in, and the segment format keeps 16 bytes after the last code.
// The 32-bit word holding row r's code, and the value decoded from it (enc_get in two halves).
__device__ __forceinline__ unsigned enc_word(const EncView& e, i64 r) {
  return *(const unsigned WS_GLOBAL*)((u64)(e.data + (r << e.lg)) & ~(u64)3);
}
__device__ __forceinline__ i64 enc_decode(const EncView& e, unsigned word, i64 r) {
  const unsigned code = (word >> ((unsigned)((u64)(e.data + (r << e.lg)) & 3) * 8)) & e.top;
  return code == e.top ? e.nul : e.base + (i64)code;
}
__device__ __forceinline__ i64 enc_get(const EncView& e, i64 r) {
  const u64 a = (u64)(e.data + (r << e.lg));
  const unsigned word = *(const unsigned WS_GLOBAL*)(a & ~(u64)3);
  const unsigned code = (word >> ((unsigned)(a & 3) * 8)) & e.top;
  return code == e.top ? e.nul : e.base + (i64)code;
}

__device__ __forceinline__ u64 mix64(u64 x) {  // splitmix64 finalizer
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}
__device__ __forceinline__ bool is_deleted(const uint32_t* del, i64 r) {
  return del && ((del[r >> 5] >> (r & 31)) & 1u);
}
__device__ __forceinline__ double agg_identity(int op) {
  return op == WS_AGG_MIN ? bitsd(0x7ff0000000000000ll) : op == WS_AGG_MAX ? bitsd((i64)0xfff0000000000000ull) : 0.0;
}
__device__ __forceinline__ i64 int_identity(int op) { return op == WS_AGG_MIN ? kI64Max : op == WS_AGG_MAX ? kI64Min : 0; }
__device__ __forceinline__ double slot_identity(bool bits, int op) { return bits ? bitsd(int_identity(op)) : agg_identity(op); }
// NULL-skipping combine of doubles (NaN = NULL): fmin / fmax ignore NaN, sums skip it.
__device__ __forceinline__ double combine_d(int op, double a, double b) {
  if (op == WS_AGG_MIN) return fmin(a, b);
  if (op == WS_AGG_MAX) return fmax(a, b);
  return b != b ? a : a != a ? b : a + b;
}
__device__ __forceinline__ i64 combine_i(int op, i64 a, i64 b) {
  return op == WS_AGG_MIN ? (b < a ? b : a) : op == WS_AGG_MAX ? (b > a ? b : a) : (i64)((u64)a + (u64)b);
}
// A slot value as stored: exact int64 bits (agg_bits) or a double.
__device__ __forceinline__ double combine_slot(bool bits, int op, double a, double b) {
  return bits ? bitsd(combine_i(op, dbits(a), dbits(b))) : combine_d(op, a, b);
}
// Integer aggregate argument as the slot accumulates it (reference kernels.cuh:65-69): a NULL of a
// nullable int column becomes the neutral value; a split BIGINT SUM keeps its low / high half.
// (The slot properties come from the evaluator: constants in a compiled plan, so the hot loops
// load nothing from the plan per row.)
template <class Ev>
__device__ __forceinline__ i64 int_part(const Ev& ev, int a, i64 v) {
  const int nul = ev.agg_null(a);
  if (nul && v == (nul == WS_T_I32N ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64)) return int_identity(ev.op(a));
  if (!ev.agg_bits(a)) return v;
  return ev.agg_part(a) == 1 ? (v & 0xffffffffll) : ev.agg_part(a) == 2 ? (v >> 32) : v;
}

// Atomic combines into a slot. SCOPE is __HIP_MEMORY_SCOPE_WORKGROUP for LDS tables and
// __HIP_MEMORY_SCOPE_AGENT for global ones. Double sums use the hardware FP64 atomic add (the
// buffers are coarse-grained device memory); double min / max are compare-and-swap loops; NaN
// inputs (NULL) are skipped.
template <int SCOPE>
__device__ __forceinline__ void atomic_combine_d(int op, double* addr, double v) {
  if (v != v) return;
  // adding zero to a global sum changes nothing (a sum slot starts at +0.0, so it never holds -0.0)
  if (SCOPE == __HIP_MEMORY_SCOPE_AGENT && op == WS_AGG_SUM && v == 0.0) return;
  if (op == WS_AGG_SUM || op == WS_AGG_COUNT) {
    __hip_atomic_fetch_add(addr, v, __ATOMIC_RELAXED, SCOPE);
    return;
  }
  u64* a = (u64*)addr;
  u64 old = __hip_atomic_load(a, __ATOMIC_RELAXED, SCOPE);
  for (;;) {
    const double cur = bitsd((i64)old);
    const double nv = op == WS_AGG_MIN ? fmin(cur, v) : fmax(cur, v);
    if (nv == cur) return;
    if (__hip_atomic_compare_exchange_strong(a, &old, (u64)dbits(nv), __ATOMIC_RELAXED, __ATOMIC_RELAXED, SCOPE)) return;
  }
}
template <int SCOPE>
__device__ __forceinline__ void atomic_combine_i(int op, i64* addr, i64 v) {
  if (SCOPE == __HIP_MEMORY_SCOPE_AGENT && op != WS_AGG_MIN && op != WS_AGG_MAX && v == 0) return;  // + 0: no change
  if (SCOPE == __HIP_MEMORY_SCOPE_AGENT && (op == WS_AGG_MIN || op == WS_AGG_MAX)) {
    // A global MIN only decreases (MAX only grows) while a kernel runs, and caches start clean at
    // every launch, so even a stale cached value that already beats v proves the update
    // unnecessary: most rows of a group then skip the read-modify-write.
    const i64 cur = __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (op == WS_AGG_MIN ? cur <= v : cur >= v) return;
  }
  if (op == WS_AGG_MIN) __hip_atomic_fetch_min(addr, v, __ATOMIC_RELAXED, SCOPE);
  else if (op == WS_AGG_MAX) __hip_atomic_fetch_max(addr, v, __ATOMIC_RELAXED, SCOPE);
  else __hip_atomic_fetch_add((u64*)addr, (u64)v, __ATOMIC_RELAXED, SCOPE);
}
template <int SCOPE>
__device__ __forceinline__ void atomic_combine_slot(bool bits, int op, double* addr, double v) {
  if (bits) atomic_combine_i<SCOPE>(op, (i64*)addr, dbits(v));
  else atomic_combine_d<SCOPE>(op, addr, v);
}
#define WS_LDS __HIP_MEMORY_SCOPE_WORKGROUP
#define WS_DEV __HIP_MEMORY_SCOPE_AGENT

// ------------------------------------------------------------------ wave / block reductions

__device__ __forceinline__ i64 wave_sum_i64(i64 v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ i64 wave_combine_i(int op, i64 v) {
  for (int o = 32; o > 0; o >>= 1) v = combine_i(op, v, (i64)__shfl_xor((long long)v, o, kWave));
  return v;
}
__device__ __forceinline__ double wave_combine_d(int op, double v) {
  for (int o = 32; o > 0; o >>= 1) v = combine_d(op, v, __shfl_xor(v, o, kWave));
  return v;
}
// 128-bit signed sum carried as (lo, hi).
__device__ __forceinline__ void add128(u64& lo, i64& hi, u64 blo, i64 bhi) {
  const u64 s = lo + blo;
  hi += bhi + (s < lo ? 1 : 0);
  lo = s;
}
__device__ __forceinline__ void wave_sum128(u64& lo, i64& hi) {
  for (int o = 32; o > 0; o >>= 1) {
    const u64 olo = __shfl_xor(lo, o, kWave);
    const i64 ohi = (i64)__shfl_xor((long long)hi, o, kWave);
    add128(lo, hi, olo, ohi);
  }
}

// ------------------------------------------------------------------ tile walking

// Per-block cursor over increasing tile ids: which batch a tile is in, and its rows.
struct TileRows {
  int b;
  i64 first, count;
};
// Tiles are block-uniform, so their batch, row range and delete bitmap are too; readfirstlane
// tells the compiler (scalar registers, scalar branches, and column loads of the batch from a
// scalar base) instead of it treating values loaded per lane as divergent.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ i64 uni(i64 x) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)x);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((u64)x >> 32));
  return (i64)(((u64)hi << 32) | lo);
}
template <class T>
__device__ __forceinline__ T* uni(T* p) { return (T*)uni((i64)p); }

// Set membership (compiled plans: the bitmap bases are scalar). Set bitmaps staged in LDS. A kernel that has LDS to spare copies the leading words of each set
// there (stage_set); a lookup reads words below `staged` from LDS and the rest from global memory.
// Both loads always issue (branch-free): a lane served by LDS reads global word 0, which all such
// lanes of a wave share (one cache line), and a lane served globally reads LDS word 0. Unstaged
// sets point at lds_spare_word() with staged = 0.
#define WS_LOCAL __attribute__((address_space(3)))
__device__ __forceinline__ const unsigned WS_LOCAL* lds_spare_word() {
  __shared__ unsigned spare;  // (never written: its value is always masked off)
  return (const unsigned WS_LOCAL*)&spare;
}
// Copies min(set words, words - off) words of a set to lds[off...]; returns the count. Each
// thread issues all loads of a round before its first LDS store, so a round costs one memory
// round trip: 16-byte loads when source and destination allow, else 4-byte ones.
__device__ __forceinline__ int stage_set(unsigned* lds, int off, int words, const unsigned WS_GLOBAL* g, i64 size) {
  const i64 need = size > 0 ? (size + 31) >> 5 : 0;
  const i64 room = words - off > 0 ? words - off : 0;
  const int w = (int)(need < room ? need : room);
  constexpr int U = 8;
  const int T = blockDim.x;
  int done = 0;
  if ((((unsigned long long)g | (unsigned long long)(lds + off)) & 15) == 0) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const int w4 = w >> 2;
    for (int base = 0; base < w4; base += U * T) {
      u4 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * T + (int)threadIdx.x;
        if (i < w4) v[u] = ((const u4 WS_GLOBAL*)g)[i];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * T + (int)threadIdx.x;
        if (i < w4) ((u4*)(lds + off))[i] = v[u];
      }
    }
    done = w4 << 2;
  }
  for (int base = done; base < w; base += U * T) {
    unsigned v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * T + (int)threadIdx.x;
      if (i < w) v[u] = g[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * T + (int)threadIdx.x;
      if (i < w) lds[off + i] = v[u];
    }
  }
  return w;
}
// size is the set's bit count (0 <= size < 2^31), so a member's index fits 32 bits.
__device__ __forceinline__ i64 inset_s(const unsigned WS_LOCAL* l, int staged, const unsigned WS_GLOBAL* g, i64 size, i64 c) {
  const bool in = (u64)c < (u64)size;  // also c >= 0
  const unsigned cc = in ? (unsigned)c : 0u;
  const unsigned w = cc >> 5;
  const bool loc = w < (unsigned)staged;
  const unsigned a = l[loc ? w : 0u];
  const unsigned b = g[loc ? 0u : w];
  return (i64)(((loc ? a : b) >> (cc & 31)) & (unsigned)in);
}
__device__ __forceinline__ double inset_sd(const unsigned WS_LOCAL* l, int staged, const unsigned WS_GLOBAL* g, i64 size,
                                           double x) {
  if (x != x) return x;
  return (double)inset_s(l, staged, g, size, (i64)x);
}

__device__ __forceinline__ int batch_of_tile(const ScanDesc& d, i64 t, int hint) {
  int b = hint < 0 ? 0 : hint;
  while (b + 1 < d.nb && d.tile0[b + 1] <= t) ++b;
  return uni(b);
}
__device__ __forceinline__ int batch_of_tile_bsearch(const ScanDesc& d, i64 t) {
  int lo = 0, hi = d.nb - 1;  // last batch with tile0 <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.tile0[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return uni(lo);
}
__device__ __forceinline__ TileRows tile_rows(const ScanDesc& d, i64 t, int b) {
  TileRows tr;
  tr.b = b;
  tr.first = uni((t - d.tile0[b]) * (i64)d.tile_rows);
  const i64 left = d.n[b] - tr.first;
  tr.count = uni(left < d.tile_rows ? left : (i64)d.tile_rows);
  return tr;
}
// The batch of a block's current tile, with the fields its tile loop needs held in scalar
// registers. The descriptor is read only when the batch changes: a read of it inside a loop that
// stores to global memory is a vector load (the compiler cannot prove the descriptor unchanged,
// which a scalar load needs), and waiting for one waits for every load issued before it, a
// prefetch of the next tile's columns included.
struct BatchCursor {
  int b = -1;
  i64 t_lo = 0, t_hi = 0, n = 0, tr = 0, row0 = 0;
  const uint32_t* del = nullptr;
  // Moves to tile t's batch; true when it differs from the previous one (block-uniform).
  __device__ __forceinline__ bool seek(const ScanDesc& d, i64 t) {
    if (b >= 0 && t >= t_lo && t < t_hi) return false;
    b = b < 0 || t < t_lo ? batch_of_tile_bsearch(d, t) : batch_of_tile(d, t, b);
    t_lo = uni((i64)d.tile0[b]);
    t_hi = uni((i64)d.tile0[b + 1]);
    n = uni((i64)d.n[b]);
    tr = uni((i64)d.tile_rows);
    row0 = d.row0 ? uni((i64)d.row0[b]) : 0;
    del = uni(d.del[b]);
    return true;
  }
  __device__ __forceinline__ TileRows rows(i64 t) const {
    TileRows r;
    r.b = b;
    r.first = (t - t_lo) * tr;
    const i64 left = n - r.first;
    r.count = left < tr ? left : tr;
    return r;
  }
};
// Block-wide: stage batch b's column views in LDS (all threads call).
__device__ __forceinline__ void stage_batch(const ScanDesc& d, int b, ws_colview* scols) {
  __syncthreads();
  for (int i = threadIdx.x; i < d.ncols; i += blockDim.x) scols[i] = d.cols[(i64)b * d.ncols + i];
  __syncthreads();
}

// ------------------------------------------------------------------ interpreter evaluator

// Out-of-line VM call: inlined into every unrolled key / aggregate evaluation, its stack and the
// opcode switch would multiply register pressure (and spill) across the whole kernel.
template <bool kFull>
__device__ __noinline__ double eval_prog_ni(const ws_expr_env* env, const ws_colview* cols, ws_prog pg, i64 row) {
  return eval_prog_t<kFull>(*env, cols, pg, row);
}

// Runs the engine's bytecode per row (AOT fallback while a plan compiles, or when codegen is off).
// Plan limits: WS_INTERP_KEYS keys, WS_INTERP_AGGS aggregates. A program that is one column or one
// literal (the common case for keys and aggregate arguments) is read inline. kFull = false is the
// lean variant for plans without scalar-function opcodes (see prog_is_simple).
template <bool kFull>
struct InterpEvalT {
  static constexpr int kKeys = WS_INTERP_KEYS, kAggs = WS_INTERP_AGGS, kRows = 1;
  static constexpr bool kPureKeys = false, kPureVals = false;  // (see the compiled evaluator)
  static constexpr bool kPureWhere = false;
  static constexpr bool kGlobalCols = false;
  const ws_expr_env* env;
  const ws_agg_plan* plan;
  const ws_colview* cols = nullptr;
  __device__ InterpEvalT(const ScanDesc& d) : env(d.env), plan(d.plan) {}
  __device__ void bind(const ws_colview* c) { cols = c; }
  __device__ int nkeys() const { return plan->nkeys; }
  __device__ int naggs() const { return plan->naggs; }
  __device__ int op(int a) const { return plan->ops[a]; }
  __device__ bool agg_int(int a) const { return plan->agg_int[a] != 0; }
  __device__ bool agg_bits(int a) const { return plan->agg_bits[a] != 0; }
  __device__ int agg_part(int a) const { return plan->agg_part[a]; }
  __device__ int agg_null(int a) const { return plan->agg_null[a]; }
  __device__ bool key_int(int k) const { return plan->key_int[k] != 0; }
  __device__ bool has_where() const { return plan->has_where != 0; }
  __device__ uint8_t dkind(int a, const ws_dense_spec& ds) const { return ds.kind[a]; }
  __device__ double run(ws_prog p, i64 r) const {
    if (p.len == 1) {
      const ws_instr in = env->code[p.start];
      if (in.op == WS_OP_COL) return load_col(cols[in.arg], r);
      if (in.op == WS_OP_CONST) return in.is_int ? (double)dbits(env->consts[in.arg]) : env->consts[in.arg];
    }
    return eval_prog_ni<kFull>(env, cols, p, r);
  }
  __device__ bool where(i64 r) const { return !plan->has_where || truthy(run(plan->where, r)); }
  __device__ double key(int k, i64 r) const { return run(plan->keys[k], r); }
  // A plain integer column reads exactly (BIGINT above 2^53); other integer programs round.
  __device__ i64 ikey(int k, i64 r) const {
    const ws_prog p = plan->keys[k];
    const ws_instr in = env->code[p.start];
    if (p.len == 1 && in.op == WS_OP_COL && cols[in.arg].t != WS_T_F64) return col_int(cols[in.arg], r);
    return (i64)key(k, r);
  }
  __device__ double agg(int a, i64 r) const { return plan->ops[a] == WS_AGG_COUNT ? 1.0 : run(plan->aggs[a], r); }
  __device__ i64 iagg(int a, i64 r) const {
    const ws_prog p = plan->aggs[a];
    const ws_instr in = env->code[p.start];
    if (plan->ops[a] != WS_AGG_COUNT && p.len == 1 && in.op == WS_OP_COL && cols[in.arg].t != WS_T_F64)
      return col_int(cols[in.arg], r);
    return (i64)agg(a, r);
  }
};
typedef InterpEvalT<true> InterpEval;
typedef InterpEvalT<false> InterpEvalLean;

// Interpreter for plans whose every program (WHERE, keys, aggregate arguments) is a single column
// or literal: no VM call at all, so rows can be unrolled and their loads issued together.
struct InterpEvalCols {
  static constexpr int kKeys = WS_INTERP_KEYS, kAggs = WS_INTERP_AGGS, kRows = 4;
  static constexpr bool kPureKeys = true, kPureVals = true, kPureWhere = true;
  static constexpr bool kGlobalCols = false;
  const ws_expr_env* env;
  const ws_agg_plan* plan;
  const ws_colview* cols = nullptr;
  __device__ InterpEvalCols(const ScanDesc& d) : env(d.env), plan(d.plan) {}
  __device__ void bind(const ws_colview* c) { cols = c; }
  __device__ int nkeys() const { return plan->nkeys; }
  __device__ int naggs() const { return plan->naggs; }
  __device__ int op(int a) const { return plan->ops[a]; }
  __device__ bool agg_int(int a) const { return plan->agg_int[a] != 0; }
  __device__ bool agg_bits(int a) const { return plan->agg_bits[a] != 0; }
  __device__ int agg_part(int a) const { return plan->agg_part[a]; }
  __device__ int agg_null(int a) const { return plan->agg_null[a]; }
  __device__ bool key_int(int k) const { return plan->key_int[k] != 0; }
  __device__ bool has_where() const { return plan->has_where != 0; }
  __device__ uint8_t dkind(int a, const ws_dense_spec& ds) const { return ds.kind[a]; }
  __device__ double run(ws_prog p, i64 r) const {
    const ws_instr in = env->code[p.start];
    if (in.op == WS_OP_COL) return load_col(cols[in.arg], r);
    return in.is_int ? (double)dbits(env->consts[in.arg]) : env->consts[in.arg];
  }
  __device__ i64 irun(ws_prog p, i64 r) const {
    const ws_instr in = env->code[p.start];
    if (in.op == WS_OP_COL && cols[in.arg].t != WS_T_F64) return col_int(cols[in.arg], r);
    return (i64)run(p, r);
  }
  __device__ bool where(i64 r) const { return !plan->has_where || truthy(run(plan->where, r)); }
  __device__ double key(int k, i64 r) const { return run(plan->keys[k], r); }
  __device__ i64 ikey(int k, i64 r) const { return irun(plan->keys[k], r); }
  __device__ double agg(int a, i64 r) const { return plan->ops[a] == WS_AGG_COUNT ? 1.0 : run(plan->aggs[a], r); }
  __device__ i64 iagg(int a, i64 r) const { return plan->ops[a] == WS_AGG_COUNT ? 1 : irun(plan->aggs[a], r); }
};

// ------------------------------------------------------------------ no GROUP BY

// Per-block partial of a no-keys aggregation: per slot a 128-bit (lo, hi) pair for integer sums,
// the int64 min / max, or the double value; slot 0 counts rows.
struct NoKeysPartial {
  u64 lo[WS_MAX_AGGS];
  i64 hi[WS_MAX_AGGS];
};

// Filter + aggregate without GROUP BY. Each thread accumulates exactly (int128 sums, int64
// min / max, NULL-skipping doubles), the block reduces through wave shuffles and LDS, writes one
// partial, and the last block to finish folds all partials in block order into `out` (so FP sums
// are deterministic for a given input and grid). out[a] holds the low 64 bits / value, out[64 + a]
// the high 64 bits of exact integer sums (reference kernels.cuh:188-302 keeps the same layout).
template <class Ev>
__device__ void nokeys_body(const ScanDesc& d, double* out, NoKeysPartial* parts, unsigned* done) {
  __shared__ ws_colview scols[WS_MAX_COLS];
  __shared__ u64 s_lo[WS_MAX_AGGS][4];
  __shared__ i64 s_hi[WS_MAX_AGGS][4];
  __shared__ bool s_last;
  Ev ev(d);
  const ws_agg_plan& plan = *d.plan;
  const int NA = ev.naggs();
  u64 lo[Ev::kAggs];
  i64 hi[Ev::kAggs];
  double facc[Ev::kAggs];
#pragma unroll
  for (int a = 0; a < Ev::kAggs; ++a) {
    const int op = a < NA ? ev.op(a) : WS_AGG_SUM;
    lo[a] = (u64)int_identity(op);
    hi[a] = 0;
    facc[a] = agg_identity(op);
  }
  i64 cnt = 0;
  BatchCursor bc;
  const i64 ntiles = uni(d.ntiles);
  for (i64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    if (bc.seek(d, t)) {
      stage_batch(d, bc.b, scols);
      ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
    }
    const TileRows tr = bc.rows(t);
    const uint32_t* del = bc.del;
#pragma unroll
    for (int j = 0; j < Ev::kRows; ++j) {
      for (i64 i = threadIdx.x + (i64)j * blockDim.x; i < tr.count; i += (i64)Ev::kRows * blockDim.x) {
        const i64 r = tr.first + i;
        if (is_deleted(del, r)) continue;
        if (ev.has_where() && !ev.where(r)) continue;
        ++cnt;
#pragma unroll
        for (int a = 0; a < Ev::kAggs; ++a) {
          if (a >= NA) break;
          const int op = ev.op(a);
          if (op == WS_AGG_COUNT) continue;
          if (ev.agg_int(a)) {
            const i64 v = int_part(ev, a, ev.iagg(a, r));
            if (op == WS_AGG_SUM) add128(lo[a], hi[a], (u64)v, v < 0 ? -1 : 0);
            else lo[a] = (u64)combine_i(op, (i64)lo[a], v);
          } else {
            facc[a] = combine_d(op, facc[a], ev.agg(a, r));
          }
        }
      }
    }
  }
  // Block reduction: wave shuffles, then LDS across the block's waves (<= 4 waves of 64).
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  for (int a = 0; a < NA && a < Ev::kAggs; ++a) {
    const int op = ev.op(a);
    u64 vlo;
    i64 vhi = 0;
    if (op == WS_AGG_COUNT) {
      vlo = (u64)wave_sum_i64(cnt);
    } else if (ev.agg_int(a) && op == WS_AGG_SUM) {
      vlo = lo[a];
      vhi = hi[a];
      wave_sum128(vlo, vhi);
    } else if (ev.agg_int(a)) {
      vlo = (u64)wave_combine_i(op, (i64)lo[a]);
    } else {
      vlo = (u64)dbits(wave_combine_d(op, facc[a]));
    }
    if (lane == 0) {
      s_lo[a][wave] = vlo;
      s_hi[a][wave] = vhi;
    }
  }
  __syncthreads();
  if (threadIdx.x < NA) {
    const int a = threadIdx.x;
    const int op = plan.ops[a];
    u64 vlo = s_lo[a][0];
    i64 vhi = s_hi[a][0];
    for (int w = 1; w < nw; ++w) {
      if (op == WS_AGG_COUNT || (plan.agg_int[a] && op == WS_AGG_SUM)) add128(vlo, vhi, s_lo[a][w], s_hi[a][w]);
      else if (plan.agg_int[a]) vlo = (u64)combine_i(op, (i64)vlo, (i64)s_lo[a][w]);
      else vlo = (u64)dbits(combine_d(op, bitsd((i64)vlo), bitsd((i64)s_lo[a][w])));
    }
    parts[blockIdx.x].lo[a] = vlo;
    parts[blockIdx.x].hi[a] = vhi;
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  // Last block: fold the partials in block order; thread a owns slot a.
  if (threadIdx.x < NA) {
    const int a = threadIdx.x;
    const int op = plan.ops[a];
    const bool isum = op == WS_AGG_COUNT || (plan.agg_int[a] && op == WS_AGG_SUM);
    u64 vlo = isum ? 0 : (u64)(plan.agg_int[a] ? int_identity(op) : dbits(agg_identity(op)));
    i64 vhi = 0;
    for (unsigned blk = 0; blk < gridDim.x; ++blk) {
      const volatile NoKeysPartial* p = parts + blk;
      const u64 plo = p->lo[a];
      const i64 phi = p->hi[a];
      if (isum) add128(vlo, vhi, plo, phi);
      else if (plan.agg_int[a]) vlo = (u64)combine_i(op, (i64)vlo, (i64)plo);
      else vlo = (u64)dbits(combine_d(op, bitsd((i64)vlo), bitsd((i64)plo)));
    }
    // Fold into the running result (earlier windows of the same aggregation).
    if (op == WS_AGG_COUNT) {
      out[a] = out[a] + (double)(i64)vlo;
    } else if (plan.agg_int[a] && op == WS_AGG_SUM) {
      u64 olo = (u64)dbits(out[a]);
      i64 ohi = dbits(out[WS_MAX_AGGS + a]);
      add128(olo, ohi, vlo, vhi);
      out[a] = bitsd((i64)olo);
      out[WS_MAX_AGGS + a] = bitsd(ohi);
    } else if (plan.agg_int[a]) {
      out[a] = bitsd(combine_i(op, dbits(out[a]), (i64)vlo));
    } else {
      out[a] = combine_d(op, out[a], bitsd((i64)vlo));
    }
  }
}

// ------------------------------------------------------------------ dense GROUP BY

// Row's dense group index g = sum((key_k - kmin_k) * stride_k); -1 if a key is outside its range
// (stale statistics: the caller is told through flags[1] instead of writing out of bounds).
template <class Ev>
__device__ __forceinline__ i64 dense_index(const Ev& ev, const ws_dense_spec& ds, int NK, i64 r) {
  i64 g = 0;
#pragma unroll
  for (int k = 0; k < Ev::kKeys; ++k) {
    if (k >= NK) break;
    const i64 o = ev.ikey(k, r) - ds.kmin[k];
    if (o < 0 || o >= ds.range[k]) return -1;
    g += o * ds.stride[k];
  }
  return g;
}
__device__ __forceinline__ int dk_op(uint8_t k) {
  return (k == WS_DK_MIN_I || k == WS_DK_MIN_D) ? WS_AGG_MIN : (k == WS_DK_MAX_I || k == WS_DK_MAX_D) ? WS_AGG_MAX : WS_AGG_SUM;
}
__device__ __forceinline__ bool dk_int(uint8_t k) {
  return k == WS_DK_COUNT || k == WS_DK_SUM_I || k == WS_DK_MIN_I || k == WS_DK_MAX_I;
}
// Per-row value of slot a as the dense accumulators hold it: int64 for integer kinds, double
// otherwise (NaN = NULL, skipped by the combine).
template <class Ev>
__device__ __forceinline__ void dense_value(const Ev& ev, const ws_agg_plan& plan, uint8_t k, int a, i64 r, i64& iv, double& dv) {
  if (k == WS_DK_COUNT) iv = 1;
  else if (dk_int(k)) iv = int_part(ev, a, ev.iagg(a, r));
  else dv = ev.agg(a, r);
}
// Block partial of one dense slot -> the value written to the partials (or combined globally):
// integer kinds as int64 bits, double kinds as double bits.
// Final slot encoding (reference kernels.cuh:489-506, 616-630): agg_bits -> int64 bits; otherwise a
// double ((double) of the exact integer).

// The dense LDS copies (layout mode 2): d.lds_slots copies of d.copy_stride bytes, slot a's array
// at d.slot_off[a] within a copy (u32 counts, int64 or double otherwise). Block-wide.
__device__ void dense_init_copies(const ScanDesc& d, const ws_dense_spec& ds, int NA, unsigned char* smem) {
  const i64 G = ds.G;
  for (int c = 0; c < d.lds_slots; ++c) {
    unsigned char* cb = smem + (i64)c * d.copy_stride;
    for (int a = 0; a < NA; ++a) {
      const uint8_t k = ds.kind[a];
      unsigned char* arr = cb + d.slot_off[a];
      for (i64 g = threadIdx.x; g < G; g += blockDim.x) {
        if (k == WS_DK_COUNT) ((unsigned*)arr)[g] = 0u;
        else if (dk_int(k)) ((i64*)arr)[g] = int_identity(dk_op(k));
        else ((double*)arr)[g] = agg_identity(dk_op(k));
      }
    }
  }
}
// The block's partial of every (group, slot) from its LDS copies, written coalesced to
// out[g * NA + a] (int64 or double bits). Block-wide, after a barrier.
__device__ void dense_write_copies(const ScanDesc& d, const ws_dense_spec& ds, int NA, const unsigned char* smem, i64* out) {
  const i64 S = ds.G * NA;
  const int C = d.lds_slots;
  for (i64 s = threadIdx.x; s < S; s += blockDim.x) {
    const i64 g = s / NA;
    const int a = (int)(s - g * NA);
    const uint8_t k = ds.kind[a];
    const int op = dk_op(k);
    if (k == WS_DK_COUNT) {
      i64 x = 0;
      for (int c = 0; c < C; ++c) x += ((const unsigned*)(smem + (i64)c * d.copy_stride + d.slot_off[a]))[g];
      out[s] = x;
    } else if (dk_int(k)) {
      i64 x = int_identity(op);
      for (int c = 0; c < C; ++c) x = combine_i(op, x, ((const i64*)(smem + (i64)c * d.copy_stride + d.slot_off[a]))[g]);
      out[s] = x;
    } else {
      double x = agg_identity(op);
      for (int c = 0; c < C; ++c) x = combine_d(op, x, ((const double*)(smem + (i64)c * d.copy_stride + d.slot_off[a]))[g]);
      out[s] = dbits(x);
    }
  }
}

// Dense filter + GROUP BY. Three physical layouts, chosen by the host planner per launch:
//   mode 1: per-thread private accumulators in LDS ([slot][thread], no atomics per row);
//   mode 2: `copies` LDS copies of all group accumulators, copy = wave % copies, LDS atomics;
//   mode 0: straight to global memory with atomics.
// Modes 1 and 2 end with one block partial per slot written to `parts` ([block][G * NA], int64
// or double bits), folded by `dense_reduce` in block order.
template <class Ev>
__device__ void dense_body(const ScanDesc& d, double* gacc, i64* parts, int* flags) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  Ev ev(d);
  const ws_agg_plan& plan = *d.plan;
  const ws_dense_spec& ds = *d.ds;
  const int NK = ev.nkeys(), NA = ev.naggs();
  const i64 G = ds.G;
  const int T = blockDim.x, tid = threadIdx.x;
  const int mode = d.mode;
  const int C = d.lds_slots;
  // ---- init the LDS accumulators
  if (mode == 1) {
    i64* sl = (i64*)ws_smem;
    const int S = (int)G * NA;
    for (int s = 0; s < S; ++s) {
      const uint8_t k = ds.kind[s % NA];
      const i64 init = dk_int(k) ? int_identity(dk_op(k)) : dbits(agg_identity(dk_op(k)));
      sl[(i64)s * T + tid] = init;
    }
  } else if (mode == 2) {
    dense_init_copies(d, ds, NA, ws_smem);
  }
  const int wave = tid / kWave;
  unsigned char* mine = mode == 2 ? ws_smem + (i64)(wave % C) * d.copy_stride : nullptr;
  bool bad = false;
  BatchCursor bc;
  const i64 ntiles = uni(d.ntiles);
  for (i64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    if (bc.seek(d, t)) {
      stage_batch(d, bc.b, scols);
      ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
    }
    const TileRows tr = bc.rows(t);
    const uint32_t* del = bc.del;
    for (i64 i = tid; i < tr.count; i += T) {
      const i64 r = tr.first + i;
      if (is_deleted(del, r)) continue;
      if (ev.has_where() && !ev.where(r)) continue;
      const i64 g = dense_index(ev, ds, NK, r);
      if (g < 0) {
        bad = true;
        continue;
      }
      if (mode == 1) {
        i64* sl = (i64*)ws_smem;
        const i64 base = g * NA * T + tid;
#pragma unroll
        for (int a = 0; a < Ev::kAggs; ++a) {
          if (a >= NA) break;
          const uint8_t k = ev.dkind(a, ds);
          i64 iv = 0;
          double dv = 0;
          dense_value(ev, plan, k, a, r, iv, dv);
          i64& slot = sl[base + (i64)a * T];
          if (dk_int(k)) slot = combine_i(dk_op(k), slot, iv);
          else slot = dbits(combine_d(dk_op(k), bitsd(slot), dv));
        }
      } else if (mode == 2) {
#pragma unroll
        for (int a = 0; a < Ev::kAggs; ++a) {
          if (a >= NA) break;
          const uint8_t k = ev.dkind(a, ds);
          unsigned char* arr = mine + d.slot_off[a];
          i64 iv = 0;
          double dv = 0;
          dense_value(ev, plan, k, a, r, iv, dv);
          if (k == WS_DK_COUNT) __hip_atomic_fetch_add(&((unsigned*)arr)[g], 1u, __ATOMIC_RELAXED, WS_LDS);
          else if (dk_int(k)) atomic_combine_i<WS_LDS>(dk_op(k), &((i64*)arr)[g], iv);
          else atomic_combine_d<WS_LDS>(dk_op(k), &((double*)arr)[g], dv);
        }
      } else {
        double* row = gacc + g * NA;
#pragma unroll
        for (int a = 0; a < Ev::kAggs; ++a) {
          if (a >= NA) break;
          const uint8_t k = ev.dkind(a, ds);
          i64 iv = 0;
          double dv = 0;
          dense_value(ev, plan, k, a, r, iv, dv);
          if (ev.agg_bits(a)) atomic_combine_i<WS_DEV>(dk_op(k), (i64*)&row[a], iv);
          else if (dk_int(k)) atomic_combine_d<WS_DEV>(a == 0 ? WS_AGG_SUM : dk_op(k), &row[a], (double)iv);
          else atomic_combine_d<WS_DEV>(dk_op(k), &row[a], dv);
        }
      }
    }
  }
  if (bad) flags[1] = 1;
  if (mode == 0) return;
  __syncthreads();
  // ---- block partial: one value per (group, slot), written coalesced to parts[block][g * NA + a]
  const i64 S = G * NA;
  i64* out = parts + (i64)blockIdx.x * S;
  if (mode == 1) {
    const i64* sl = (const i64*)ws_smem;
    // One wave per slot: lanes sum / combine the T private copies.
    const int lane = tid & (kWave - 1), nw = T / kWave;
    for (i64 s = wave; s < S; s += nw) {
      const uint8_t k = ds.kind[s % NA];
      const int op = dk_op(k);
      if (dk_int(k)) {
        i64 x = int_identity(op);
        for (int t = lane; t < T; t += kWave) x = combine_i(op, x, sl[s * T + t]);
        x = wave_combine_i(op, x);
        if (lane == 0) out[s] = x;
      } else {
        double x = agg_identity(op);
        for (int t = lane; t < T; t += kWave) x = combine_d(op, x, bitsd(sl[s * T + t]));
        x = wave_combine_d(op, x);
        if (lane == 0) out[s] = dbits(x);
      }
    }
  } else {
    dense_write_copies(d, ds, NA, ws_smem, out);
  }
}

// Folds the dense block partials in block order into gacc (row-major [g][NA]), on top of what
// earlier windows left there. A block whose count for g is 0 contributes nothing to g. Lane l of a
// 1024-thread block owns slot blockIdx.x * 64 + l, wave w folds partials w, w + 16, ... in order
// (adjacent lanes read adjacent slots of a partial: coalesced), then wave 0 folds the 16 wave
// results in wave order (a fixed order: deterministic for doubles).
constexpr int kReduceWaves = 16;
__device__ void dense_reduce_cols_body(const ws_agg_plan& plan, const ws_dense_spec& ds, const i64* parts, int nparts,
                                       double* gacc) {
  __shared__ i64 s_v[kReduceWaves][kWave];
  __shared__ unsigned char s_any[kReduceWaves][kWave];
  const int NA = plan.naggs;
  const i64 S = ds.G * NA;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const i64 s = (i64)blockIdx.x * kWave + lane;
  const bool live = s < S;
  const i64 g = live ? s / NA : 0;
  const int a = live ? (int)(s - g * NA) : 0;
  const uint8_t k = ds.kind[a];
  const int op = dk_op(k);
  const bool isint = dk_int(k);
  i64 xi = op == WS_AGG_SUM ? 0 : int_identity(op);
  double xd = agg_identity(op);
  bool any = false;
  if (live) {
    // 8 partials per step, their count and value loads issued before any is used
    constexpr int U = 8;
    for (int p0 = w; p0 < nparts; p0 += kReduceWaves * U) {
      i64 c[U], v[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const int p = p0 + j * kReduceWaves;
        const i64* pp = parts + (i64)(p < nparts ? p : w) * S;
        c[j] = p < nparts ? pp[g * NA] : 0;
        v[j] = pp[s];
      }
#pragma unroll
      for (int j = 0; j < U; ++j) {
        if (c[j] == 0) continue;  // the block saw no row of g (or past the last partial)
        any = true;
        if (isint) xi = combine_i(op, xi, v[j]);
        else xd = combine_d(op, xd, bitsd(v[j]));
      }
    }
  }
  s_v[w][lane] = isint ? xi : dbits(xd);
  s_any[w][lane] = any ? 1 : 0;
  __syncthreads();
  if (w != 0 || !live) return;
  bool seen = false;
  for (int v = 0; v < kReduceWaves; ++v) {
    if (!s_any[v][lane]) continue;
    const i64 o = s_v[v][lane];
    if (!seen) xi = o, xd = bitsd(o);
    else if (isint) xi = combine_i(op, xi, o);
    else xd = combine_d(op, xd, bitsd(o));
    seen = true;
  }
  if (!seen) return;
  if (isint) {
    if (plan.agg_bits[a]) gacc[s] = bitsd(combine_i(op, dbits(gacc[s]), xi));
    else gacc[s] = combine_d(a == 0 ? WS_AGG_SUM : op, gacc[s], (double)xi);
  } else {
    gacc[s] = combine_d(op, gacc[s], xd);
  }
}


// The same fold with one block per slot (few slots, many partials): the block's threads combine
// strided partials, then a fixed-order tree in LDS (deterministic for doubles) finishes the slot.
__device__ void dense_reduce_slot_body(const ws_agg_plan& plan, const ws_dense_spec& ds, const i64* parts, int nparts,
                                       double* gacc) {
  __shared__ i64 s_v[256];
  __shared__ int s_any[256];
  const int NA = plan.naggs;
  const i64 S = ds.G * NA;
  const i64 s = blockIdx.x;
  const i64 g = s / NA;
  const int a = (int)(s - g * NA);
  const uint8_t k = ds.kind[a];
  const int op = dk_op(k);
  const bool isint = dk_int(k);
  i64 xi = op == WS_AGG_SUM ? 0 : int_identity(op);
  double xd = agg_identity(op);
  bool any = false;
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
    if (parts[(i64)p * S + g * NA] == 0) continue;  // the block saw no row of g
    any = true;
    if (isint) xi = combine_i(op, xi, parts[(i64)p * S + s]);
    else xd = combine_d(op, xd, bitsd(parts[(i64)p * S + s]));
  }
  s_v[threadIdx.x] = isint ? xi : dbits(xd);
  s_any[threadIdx.x] = any ? 1 : 0;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && s_any[threadIdx.x + w]) {
      const i64 o = s_v[threadIdx.x + w];
      s_v[threadIdx.x] = !s_any[threadIdx.x] ? o
                         : isint         ? combine_i(op, s_v[threadIdx.x], o)
                                         : dbits(combine_d(op, bitsd(s_v[threadIdx.x]), bitsd(o)));
      s_any[threadIdx.x] = 1;
    }
    __syncthreads();
  }
  if (threadIdx.x != 0 || !s_any[0]) return;
  const bool bits = plan.agg_bits[a] != 0;
  if (isint) {
    const i64 x = s_v[0];
    if (bits) gacc[s] = bitsd(combine_i(op, dbits(gacc[s]), x));
    else gacc[s] = combine_d(a == 0 ? WS_AGG_SUM : op, gacc[s], (double)x);
  } else {
    gacc[s] = combine_d(op, gacc[s], bitsd(s_v[0]));
  }
}

// ------------------------------------------------------------------ hash GROUP BY

// Inserts one group contribution into the global table; false when the table overflowed (too
// full or a probe chain longer than 256), which the caller answers by retrying bigger.
__device__ __forceinline__ bool global_insert(const GTable& t, const ws_agg_plan& plan, int NK, int NA, u64 h,
                                              const double* kv, const double* vals, unsigned* new_groups) {
  u64 s = h & t.mask;
  for (int probe = 0; probe < 256; ++probe) {
    u64 prev = kEmpty;
    const bool claimed = __hip_atomic_compare_exchange_strong(&t.keys[s], &prev, h, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
    if (claimed) {
      for (int k = 0; k < NK; ++k) t.kv[s * NK + k] = kv[k];
      __hip_atomic_fetch_add(new_groups, 1u, __ATOMIC_RELAXED, WS_LDS);
    }
    if (claimed || prev == h) {
      for (int a = 0; a < NA; ++a) atomic_combine_slot<WS_DEV>(plan.agg_bits[a] != 0, plan.ops[a], &t.acc[s * NA + a], vals[a]);
      return true;
    }
    s = (s + 1) & t.mask;
  }
  t.flags[0] = 1;
  return false;
}

// Key values of row r (integer keys as int64 bits, doubles normalised with +0.0) and the group
// hash over them: h = fold(mix64(h ^ bits(key_k)) + k) from 0x9e3779b97f4a7c15, with the free-slot
// marker mapped away (reference kernels.cuh:340-347).
template <class Ev>
__device__ __forceinline__ u64 row_keys(const Ev& ev, int NK, i64 r, double* kv) {
  u64 h = 0x9e3779b97f4a7c15ull;
#pragma unroll
  for (int k = 0; k < Ev::kKeys; ++k) {
    if (k >= NK) break;
    kv[k] = ev.key_int(k) ? bitsd(ev.ikey(k, r)) : ev.key(k, r) + 0.0;
    h = mix64(h ^ (u64)dbits(kv[k])) + (u64)k;
  }
  return h == kEmpty ? h ^ 1 : h;
}
// Slot values row r contributes (exact int64 bits for agg_bits slots, else doubles).
template <class Ev>
__device__ __forceinline__ void row_values(const Ev& ev, const ws_agg_plan& plan, int NA, i64 r, double* av) {
#pragma unroll
  for (int a = 0; a < Ev::kAggs; ++a) {
    if (a >= NA) break;
    av[a] = ev.agg_bits(a) ? bitsd(int_part(ev, a, ev.iagg(a, r))) : ev.agg(a, r);
  }
}

// The same from a row's loaded column words (a split evaluator's VRow).
template <class Ev>
__device__ __forceinline__ u64 row_keys_v(const Ev& ev, int NK, const typename Ev::VRow& x, i64 r, double* kv) {
  u64 h = 0x9e3779b97f4a7c15ull;
#pragma unroll
  for (int k = 0; k < Ev::kKeys; ++k) {
    if (k >= NK) break;
    kv[k] = ev.key_int(k) ? bitsd(ev.ikey_v(k, x, r)) : ev.key_v(k, x, r) + 0.0;
    h = mix64(h ^ (u64)dbits(kv[k])) + (u64)k;
  }
  return h == kEmpty ? h ^ 1 : h;
}
template <class Ev>
__device__ __forceinline__ void row_values_v(const Ev& ev, int NA, const typename Ev::VRow& x, i64 r, double* av) {
#pragma unroll
  for (int a = 0; a < Ev::kAggs; ++a) {
    if (a >= NA) break;
    av[a] = ev.agg_bits(a) ? bitsd(int_part(ev, a, ev.iagg_v(a, x, r))) : ev.agg_v(a, x, r);
  }
}
// Whether an evaluator has load halves for a whole row (WHERE words without loads of their own,
// key and aggregate words): compiled evaluators; the interpreters have none.
template <class Ev, class = void>
struct RowSplit {
  static constexpr bool value = false;
};
template <class Ev>
struct RowSplit<Ev, decltype((void)Ev::kRowSplit)> {
  static constexpr bool value = Ev::kRowSplit && Ev::kVSplit && Ev::kPureKeys && Ev::kPureVals && Ev::kPureWhere;
};
// The same for the WHERE and the keys alone (KRow).
template <class Ev, class = void>
struct KeySplit {
  static constexpr bool value = false;
};
template <class Ev>
struct KeySplit<Ev, decltype((void)Ev::kKSplit)> {
  static constexpr bool value = Ev::kRowSplit && Ev::kKSplit && Ev::kPureKeys && Ev::kPureWhere;
};
template <class Ev>
__device__ __forceinline__ u64 row_keys_k(const Ev& ev, int NK, const typename Ev::KRow& x, i64 r, double* kv) {
  u64 h = 0x9e3779b97f4a7c15ull;
#pragma unroll
  for (int k = 0; k < Ev::kKeys; ++k) {
    if (k >= NK) break;
    kv[k] = ev.key_int(k) ? bitsd(ev.ikey_k(k, x, r)) : ev.key_k(k, x, r) + 0.0;
    h = mix64(h ^ (u64)dbits(kv[k])) + (u64)k;
  }
  return h == kEmpty ? h ^ 1 : h;
}

// Hash filter + GROUP BY. Group identity is the 64-bit hash of the keys (as in the reference,
// kernels.cuh:340-347): h = fold(mix64(h ^ bits(key_k)) + k) from 0x9e3779b97f4a7c15.
// Each block keeps an open-addressing table of `lds_slots` groups in LDS for the whole launch
// (hot groups of skewed keys are combined there with LDS atomics); a row whose group finds no
// LDS slot within 8 probes goes straight to the global table. At the end the block flushes its
// LDS groups into the global table, and adds its count of new global groups once.
template <class Ev>
__device__ void grouped_body(const ScanDesc& d, GTable gt) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  __shared__ unsigned s_new;
  __shared__ int s_stop;
  Ev ev(d);
  const ws_agg_plan& plan = *d.plan;
  const int NK = ev.nkeys(), NA = ev.naggs();
  const int S = d.lds_slots;
  u64* lk = (u64*)ws_smem;
  double* lkv = (double*)(lk + S);
  double* lacc = lkv + (i64)S * NK;
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    lk[i] = kEmpty;
    for (int a = 0; a < NA; ++a) lacc[(i64)i * NA + a] = slot_identity(ev.agg_bits(a), ev.op(a));
  }
  if (threadIdx.x == 0) s_new = 0;
  BatchCursor bc;
  const i64 ntiles = uni(d.ntiles);
  for (i64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // Overflowed: the whole aggregation is redone bigger, stop early (block-uniform decision).
    __syncthreads();
    if (threadIdx.x == 0) s_stop = *(volatile int*)gt.flags;
    __syncthreads();
    if (s_stop) break;
    if (bc.seek(d, t)) {
      stage_batch(d, bc.b, scols);
      ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
    }
    const TileRows tr = bc.rows(t);
    const uint32_t* del = bc.del;
    for (i64 i = threadIdx.x; i < tr.count; i += blockDim.x) {
      const i64 r = tr.first + i;
      if (is_deleted(del, r)) continue;
      if (ev.has_where() && !ev.where(r)) continue;
      double kv[Ev::kKeys];
      double av[Ev::kAggs];
      const u64 h = row_keys(ev, NK, r, kv);
      row_values(ev, plan, NA, r, av);
      unsigned s = (unsigned)(h & (u64)(S - 1));
      bool done = false;
      for (int probe = 0; probe < 8; ++probe) {
        u64 prev = kEmpty;
        const bool claimed = __hip_atomic_compare_exchange_strong(&lk[s], &prev, h, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
        if (claimed)
          for (int k = 0; k < NK; ++k) lkv[(i64)s * NK + k] = kv[k];
        if (claimed || prev == h) {
#pragma unroll
          for (int a = 0; a < Ev::kAggs; ++a) {
            if (a >= NA) break;
            atomic_combine_slot<WS_LDS>(ev.agg_bits(a), ev.op(a), &lacc[(i64)s * NA + a], av[a]);
          }
          done = true;
          break;
        }
        s = (s + 1) & (unsigned)(S - 1);
      }
      if (!done) global_insert(gt, plan, NK, NA, h, kv, av, &s_new);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) s_stop = *(volatile int*)gt.flags;
  __syncthreads();
  for (int i = threadIdx.x; i < S && !s_stop; i += blockDim.x)
    if (lk[i] != kEmpty) global_insert(gt, plan, NK, NA, lk[i], &lkv[(i64)i * NK], &lacc[(i64)i * NA], &s_new);
  __syncthreads();
  if (threadIdx.x == 0 && s_new) {
    const u64 total = atomicAdd((unsigned long long*)gt.count, (unsigned long long)s_new) + s_new;
    if (total * 10 > (gt.mask + 1) * 7) gt.flags[0] = 1;  // load above 0.7: retry bigger
  }
}

// ------------------------------------------------------------------ batched dense GROUP BY

// Many dense aggregations (one plan shape, each with its own batches, constants and table) in one
// launch: the host lists every tile of every query; blocks take tiles in a grid-stride loop (no
// queue counter) and pre-aggregate each tile in an LDS table keyed by the dense group index, then
// fold it into the query's accumulators with device-scope atomics. A query with at most S groups
// keeps each group at its own slot (no hashing, no probes); otherwise the table holds twice the
// tile's possible groups (min(rows, G)), hashed, when that fits S; a tile whose possible groups do
// not fit, or of a query with many more groups than rows (little to combine), folds its rows
// straight into the query's accumulators. Values are combined in the final slot encoding of the shared-table
// dense layout (int64 bits for agg_bits slots, doubles otherwise), so the results equal a plain
// run of each query (up to double summation order).
struct BatchQuery {
  ScanDesc d;    // batches, column views, env, plan and dense spec of the query
  double* gacc;  // [G][NA] accumulators
  int* flags;    // [4] flags; [1] dense key out of range
  u64* count;    // group counter of the aggregation
  int32_t init;  // the accumulators, flags and counter still need their initial values
  int32_t pad;
};
// Initial values of the batch's accumulators, flags and counters (queries with init set), one
// launch for the whole batch: block (x, q) takes a share of query q's [G][NA] slots.
__device__ inline void batch_init_body(const BatchQuery* qs) {
  const BatchQuery& Q = qs[blockIdx.y];
  if (!Q.init) return;
  const ws_agg_plan& plan = *Q.d.plan;
  const int NA = plan.naggs > 0 ? plan.naggs : 1;
  if (blockIdx.x == 0) {
    if (threadIdx.x < 4) Q.flags[threadIdx.x] = 0;
    if (threadIdx.x == 4) *Q.count = 0;
  }
  const i64 n = Q.d.ds->G * NA;
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    const int a = (int)(i % NA);
    Q.gacc[i] = a < plan.naggs ? slot_identity(plan.agg_bits[a] != 0, plan.ops[a]) : 0.0;
  }
}
struct BatchTile {
  int32_t q, b;   // query, batch of the query
  int32_t count;  // rows
  int32_t glo;    // -1; or a key-range tile: only groups [glo, glo + S) of the query, in LDS
  int64_t first;  // first row in the batch
};

template <int SCOPE>
__device__ __forceinline__ void dense_fold(const ws_agg_plan& plan, uint8_t k, int a, double* dst, i64 iv, double dv) {
  if (plan.agg_bits[a]) atomic_combine_i<SCOPE>(dk_op(k), (i64*)dst, iv);
  else if (dk_int(k)) atomic_combine_d<SCOPE>(a == 0 ? WS_AGG_SUM : dk_op(k), dst, (double)iv);
  else atomic_combine_d<SCOPE>(dk_op(k), dst, dv);
}

constexpr int kBatchRows = 4;
constexpr i64 kSkipRow = -2;  // a row the step does not aggregate (dense_index gives -1 for a key out of range)
constexpr i64 kCopyGroups = 64;  // direct tables of at most this many groups are replicated per lane group
// Folds the LDS table (slots [0, St) of query Q; lg[i] >= 0: slot i holds group lg[i], a group may
// have several slots) into Q's accumulators.
__device__ inline void batch_flush(const BatchQuery& Q, const int* lg, const double* lacc, int St) {
  const ws_agg_plan& plan = *Q.d.plan;
  const ws_dense_spec& ds = *Q.d.ds;
  const int NA = plan.naggs;
  for (int i = threadIdx.x; i < St; i += blockDim.x) {
    const int g = lg[i];
    if (g < 0) continue;
    for (int a = 0; a < NA; ++a) {
      const double v = lacc[(i64)i * NA + a];
      const uint8_t k = ds.kind[a];
      if (plan.agg_bits[a]) atomic_combine_i<WS_DEV>(dk_op(k), (i64*)&Q.gacc[(i64)g * NA + a], dbits(v));
      else atomic_combine_d<WS_DEV>(a == 0 ? WS_AGG_SUM : dk_op(k), &Q.gacc[(i64)g * NA + a], v);
    }
  }
}
// Each block takes a contiguous range of the tile list (a query's tiles are consecutive there), so
// a query whose groups fit the LDS table directly keeps them there across its tiles and folds them
// into its accumulators once per block, when the block moves on to another query.
template <class Ev>
__device__ void batch_dense_body(const BatchQuery* qs, const BatchTile* tiles, i64 ntiles, int S) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  int* lg = (int*)ws_smem;                                    // [S] group index, -1 = free
  double* lacc = (double*)(ws_smem + (((i64)S * 4 + 15) & ~(i64)15));  // [S][NA]
  const int lane = threadIdx.x & (kWave - 1);
  const i64 per = (ntiles + gridDim.x - 1) / gridDim.x;
  const i64 t0 = (i64)blockIdx.x * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
  int tab_q = -1, tab_St = 0;  // the query whose groups the table holds, and its slots in use
  int tab_glo = -1;            // and its key range (key-range tiles)
  bool tab_keep = false;       // the table holds a direct layout that the query's next tile continues
  for (i64 t = t0; t < t1; ++t) {
    const BatchTile bt = tiles[t];
    const BatchQuery& Q = qs[bt.q];
    const ScanDesc& d = Q.d;
    const ws_agg_plan& plan = *d.plan;
    const ws_dense_spec& ds = *d.ds;
    Ev ev(d);
    const int NK = ev.nkeys(), NA = ev.naggs();
    // this tile's table: every group of the query at its own slot when G <= S (direct, kept across
    // the query's tiles); else 2 x the tile's possible groups, hashed (a power of two, >= 64,
    // <= S); else none
    const i64 G = ds.G, reach = bt.count < G ? (i64)bt.count : G;
    // a key-range tile: the groups [glo, glo + S) of a query with more groups than S, directly
    // at slot g - glo (the host gives such a query one tile list per range: each range reads the
    // query's rows again instead of sending them to device-scope atomics)
    const int glo = bt.glo;
    const bool ranged = glo >= 0;
    const bool direct = ranged || G <= (i64)S;
    // a few groups: C copies of the direct table, lane l using copy l % C, so the lanes of one LDS
    // atomic spread over C x G slots instead of contending for G
    const int C = !ranged && direct && G <= kCopyGroups ? (int)(S / G < 8 ? S / G : 8) : 1;
    int St = ranged ? (int)(G - glo < (i64)S ? G - glo : (i64)S) : direct ? (int)G * C : 0;
    if (!direct && G <= 4 * (i64)bt.count && 2 * reach <= S) {
      St = 64;
      while (St < S && St < 2 * reach) St <<= 1;
      St = St < S ? St : S;
    }
    const bool cont = tab_keep && tab_q == bt.q && tab_glo == glo;
    __syncthreads();  // the previous tile's rows are folded; its column views are no longer read
    if (!cont) {
      if (tab_q >= 0) {
        batch_flush(qs[tab_q], lg, lacc, tab_St);
        __syncthreads();
      }
      for (int i = threadIdx.x; i < St; i += blockDim.x) {
        lg[i] = -1;
        for (int a = 0; a < NA; ++a) lacc[(i64)i * NA + a] = slot_identity(plan.agg_bits[a] != 0, dk_op(ds.kind[a]));
      }
      tab_q = St > 0 ? bt.q : -1;
      tab_St = St;
      tab_glo = glo;
      tab_keep = direct;
    }
    for (int i = threadIdx.x; i < d.ncols; i += blockDim.x) scols[i] = d.cols[(i64)bt.b * d.ncols + i];
    __syncthreads();
    ev.bind(Ev::kGlobalCols ? d.cols + (i64)bt.b * d.ncols : scols);
    const uint32_t* del = uni(d.del[bt.b]);
    const int copy_base = (lane % C) * (int)G;
    bool bad = false;
    // kBatchRows rows per thread per step: their filter and key loads are issued before any fold
    // (evaluated for every row, rows past the tile re-reading its first, when the plan has no side
    // effects and the batch no deleted rows; row by row otherwise)
    constexpr int U = kBatchRows;
    const bool spec = Ev::kPureKeys && Ev::kPureWhere && !del;
    for (int i0 = threadIdx.x; i0 < bt.count; i0 += (int)blockDim.x * U) {
      i64 gs[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * (int)blockDim.x;
        gs[u] = kSkipRow;
        if (spec) {
          const i64 r = bt.first + (i < bt.count ? i : 0);
          const bool w = !ev.has_where() || ev.where(r);
          const i64 g = dense_index(ev, ds, NK, r);
          gs[u] = i < bt.count && w ? g : kSkipRow;
        } else if (i < bt.count) {
          const i64 r = bt.first + i;
          if (!is_deleted(del, r) && (!ev.has_where() || ev.where(r))) gs[u] = dense_index(ev, ds, NK, r);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const i64 g = gs[u];
        const i64 r = bt.first + i0 + u * (int)blockDim.x;
        if (g == -1) bad = true;
        const bool valid = g >= 0;
        if (!valid) continue;
        int slot = -1;
        if (ranged) {
          if (g < glo || g >= glo + St) continue;  // another range's group
          slot = (int)g - glo;
          if (lg[slot] < 0) lg[slot] = (int)g;
        } else if (direct) {
          slot = copy_base + (int)g;
          if (lg[slot] < 0) lg[slot] = (int)g;  // occupied (a plain store: every writer stores the same)
        } else if (St > 0) {
          int s = (int)(g & (St - 1));
          for (int probe = 0; probe < 8; ++probe) {
            int prev = __hip_atomic_load(&lg[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prev == -1 && __hip_atomic_compare_exchange_strong(&lg[s], &prev, (int)g, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_WORKGROUP))
              prev = (int)g;
            if (prev == (int)g) {
              slot = s;
              break;
            }
            s = (s + 1) & (St - 1);
          }
        }
        double* dst = slot >= 0 ? lacc + (i64)slot * NA : Q.gacc + g * NA;
#pragma unroll
        for (int a = 0; a < Ev::kAggs; ++a) {
          if (a >= NA) break;
          const uint8_t k = ev.dkind(a, ds);
          i64 iv = 0;
          double dv = 0;
          dense_value(ev, plan, k, a, r, iv, dv);
          if (slot >= 0) dense_fold<WS_LDS>(plan, k, a, &dst[a], iv, dv);
          else dense_fold<WS_DEV>(plan, k, a, &dst[a], iv, dv);
        }
      }
    }
    if (bad) Q.flags[1] = 1;
  }
  __syncthreads();
  if (tab_q >= 0) batch_flush(qs[tab_q], lg, lacc, tab_St);
}

// ------------------------------------------------------------------ partitioned hash GROUP BY

// High-cardinality GROUP BY without global atomics on a shared table. Three passes over a plan:
//   count   (NAME_pcount): block c scans chunk c (chunk_tiles tiles) and counts its surviving
//           rows per partition (top pbits of the group hash) in LDS -> part_cnt[c][p];
//   (host-side kernels turn the counts into part_start[p] and per-chunk offsets part_off[c][p]);
//   scatter (NAME_pscatter): block c scans chunk c again and writes one record per row into its
//           partition's run [part_start[p] + part_off[c][p], + part_cnt[c][p]). Groups that repeat
//           inside the chunk are first combined in a small LDS table (first come, first served),
//           which flushes one record per group at the end; the unused tail of each run is marked
//           with the free hash. A run that would exceed its count raises part_err instead of
//           writing outside it.
//   merge   (part_agg_k): one block per partition combines its records in an LDS table and
//           appends the finished groups to the output.
// A record is [hash, key values (nk), slot values (na)], 8 bytes each.
// Rows per thread per step of the partition passes: loads of several rows overlap, but the
// 1024-thread blocks leave each thread at most 128 VGPRs, and past that the passes spill to
// scratch memory (at 4 rows, the web table's compiled plan did), which costs more than the
// overlap gains.
constexpr int kPartRows = 4;

template <class Ev>
__device__ void part_count_body(const ScanDesc& d) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  Ev ev(d);
  const int NK = ev.nkeys();
  const int P = 1 << d.pbits, shift = 64 - d.pbits;
  uint32_t* hist = (uint32_t*)ws_smem;
  for (int i = threadIdx.x; i < P; i += blockDim.x) hist[i] = 0;
  // key ranges for the packed record layout (integer keys only; the host sets key_range then):
  // per thread, then per block in LDS, one partial per chunk
  __shared__ long long s_rng[2 * kPackKeys];
  const bool ranged = d.key_range != nullptr;
  if (threadIdx.x < 2 * kPackKeys) s_rng[threadIdx.x] = threadIdx.x < kPackKeys ? kI64Max : kI64Min;
  i64 kmin[kPackKeys], kmax[kPackKeys];
#pragma unroll
  for (int k = 0; k < kPackKeys; ++k) kmin[k] = kI64Max, kmax[k] = kI64Min;
  const i64 c = blockIdx.x;
  const i64 t0 = c * d.chunk_tiles, t1 = t0 + d.chunk_tiles < d.ntiles ? t0 + d.chunk_tiles : d.ntiles;
  // U rows per thread per step, evaluated before any LDS update so their loads overlap (the
  // interpreter's out-of-line VM calls gain nothing from it: one row there)
  constexpr int U = Ev::kRows > 1 ? kPartRows : 1;
  auto widen = [&](const double* kv) {
#pragma unroll
    for (int k = 0; k < kPackKeys; ++k) {
      if (k >= NK || k >= Ev::kKeys) break;
      const i64 v = dbits(kv[k]);
      kmin[k] = v < kmin[k] ? v : kmin[k];
      kmax[k] = v > kmax[k] ? v : kmax[k];
    }
  };
  BatchCursor bc;
  if constexpr (KeySplit<Ev>::value) {
    // Compiled plans in load and compute halves: tile t + 1's WHERE and key words load while tile
    // t's rows are hashed and counted (one tile = one step of U rows per thread; the loop
    // alternates two register sets, as a copy between them would wait for the loads).
    struct Step {
      typename Ev::Row w[U];
      typename Ev::KRow k[U];
    };
    auto load_step = [&](const TileRows& tr, Step& x) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const i64 ii = threadIdx.x + (i64)u * blockDim.x;
        const i64 r = tr.first + (ii < tr.count ? ii : 0);
        ev.load_row(r, x.w[u]);
        ev.load_krow(r, x.k[u]);
      }
    };
    bool have = false;
    auto tile = [&](i64 t, Step& xc, Step& xn) {
      if (bc.seek(d, t)) {
        have = false;
        stage_batch(d, bc.b, scols);  // also orders the histogram reset before the first update
        ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
      }
      const TileRows tr = bc.rows(t);
      if (bc.del || bc.tr != (i64)blockDim.x * U) {  // (block-uniform) row by row
        have = false;
        for (i64 i = threadIdx.x; i < tr.count; i += blockDim.x) {
          const i64 r = tr.first + i;
          if (is_deleted(bc.del, r)) continue;
          double kv[Ev::kKeys];
          const u64 h = row_keys(ev, NK, r, kv);
          if (ev.has_where() && !ev.where(r)) continue;
          if (ranged) widen(kv);
          __hip_atomic_fetch_add(&hist[h >> shift], 1u, __ATOMIC_RELAXED, WS_LDS);
        }
        return;
      }
      if (!have) load_step(tr, xc);
      // the next tile of this batch (else this tile again: the loads stay unconditional)
      const bool nxt = t + 1 < t1 && t + 1 < bc.t_hi;
      load_step(nxt ? bc.rows(t + 1) : tr, xn);
      u64 h[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const i64 ii = threadIdx.x + (i64)u * blockDim.x;
        const i64 r = tr.first + (ii < tr.count ? ii : 0);
        double kv[Ev::kKeys];
        const u64 hk = row_keys_k(ev, NK, xc.k[u], r, kv);
        h[u] = ii < tr.count && ev.where_row(xc.w[u], r) ? hk : kEmpty;
        if (ranged && h[u] != kEmpty) widen(kv);
      }
      have = nxt;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (h[u] != kEmpty) __hip_atomic_fetch_add(&hist[h[u] >> shift], 1u, __ATOMIC_RELAXED, WS_LDS);
    };
    Step xa, xb;
    for (i64 t = t0; t < t1; t += 2) {
      tile(t, xa, xb);
      if (t + 1 < t1) tile(t + 1, xb, xa);
    }
  } else {
    for (i64 t = t0; t < t1; ++t) {
      if (bc.seek(d, t)) {
        stage_batch(d, bc.b, scols);  // also orders the histogram reset before the first update
        ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
      }
      const TileRows tr = bc.rows(t);
      const uint32_t* del = bc.del;
      for (i64 i = threadIdx.x; i < tr.count; i += (i64)blockDim.x * U) {
        u64 h[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const i64 ii = i + (i64)u * blockDim.x;
          const i64 r = tr.first + ii;
          h[u] = kEmpty;
          double kv[Ev::kKeys];
          if (Ev::kPureKeys && Ev::kPureWhere && !del) {
            // (block-uniform) no side effects, no deleted rows: evaluated without a branch (rows
            // past the tile re-read its first row), so the step's rows keep their loads together
            const i64 rr = tr.first + (ii < tr.count ? ii : 0);
            const u64 hk = row_keys(ev, NK, rr, kv);
            const bool w = !ev.has_where() || ev.where(rr);
            h[u] = ii < tr.count && w ? hk : kEmpty;
          } else if (Ev::kPureKeys) {  // key loads issued with the WHERE loads, not after them
            if (ii < tr.count) {
              const u64 hk = row_keys(ev, NK, r, kv);
              if (!is_deleted(del, r) && (!ev.has_where() || ev.where(r))) h[u] = hk;
            }
          } else if (ii < tr.count && !is_deleted(del, r) && (!ev.has_where() || ev.where(r))) {
            h[u] = row_keys(ev, NK, r, kv);
          }
          if (ranged && h[u] != kEmpty) widen(kv);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (h[u] != kEmpty) __hip_atomic_fetch_add(&hist[h[u] >> shift], 1u, __ATOMIC_RELAXED, WS_LDS);
      }
    }
  }
  if (ranged) {  // one LDS atomic per wave and bound
#pragma unroll
    for (int k = 0; k < kPackKeys; ++k) {
      if (k >= NK || k >= Ev::kKeys) break;
      i64 lo = kmin[k], hi = kmax[k];
      for (int o = 32; o > 0; o >>= 1) {
        const i64 a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
      }
      if ((threadIdx.x & 63) == 0 && lo <= hi) {
        atomicMin(&s_rng[k], (long long)lo);
        atomicMax(&s_rng[kPackKeys + k], (long long)hi);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += blockDim.x) d.part_cnt[c * P + i] = hist[i];
  if (ranged && threadIdx.x < 2 * NK) {
    const int k = threadIdx.x < NK ? threadIdx.x : threadIdx.x - NK;
    d.key_range[c * 2 * NK + threadIdx.x] = s_rng[(threadIdx.x < NK ? 0 : kPackKeys) + k];
  }
}

// Writes one record: [hash, key values, slot values], or with a packing [packed keys, slot values].
struct RecLayout {
  bool packed;
  int words;  // record stride
  int32_t shift[kPackKeys];
  int64_t kmin[kPackKeys];
};
__device__ __forceinline__ RecLayout rec_layout(const ScanDesc& d, int NK, int NA) {
  RecLayout L;
  L.packed = uni(d.pack != nullptr && d.pack->ok ? 1 : 0) != 0;
  L.words = L.packed ? 1 + NA : d.rec_words;
#pragma unroll
  for (int k = 0; k < kPackKeys; ++k) {
    L.shift[k] = uni(L.packed ? d.pack->shift[k] : 0);
    L.kmin[k] = uni(L.packed ? (i64)d.pack->kmin[k] : (i64)0);
  }
  return L;
}
__device__ __forceinline__ void put_record(const RecLayout& L, u64 WS_GLOBAL* rec, u64 h, int NK, int NA, const double* kv,
                                           const double* av) {
  if (L.packed) {
    u64 pk = 0;
#pragma unroll
    for (int k = 0; k < kPackKeys; ++k)
      if (k < NK) pk |= (u64)(dbits(kv[k]) - L.kmin[k]) << L.shift[k];
    rec[0] = pk;
    for (int a = 0; a < NA; ++a) rec[1 + a] = (u64)dbits(av[a]);
    return;
  }
  rec[0] = h;
  for (int k = 0; k < NK; ++k) rec[1 + k] = (u64)dbits(kv[k]);
  for (int a = 0; a < NA; ++a) rec[1 + NK + a] = (u64)dbits(av[a]);
}

template <class Ev>
__device__ void part_scatter_body(const ScanDesc& d) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  Ev ev(d);
  const ws_agg_plan& plan = *d.plan;
  const int NK = ev.nkeys(), NA = ev.naggs();
  const RecLayout L = rec_layout(d, NK, NA);
  const int W = L.words;
  u64 WS_GLOBAL* recs = (u64 WS_GLOBAL*)d.recs;
  const int P = 1 << d.pbits, shift = 64 - d.pbits, S = d.pre_slots;
  const i64 c = blockIdx.x;
  uint32_t* cur_pos = (uint32_t*)ws_smem;  // [P] next record of this chunk's run
  uint32_t* end_pos = cur_pos + P;         // [P] end of the run
  u64* pk = (u64*)(end_pos + P);           // pre-aggregation table: hashes [S]
  double* pkv = (double*)(pk + S);         // [S * NK]
  double* pacc = pkv + (i64)S * NK;        // [S * NA]
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    const uint32_t b0 = d.part_start[p] + d.part_off[c * P + p];
    cur_pos[p] = b0;
    end_pos[p] = b0 + d.part_cnt[c * P + p];
  }
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    pk[i] = kEmpty;
    for (int a = 0; a < NA; ++a) pacc[(i64)i * NA + a] = slot_identity(ev.agg_bits(a), ev.op(a));
  }
  const i64 t0 = c * d.chunk_tiles, t1 = t0 + d.chunk_tiles < d.ntiles ? t0 + d.chunk_tiles : d.ntiles;
  // Ev::kRows rows per thread per step: all are evaluated (their loads overlap) before the LDS
  // table and the record writes see them.
  constexpr int U = Ev::kRows > 1 && Ev::kKeys + Ev::kAggs <= 6 ? kPartRows : 1;  // register budget
  // A run that would overflow is noted in a register and reported once at the end: a store through
  // the descriptor's generic pointer inside the loop would be a flat store, after which the
  // compiler waits for every load in flight (the next tile's included) at the next use of any.
  bool perr = false;
  // One row into the pre-aggregation table (two probes; a plain LDS read first, so a full table
  // costs reads, not compare-and-swaps), else a record in its partition's run.
  auto insert = [&](u64 h, const double* kv, const double* av) {
    unsigned s = (unsigned)(h & (u64)(S - 1));
    for (int probe = 0; probe < 2; ++probe) {
      u64 prev = __hip_atomic_load(&pk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      bool claimed = false;
      if (prev == kEmpty) {
        claimed = __hip_atomic_compare_exchange_strong(&pk[s], &prev, h, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        if (claimed)
          for (int k = 0; k < NK; ++k) pkv[(i64)s * NK + k] = kv[k];
      }
      if (claimed || prev == h) {
#pragma unroll
        for (int a = 0; a < Ev::kAggs; ++a) {
          if (a >= NA) break;
          atomic_combine_slot<WS_LDS>(ev.agg_bits(a), ev.op(a), &pacc[(i64)s * NA + a], av[a]);
        }
        return;
      }
      s = (s + 1) & (unsigned)(S - 1);
    }
    const unsigned p = (unsigned)(h >> shift);
    const uint32_t pos = __hip_atomic_fetch_add(&cur_pos[p], 1u, __ATOMIC_RELAXED, WS_LDS);
    if (pos >= end_pos[p]) {
      perr = true;
      return;
    }
    put_record(L, recs + (i64)pos * W, h, NK, NA, kv, av);
  };
  // The rows of tile tr, steps of U rows per thread (the evaluation of a step's rows issues all
  // their loads before any result is needed when the plan allows).
  auto tile_rows = [&](const TileRows& tr, const uint32_t* del) {
    for (i64 i = threadIdx.x; i < tr.count; i += (i64)blockDim.x * U) {
      double kv[U][Ev::kKeys];
      double av[U][Ev::kAggs];
      u64 h[U];
      if (Ev::kPureKeys && Ev::kPureVals && Ev::kPureWhere && !del) {
        // (block-uniform) No side effects and no deleted rows: every row of the step is evaluated
        // (rows past the tile re-read its first row) with no branch between them.
        bool w[U];
        u64 hk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const i64 ii = i + (i64)u * blockDim.x;
          const i64 r = tr.first + (ii < tr.count ? ii : 0);
          hk[u] = row_keys(ev, NK, r, kv[u]);
          row_values(ev, plan, NA, r, av[u]);
          w[u] = !ev.has_where() || ev.where(r);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) h[u] = i + (i64)u * blockDim.x < tr.count && w[u] ? hk[u] : kEmpty;
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const i64 ii = i + (i64)u * blockDim.x;
          const i64 r = tr.first + ii;
          h[u] = kEmpty;
          if (Ev::kPureKeys && Ev::kPureVals) {  // all loads of the row issued together
            if (ii < tr.count) {
              const u64 hk = row_keys(ev, NK, r, kv[u]);
              row_values(ev, plan, NA, r, av[u]);
              if (!is_deleted(del, r) && (!ev.has_where() || ev.where(r))) h[u] = hk;
            }
          } else if (ii < tr.count && !is_deleted(del, r) && (!ev.has_where() || ev.where(r))) {
            h[u] = row_keys(ev, NK, r, kv[u]);
            row_values(ev, plan, NA, r, av[u]);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (h[u] != kEmpty) insert(h[u], kv[u], av[u]);
    }
  };
  BatchCursor bc;
  if constexpr (RowSplit<Ev>::value) {
    // Compiled plans in load and compute halves: while tile t's rows go into the table and their
    // records are written, tile t + 1's column words are already loading (its loads are issued
    // before this tile's record stores, so waiting for them does not wait for those stores). A tile
    // is one step (tile rows = U per thread); the loop alternates two register sets (a copy from
    // one to the other would wait for the loads).
    struct Step {
      typename Ev::Row w[U];
      typename Ev::VRow v[U];
    };
    auto load_step = [&](const TileRows& tr, Step& x) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const i64 ii = threadIdx.x + (i64)u * blockDim.x;
        const i64 r = tr.first + (ii < tr.count ? ii : 0);
        ev.load_row(r, x.w[u]);
        ev.load_vrow(r, x.v[u]);
      }
    };
    bool have = false;
    auto tile = [&](i64 t, Step& xc, Step& xn) {
      if (bc.seek(d, t)) {
        have = false;
        stage_batch(d, bc.b, scols);  // also orders the table set-up before the first row
        ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
      }
      const TileRows tr = bc.rows(t);
      if (bc.del || bc.tr != (i64)blockDim.x * U) {  // (block-uniform)
        // row by row (both register sets stay live across this branch: no room for U rows)
        have = false;
        for (i64 i = threadIdx.x; i < tr.count; i += blockDim.x) {
          const i64 r = tr.first + i;
          if (is_deleted(bc.del, r)) continue;
          double kv[Ev::kKeys], av[Ev::kAggs];
          const u64 h = row_keys(ev, NK, r, kv);
          row_values(ev, plan, NA, r, av);
          if (!ev.has_where() || ev.where(r)) insert(h, kv, av);
        }
        return;
      }
      if (!have) load_step(tr, xc);
      // the next tile of this batch (else this tile again: the loads stay unconditional, since a
      // branch around them would make the compiler wait for every load in flight)
      const bool nxt = t + 1 < t1 && t + 1 < bc.t_hi;
      load_step(nxt ? bc.rows(t + 1) : tr, xn);
      have = nxt;
      // every row's hash and WHERE first (the waits for this tile's words then come before any
      // record store, which the same counter would otherwise make them wait for), then one row at
      // a time into the table (only one row's keys and values live at once)
      u64 h[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const i64 ii = threadIdx.x + (i64)u * blockDim.x;
        const i64 r = tr.first + (ii < tr.count ? ii : 0);
        double kv[Ev::kKeys];
        const u64 hk = row_keys_v(ev, NK, xc.v[u], r, kv);
        h[u] = ii < tr.count && ev.where_row(xc.w[u], r) ? hk : kEmpty;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (h[u] == kEmpty) continue;
        const i64 r = tr.first + threadIdx.x + (i64)u * blockDim.x;
        double kv[Ev::kKeys], av[Ev::kAggs];
        row_keys_v(ev, NK, xc.v[u], r, kv);
        row_values_v(ev, NA, xc.v[u], r, av);
        insert(h[u], kv, av);
      }
    };
    Step xa, xb;
    for (i64 t = t0; t < t1; t += 2) {
      tile(t, xa, xb);
      if (t + 1 < t1) tile(t + 1, xb, xa);
    }
  } else {
    for (i64 t = t0; t < t1; ++t) {
      if (bc.seek(d, t)) {
        stage_batch(d, bc.b, scols);  // also orders the table set-up before the first row
        ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
      }
      tile_rows(bc.rows(t), bc.del);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    const u64 h = pk[i];
    if (h == kEmpty) continue;
    const unsigned p = (unsigned)(h >> shift);
    const uint32_t pos = __hip_atomic_fetch_add(&cur_pos[p], 1u, __ATOMIC_RELAXED, WS_LDS);
    if (pos >= end_pos[p]) {
      perr = true;
      continue;
    }
    put_record(L, recs + (i64)pos * W, h, NK, NA, &pkv[(i64)i * NK], &pacc[(i64)i * NA]);
  }
  if (perr) *d.part_err = 1;
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x)
    for (uint32_t pos = cur_pos[p]; pos < end_pos[p]; ++pos) recs[(i64)pos * W] = kEmpty;
}

// ------------------------------------------------------------------ filter + stream compaction

// Tile status word for decoupled look-back: [63:62] 0 = not ready, 1 = tile count, 2 = inclusive
// prefix; [61:0] the value.
constexpr u64 kLbAgg = 1ull << 62, kLbInc = 2ull << 62, kLbMask = (1ull << 62) - 1;
constexpr int kSelRows = 4;  // rows per thread of a select tile: tile = 4 * blockDim rows
constexpr int kSelChunkRows = 8;  // rows per thread of a chunked select's tile: tile = 8 * blockDim rows

// Wave-wide decoupled look-back, called by one full wave of the block that owns tile t: publishes
// the tile's count, then sums the counts of earlier tiles (64 per step) until an inclusive prefix is
// found, publishes its own inclusive prefix and returns the exclusive one (in every lane).
// Predecessor tiles were handed out earlier by an atomic counter, so they are running and publish;
// a wait that still does not end is reported through *err instead of hanging the device.
__device__ __forceinline__ u64 lookback_exclusive(u64* state, i64 t, u64 tile_count, int* err) {
  const int lane = threadIdx.x & (kWave - 1);
  if (lane == 0)
    __hip_atomic_store(&state[t], (t == 0 ? kLbInc : kLbAgg) | tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  u64 excl = 0;
  i64 pos = t - 1;
  unsigned spins = 0;
  while (pos >= 0) {
    const i64 idx = pos - lane;
    const u64 st = idx >= 0 ? __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kLbInc;
    if (__ballot((st >> 62) == 0)) {
      if (++spins > (1u << 22)) {
        if (lane == 0) *err = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    const u64 incl = __ballot((st >> 62) == 2);
    if (incl) {
      const int first = __ffsll((unsigned long long)incl) - 1;  // nearest inclusive predecessor
      excl += wave_sum_u64(lane <= first ? (st & kLbMask) : 0);
      break;
    }
    excl += wave_sum_u64(st & kLbMask);
    pos -= kWave;
  }
  if (lane == 0 && t > 0)
    __hip_atomic_store(&state[t], kLbInc | (excl + tile_count), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// Slot compaction: out_idx gets, in ascending order, every i in [0, n) with pred(i); total[0] the
// count. One tile of kSelRows * blockDim slots per block, tiles in counter order.
template <class P>
__device__ void compact_body(const P& pred, i64 n, u64* state, unsigned* tile_ctr, int32_t* out_idx, u64* total,
                             int* err) {
  __shared__ unsigned s_tile;
  __shared__ unsigned s_grp[kSelRows * 16];
  __shared__ u64 s_excl;
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const i64 t = s_tile;
  const i64 tile = (i64)kSelRows * blockDim.x;
  if (t * tile >= n) return;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  bool pass[kSelRows];
  unsigned rank[kSelRows];
#pragma unroll
  for (int j = 0; j < kSelRows; ++j) {
    const i64 i = t * tile + (i64)j * blockDim.x + threadIdx.x;
    pass[j] = i < n && pred(i);
    const u64 m = __ballot(pass[j]);
    rank[j] = (unsigned)__popcll(m & ((1ull << lane) - 1));
    if (lane == 0) s_grp[j * nw + wave] = (unsigned)__popcll(m);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned run = 0;
    for (int g = 0; g < kSelRows * nw; ++g) {
      const unsigned c = s_grp[g];
      s_grp[g] = run;
      run += c;
    }
    s_excl = run;
  }
  __syncthreads();
  const u64 tile_count = s_excl;
  __syncthreads();
  if (wave == 0) {
    const u64 excl = lookback_exclusive(state, t, tile_count, err);
    if (lane == 0) {
      s_excl = excl;
      if ((t + 1) * tile >= n) *total = excl + tile_count;
    }
  }
  __syncthreads();
  const u64 base = s_excl;
#pragma unroll
  for (int j = 0; j < kSelRows; ++j)
    if (pass[j]) out_idx[base + s_grp[j * nw + wave] + rank[j]] = (int32_t)(t * tile + (i64)j * blockDim.x + threadIdx.x);
}

// Single-pass WHERE + compaction over all batches (replaces a flag array + Flagged select): each
// block takes the next tile id from a counter, evaluates the filter for its rows, ranks passing
// rows with wave ballots and an LDS scan, learns its output offset by decoupled look-back over
// earlier tiles, and writes ascending global row indices (row0[batch] + row). Per-batch counts are
// accumulated in per_batch[b]; total[0] gets the count of all selected rows.
template <class Ev>
__device__ void select_body(const ScanDesc& d, u64* state, unsigned* tile_ctr, int32_t* out_idx, u64* per_batch,
                            u64* total, int* err) {
  __shared__ ws_colview scols[WS_MAX_COLS];
  __shared__ unsigned s_tile;
  __shared__ unsigned s_wcount[16];
  __shared__ u64 s_excl;
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const i64 t = s_tile;
  if (t >= d.ntiles) return;
  const int b = batch_of_tile_bsearch(d, t);
  for (int i = threadIdx.x; i < d.ncols; i += blockDim.x) scols[i] = d.cols[(i64)b * d.ncols + i];
  __syncthreads();
  Ev ev(d);
  ev.bind(Ev::kGlobalCols ? d.cols + (i64)b * d.ncols : scols);
  const TileRows tr = tile_rows(d, t, b);
  const uint32_t* del = uni(d.del[b]);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  // Each thread owns kRows consecutive chunks of blockDim rows: row i = j * blockDim + tid.
  bool pass[kSelRows];
  unsigned mycount[kSelRows];
#pragma unroll
  for (int j = 0; j < kSelRows; ++j) {
    const i64 i = (i64)j * blockDim.x + threadIdx.x;
    const i64 r = tr.first + i;
    pass[j] = i < tr.count && !is_deleted(del, r) && (!ev.has_where() || ev.where(r));
  }
  // Rank: rows are ordered (j, wave, lane); per (j, wave) a ballot, then an LDS prefix over the
  // (j, wave) groups in that order.
  __shared__ unsigned s_grp[4 * 16];
  unsigned local_rank[kSelRows];
#pragma unroll
  for (int j = 0; j < kSelRows; ++j) {
    const u64 m = __ballot(pass[j]);
    local_rank[j] = (unsigned)__popcll(m & ((1ull << lane) - 1));
    mycount[j] = (unsigned)__popcll(m);
    if (lane == 0) s_grp[j * nw + wave] = mycount[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned run = 0;
    for (int g = 0; g < kSelRows * nw; ++g) {
      const unsigned c = s_grp[g];
      s_grp[g] = run;
      run += c;
    }
    s_wcount[0] = run;  // tile count
  }
  __syncthreads();
  const u64 tile_count = s_wcount[0];
  // Decoupled look-back (wave 0) for this tile's output offset.
  if (wave == 0) {
    const u64 excl = lookback_exclusive(state, t, tile_count, err);
    if (lane == 0) {
      s_excl = excl;
      if (tile_count) atomicAdd((unsigned long long*)&per_batch[b], (unsigned long long)tile_count);
      if (t == d.ntiles - 1) *total = excl + tile_count;
    }
  }
  __syncthreads();
  const u64 base = s_excl;
  const i64 g0 = d.row0[b] + tr.first;
#pragma unroll
  for (int j = 0; j < kSelRows; ++j) {
    if (!pass[j]) continue;
    const i64 i = (i64)j * blockDim.x + threadIdx.x;
    out_idx[base + s_grp[j * nw + wave] + local_rank[j]] = (int32_t)(g0 + i);
  }
}

// Chunked single-pass WHERE + compaction for long inputs (the tile-per-block form above pays an
// atomic tile fetch, a look-back and four barriers per 1024 rows). A block takes the next chunk
// (chunk_tiles consecutive tiles) from a counter, evaluates every row of it keeping only the
// wave ballots in LDS (8 B per 64 rows), looks back once over earlier chunks for its output
// offset, then replays the ballots in (tile, step, wave) order and writes ascending global row
// indices. Chunks are handed out in order, so every predecessor a block waits for is running.
template <class Ev>
__device__ void select_chunk_body(const ScanDesc& d, u64* state, unsigned* chunk_ctr, int32_t* out_idx, u64* per_batch,
                                  u64* total, int* err) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  __shared__ unsigned s_chunk;
  __shared__ u64 s_excl;
  __shared__ uint32_t s_w[16];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  const int per_tile = kSelChunkRows * nw;  // ballots per tile
  Ev ev(d);
  {
    if (threadIdx.x == 0) s_chunk = atomicAdd(chunk_ctr, 1u);
    __syncthreads();
    const i64 c = s_chunk;
    const i64 t0 = c * d.chunk_tiles, t1 = t0 + d.chunk_tiles < d.ntiles ? t0 + d.chunk_tiles : d.ntiles;
    if (t0 >= d.ntiles) return;  // (block-uniform)
    u64* s_mask = (u64*)ws_smem;                         // [tiles][kSelRows][nw]
    uint32_t* s_pre = (uint32_t*)(s_mask + (t1 - t0) * per_tile);  // exclusive prefix of their counts
    BatchCursor bc;
    u64 bcount = 0;  // this wave's passing rows in batch bc.b, added to per_batch when the batch ends
    for (i64 t = t0; t < t1; ++t) {
      const int prev = bc.b;
      if (bc.seek(d, t)) {
        if (prev >= 0 && lane == 0 && bcount) atomicAdd((unsigned long long*)&per_batch[prev], (unsigned long long)bcount);
        bcount = 0;
        stage_batch(d, bc.b, scols);
        ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
      }
      const TileRows tr = bc.rows(t);
      const uint32_t* del = bc.del;
      bool pass[kSelChunkRows];
      if (Ev::kPureWhere || !del) {  // (block-uniform)
        // Every row's WHERE is evaluated (rows past the tile re-read its first row), so the loads of
        // all the thread's rows issue before any result is needed. Safe when the WHERE has no side
        // effect, or when no row is deleted: then each evaluated row is a live row anyway.
        bool w[kSelChunkRows];
#pragma unroll
        for (int j = 0; j < kSelChunkRows; ++j) {
          const i64 i = (i64)j * blockDim.x + threadIdx.x;
          w[j] = !ev.has_where() || ev.where(tr.first + (i < tr.count ? i : 0));
        }
#pragma unroll
        for (int j = 0; j < kSelChunkRows; ++j) {
          const i64 i = (i64)j * blockDim.x + threadIdx.x;
          pass[j] = i < tr.count && w[j] && !is_deleted(del, tr.first + i);
        }
      } else {
#pragma unroll
        for (int j = 0; j < kSelChunkRows; ++j) {
          const i64 i = (i64)j * blockDim.x + threadIdx.x;
          const i64 r = tr.first + i;
          pass[j] = i < tr.count && !is_deleted(del, r) && (!ev.has_where() || ev.where(r));
        }
      }
#pragma unroll
      for (int j = 0; j < kSelChunkRows; ++j) {
        const u64 m = __ballot(pass[j]);
        if (lane == 0) s_mask[(t - t0) * per_tile + j * nw + wave] = m;
        bcount += (u64)__popcll(m);
      }
    }
    if (lane == 0 && bcount) atomicAdd((unsigned long long*)&per_batch[bc.b], (unsigned long long)bcount);
    __syncthreads();
    // Exclusive prefix of the ballots' counts in (tile, step, wave) order: per thread a run of
    // consecutive ballots, then a block scan of the run totals.
    const int nm = (int)(t1 - t0) * per_tile;
    const int K = (nm + (int)blockDim.x - 1) / (int)blockDim.x;
    const int m0 = threadIdx.x * K, m1 = m0 + K < nm ? m0 + K : nm;
    uint32_t run = 0;
    for (int q = m0; q < m1; ++q) run += (uint32_t)__popcll(s_mask[q]);
    uint32_t incl = run;
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0, count = 0;
    for (int w = 0; w < nw; ++w) {
      if (w < wave) wbase += s_w[w];
      count += s_w[w];
    }
    uint32_t pre = wbase + incl - run;
    for (int q = m0; q < m1; ++q) {
      s_pre[q] = pre;
      pre += (uint32_t)__popcll(s_mask[q]);
    }
    if (wave == 0) {
      const u64 excl = lookback_exclusive(state, c, count, err);
      if (lane == 0) {
        s_excl = excl;
        if (t1 == d.ntiles) *total = excl + count;
      }
    }
    __syncthreads();
    const u64 base = s_excl;
    BatchCursor wc;
    for (i64 t = t0; t < t1; ++t) {
      wc.seek(d, t);
      const TileRows tr = wc.rows(t);
      const i64 g0 = wc.row0 + tr.first;
#pragma unroll
      for (int j = 0; j < kSelChunkRows; ++j) {
        const int q = (int)(t - t0) * per_tile + j * nw + wave;
        const u64 m = s_mask[q];
        if ((m >> lane) & 1)
          out_idx[base + s_pre[q] + (u64)__popcll(m & ((1ull << lane) - 1))] = (int32_t)(g0 + (i64)j * blockDim.x + threadIdx.x);
      }
    }
  }
}

// Selection over large set bitmaps, in two kernels. select_mask_body: one block of 1024 threads per
// CU copies the leading words of the plan's set bitmaps into LDS (d.set_words of them), then scans
// a contiguous share of the tiles (kSelMaskRows rows per thread per tile) and stores each wave's
// ballots to masks[tile][tile_rows / 64], in row order; no look-back or barrier in the loop, so the
// block keeps its loads in flight, and adds each tile's count to tile_cnt[tile]. scan_counts_body
// and select_expand_body then turn the bits into ascending indices.
constexpr int kSelMaskRows = 8;
template <class Ev>
__device__ void select_mask_body(const ScanDesc& d, u64* masks, uint32_t* tile_cnt, i64* tile_row) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  const i64 words = (i64)kSelMaskRows * nw;  // ballots per tile
  Ev ev(d);
  ev.stage_sets((unsigned*)ws_smem, d.set_words);
  __syncthreads();
  const i64 t0 = (i64)blockIdx.x * d.ntiles / gridDim.x, t1 = (i64)(blockIdx.x + 1) * d.ntiles / gridDim.x;
  BatchCursor bc;
  // Compiled WHEREs in two halves (Ev::kSplit): later full tiles' column words load while this tile
  // is computed (xa, xb hold this tile's and the next one's rows when `have`).
  typename Ev::Row4 xa[kSelMaskRows / 4];
  bool have = false;
  for (i64 t = t0; t < t1; ++t) {
    if (bc.seek(d, t)) {
      have = false;
      stage_batch(d, bc.b, scols);
      ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
    }
    const TileRows tr = bc.rows(t);
    const uint32_t* del = bc.del;
    u64* tm = masks + t * words;
    if constexpr (Ev::kSplit) {
      if (!del && tr.count == bc.tr && ev.rows4_ok()) {  // (block-uniform)
        // Four consecutive rows per lane and load: lane L of wave w takes rows w * 64 R + g * 256
        // + 4 L + s of the tile (R = kSelMaskRows rows per lane, g < R / 4, s = 0..3), so a wave's
        // load instruction carries 1 KiB. Its ballots are stored as they are (bit L of word
        // (w, g, s) = w R + 4 g + s is row w * 64 R + g * 256 + 4 L + s) and bit 0 of tile_row[t] tells
        // the expansion to transpose them. The next full tile of this block and batch loads before
        // this tile's dependent lookups (without a branch: this tile's rows again when there is
        // none).
        constexpr int NG = kSelMaskRows / 4;
        const i64 tr_ = bc.tr;
        const i64 rb = tr.first + (i64)wave * (64 * kSelMaskRows) + 4 * lane;
        if (!have) {
#pragma unroll
          for (int g = 0; g < NG; ++g) ev.load_row4(rb + 256 * g, xa[g]);
        }
        const bool nxt = t + 1 < t1 && tr.first + 2 * tr_ <= bc.n;
        const i64 rn = nxt ? rb + tr_ : rb;
        typename Ev::Row4 xb[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) ev.load_row4(rn + 256 * g, xb[g]);
        bool p[kSelMaskRows];
#pragma unroll
        for (int j = 0; j < kSelMaskRows; ++j) p[j] = ev.where_row(ev.row_of4(xa[j >> 2], j & 3), rb + (j >> 2) * 256 + (j & 3));
#pragma unroll
        for (int g = 0; g < NG; ++g) xa[g] = xb[g];
        have = nxt;
        // lane j keeps ballot j: one store instruction per wave writes them all
        uint32_t tc = 0;
        u64 mine = 0;
#pragma unroll
        for (int j = 0; j < kSelMaskRows; ++j) {
          const u64 B = __ballot(p[j]);
          tc += (uint32_t)__popcll(B);
          if (lane == j) mine = B;
        }
        if (lane < kSelMaskRows) tm[wave * kSelMaskRows + lane] = mine;
        if (lane == 0 && tc) atomicAdd(&tile_cnt[t], tc);
        if (threadIdx.x == 0) tile_row[t] = ((bc.row0 + tr.first) << 1) | 1;
        continue;
      }
      have = false;
    }
    bool pass[kSelMaskRows];
    if (Ev::kPureWhere || !del) {  // (block-uniform) every row evaluated, loads issued together
      bool w[kSelMaskRows];
      if (tr.count == bc.tr) {  // (block-uniform) full tile: row = first + j * blockDim + thread
        const i64 r0 = tr.first + threadIdx.x;
#pragma unroll
        for (int j = 0; j < kSelMaskRows; ++j) w[j] = !ev.has_where() || ev.where(r0 + (i64)j * blockDim.x);
      } else {
        have = false;
#pragma unroll
        for (int j = 0; j < kSelMaskRows; ++j) {
          const i64 i = (i64)j * blockDim.x + threadIdx.x;
          w[j] = !ev.has_where() || ev.where(tr.first + (i < tr.count ? i : 0));
        }
      }
#pragma unroll
      for (int j = 0; j < kSelMaskRows; ++j) {
        const i64 i = (i64)j * blockDim.x + threadIdx.x;
        pass[j] = i < tr.count && w[j] && !is_deleted(del, tr.first + i);
      }
    } else {
      have = false;
#pragma unroll
      for (int j = 0; j < kSelMaskRows; ++j) {
        const i64 i = (i64)j * blockDim.x + threadIdx.x;
        const i64 r = tr.first + i;
        pass[j] = i < tr.count && !is_deleted(del, r) && (!ev.has_where() || ev.where(r));
      }
    }
    uint32_t tc = 0;
    u64 mine = 0;
#pragma unroll
    for (int j = 0; j < kSelMaskRows; ++j) {
      const u64 m = __ballot(pass[j]);
      if (lane == j) mine = m;
      tc += (uint32_t)__popcll(m);
    }
    if (lane < kSelMaskRows) tm[lane * nw + wave] = mine;
    if (threadIdx.x == 0) tile_row[t] = (bc.row0 + tr.first) << 1;
    if (lane == 0 && tc) atomicAdd(&tile_cnt[t], tc);  // (16 per tile address: no queue builds up)
  }
}

// Position of the k-th set bit of m (k < popcount(m)).
__device__ __forceinline__ int select_bit(u64 m, int k) {
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int c = __popcll(m & ((1ull << w) - 1));
    if (k >= c) {
      k -= c;
      m >>= w;
      pos += w;
    }
  }
  return pos;
}

// Dense GROUP BY whose WHERE reads large set bitmaps (a join filter fused into the scan, an
// IN-list over a wide key range): the ballot pass's structure, accumulating instead of storing
// ballots. One block of 1024 threads per CU holds d.lds_slots LDS copies of the accumulators (the
// dense_init_copies layout, copy = wave % copies) and, in the LDS left, the leading words of the
// plan's set bitmaps (d.set_words). Each block scans a contiguous range of tiles of kSelMaskRows
// rows per thread. On full tiles the WHERE's columns load four rows per lane, the next tile's
// while this one is evaluated. A wave's passing rows are then dealt out one per lane (the k-th
// passing row, in ballot order, to lane k); their keys and values load together, no load being
// conditional (a lane without a row reads the tile's first row), and reach the LDS copies one
// tile later, which folds the wait for them into the next tile's wait for its set lookups. Beyond
// 64 passing rows per wave and tile, the earlier rounds of 64 are added at once. Other tiles go
// row by row. Block partials are written as dense_body's.
template <class Ev>
__device__ void dense_staged_body(const ScanDesc& d, i64* parts, int* flags) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ ws_colview scols[WS_MAX_COLS];
  Ev ev(d);
  const ws_agg_plan& plan = *d.plan;
  const ws_dense_spec& ds = *d.ds;
  const int NK = ev.nkeys(), NA = ev.naggs();
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int C = d.lds_slots;
  dense_init_copies(d, ds, NA, ws_smem);
  ev.stage_sets((unsigned*)(ws_smem + (i64)C * d.copy_stride), d.set_words);
  // the dense spec and slot offsets in registers (reads of the descriptor in the loop would be
  // vector loads that wait behind the column loads)
  i64 kmin[Ev::kKeys], krange[Ev::kKeys], kstride[Ev::kKeys];
#pragma unroll
  for (int k = 0; k < Ev::kKeys; ++k) {
    kmin[k] = k < NK ? uni((i64)ds.kmin[k]) : 0;
    krange[k] = k < NK ? uni((i64)ds.range[k]) : 1;
    kstride[k] = k < NK ? uni((i64)ds.stride[k]) : 0;
  }
  int soff[Ev::kAggs];
#pragma unroll
  for (int a = 0; a < Ev::kAggs; ++a) soff[a] = a < NA ? uni((int)d.slot_off[a]) : 0;
  __syncthreads();
  unsigned char* mine = ws_smem + (i64)(wave % C) * d.copy_stride;
  bool bad = false;
  // A row's key and aggregate column words (loaded) and the row; added to the accumulators later,
  // so that nothing computed from the words makes the wave wait for them at the load.
  struct Pend {
    typename Ev::VRow x;
    i64 r;
    bool valid;
  };
  auto load = [&](i64 r, bool valid, Pend& q) {
    ev.load_vrow(r, q.x);
    q.r = r;
    q.valid = valid;
  };
  auto add = [&](const Pend& q) {
    if (!q.valid) return;
    i64 g = 0;
#pragma unroll
    for (int k = 0; k < Ev::kKeys; ++k) {
      if (k >= NK) break;
      const i64 o = ev.ikey_v(k, q.x, q.r) - kmin[k];
      if (o < 0 || o >= krange[k]) {
        bad = true;
        return;
      }
      g += o * kstride[k];
    }
#pragma unroll
    for (int a = 0; a < Ev::kAggs; ++a) {
      if (a >= NA) break;
      const uint8_t k = ev.dkind(a, ds);
      unsigned char* arr = mine + soff[a];
      if (k == WS_DK_COUNT) {
        __hip_atomic_fetch_add(&((unsigned*)arr)[g], 1u, __ATOMIC_RELAXED, WS_LDS);
      } else if (dk_int(k)) {
        atomic_combine_i<WS_LDS>(dk_op(k), &((i64*)arr)[g], int_part(ev, a, ev.iagg_v(a, q.x, q.r)));
      } else {
        atomic_combine_d<WS_LDS>(dk_op(k), &((double*)arr)[g], ev.agg_v(a, q.x, q.r));
      }
    }
  };
  Pend pend;  // the previous full tile's last round of passing rows: loaded, not yet added
  bool pending = false;
  const i64 t0 = (i64)blockIdx.x * d.ntiles / gridDim.x, t1 = (i64)(blockIdx.x + 1) * d.ntiles / gridDim.x;
  BatchCursor bc;
  bool have = false;
  // One tile: xc holds its WHERE columns when `have`; the next tile's load into xn. The loop below
  // alternates the two arrays (a copy from one to the other would wait for the loads).
  typedef typename Ev::Row4 Rows[kSelMaskRows / 4];
  auto tile = [&](i64 t, Rows& xc, Rows& xn) {
    if (bc.seek(d, t)) {
      have = false;
      stage_batch(d, bc.b, scols);
      ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
    }
    const TileRows tr = bc.rows(t);
    const uint32_t* del = bc.del;
    if constexpr (Ev::kSplit) {
      if (!del && tr.count == bc.tr && ev.rows4_ok()) {  // (block-uniform)
        // rows as in select_mask_body: lane L of wave w takes rows w * 64 R + g * 256 + 4 L + s
        constexpr int NG = kSelMaskRows / 4;
        const i64 rw = tr.first + (i64)wave * (64 * kSelMaskRows);
        const i64 rb = rw + 4 * lane;
        if (!have) {
#pragma unroll
          for (int g = 0; g < NG; ++g) ev.load_row4(rb + 256 * g, xc[g]);
        }
        const bool nxt = t + 1 < t1 && tr.first + 2 * bc.tr <= bc.n;
        const i64 rn = nxt ? rb + bc.tr : rb;
#pragma unroll
        for (int g = 0; g < NG; ++g) ev.load_row4(rn + 256 * g, xn[g]);
        u64 B[kSelMaskRows];
        int pre[kSelMaskRows + 1];
        pre[0] = 0;
#pragma unroll
        for (int j = 0; j < kSelMaskRows; ++j) {
          B[j] = __ballot(ev.where_row(ev.row_of4(xc[j >> 2], j & 3), rb + (j >> 2) * 256 + (j & 3)));
          pre[j + 1] = pre[j] + __popcll(B[j]);
        }
        have = nxt;
        if (pending) add(pend);  // (its loads were issued before this tile's lookups)
        const int total = pre[kSelMaskRows];
        // passing row k: ballot j with pre[j] <= k < pre[j + 1], then the (k - pre[j])-th set bit
        auto row_of = [&](int k) -> i64 {
          int j = 0;
#pragma unroll
          for (int q = 1; q < kSelMaskRows; ++q) j += k >= pre[q] ? 1 : 0;
          // B[j] and pre[j] by masks, not selects (a select chain becomes an indexed load from
          // scratch memory, whose wait would take the next tile's loads with it)
          u64 bj = 0;
          int pj = 0;
#pragma unroll
          for (int q = 0; q < kSelMaskRows; ++q) {
            const u64 m = 0ull - (u64)(j == q);
            bj |= B[q] & m;
            pj |= pre[q] & (int)m;
          }
          return rw + (j >> 2) * 256 + 4 * select_bit(bj, k - pj) + (j & 3);
        };
        // rounds of 64 before the last (more than 64 passing rows in the wave): added at once
        const int last = total > 0 ? (total - 1) & ~(kWave - 1) : 0;
        for (int base = 0; base < last; base += kWave) {
          Pend q;
          load(row_of(base + lane), true, q);
          add(q);
        }
        // the last round, straight-line (a loop or a branch around these loads would make the
        // compiler wait for every load in flight, the next tile's included): added next tile
        pending = total > 0;
        if (pending) {  // (wave-uniform)
          const int k = last + lane;
          const bool valid = k < total;
          const i64 r = row_of(valid ? k : total - 1);  // (every lane computes a row: no branch)
          load(valid ? r : tr.first, valid, pend);
        }
        return;
      }
      have = false;
    }
    if (pending) add(pend);
    pending = false;
    for (i64 i = threadIdx.x; i < tr.count; i += blockDim.x) {
      const i64 r = tr.first + i;
      if (is_deleted(del, r)) continue;
      if (ev.has_where() && !ev.where(r)) continue;
      Pend q;
      load(r, true, q);
      add(q);
    }
  };
  Rows xa, xb;
  for (i64 t = t0; t < t1; t += 2) {
    tile(t, xa, xb);
    if (t + 1 < t1) tile(t + 1, xb, xa);
  }
  if (pending) add(pend);
  if (bad) flags[1] = 1;
  __syncthreads();
  dense_write_copies(d, ds, NA, ws_smem, parts + (i64)blockIdx.x * ds.G * NA);
}

// Bitmap of the integer keys of the rows that pass the WHERE (the filtered build side of a join
// filter, without materializing it): bit key - lo of `bits` for each live passing row whose key is
// in [lo, lo + nbits); parts[block] (when given) gets the block's count of passing rows (summed
// afterwards: atomics from every block on one address would queue). Tiles grid-stride; a tile is
// kKeyBitsRows rows per thread, whose WHERE and key loads all issue before any bit is set when
// the plan has no side effects (rows past the tile re-read its first row), so a thread waits
// about once per tile instead of once per row.
constexpr int kKeyBitsRows = 8;
template <class Ev>
__device__ void key_bits_body(const ScanDesc& d, i64 lo, i64 nbits, uint32_t* bits, u64* parts) {
  __shared__ u64 s_cnt[16];
  __shared__ ws_colview scols[WS_MAX_COLS];
  Ev ev(d);
  u64 passed = 0;
  BatchCursor bc;
  const i64 ntiles = uni(d.ntiles);
  constexpr int U = kKeyBitsRows;
  for (i64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    if (bc.seek(d, t)) {
      stage_batch(d, bc.b, scols);
      ev.bind(Ev::kGlobalCols ? d.cols + (i64)bc.b * d.ncols : scols);
    }
    const TileRows tr = bc.rows(t);
    const uint32_t* del = bc.del;
    if (Ev::kPureKeys && Ev::kPureWhere && !del) {  // (block-uniform)
      for (i64 i0 = threadIdx.x; i0 < tr.count; i0 += (i64)blockDim.x * U) {
        bool w[U];
        i64 key[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const i64 i = i0 + (i64)u * blockDim.x;
          const i64 r = tr.first + (i < tr.count ? i : 0);
          w[u] = i < tr.count && (!ev.has_where() || ev.where(r));
          key[u] = ev.ikey(0, r);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!w[u]) continue;
          ++passed;
          const u64 o = (u64)(key[u] - lo);
          if (o < (u64)nbits) atomicOr(&bits[o >> 5], 1u << (o & 31));
        }
      }
      continue;
    }
    for (i64 i = threadIdx.x; i < tr.count; i += blockDim.x) {
      const i64 r = tr.first + i;
      if (is_deleted(del, r)) continue;
      if (ev.has_where() && !ev.where(r)) continue;
      ++passed;
      const u64 o = (u64)(ev.ikey(0, r) - lo);
      if (o < (u64)nbits) atomicOr(&bits[o >> 5], 1u << (o & 31));
    }
  }
  if (parts) {  // (block-uniform)
    for (int o = kWave / 2; o > 0; o >>= 1) passed += __shfl_down(passed, o, kWave);
    const int wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) s_cnt[wave] = passed;
    __syncthreads();
    if (threadIdx.x == 0) {
      u64 t = 0;
      for (int w = 0; w < nw; ++w) t += s_cnt[w];
      parts[blockIdx.x] = t;
    }
  }
}

// Per-batch counts from the scanned tile offsets (tiles never straddle batches); no atomics, which
// from every wave at a launch's end would queue on the few per-batch addresses.
// tile_off holds `stride` offsets per tile (the tile's own offset first).
__device__ void batch_counts_body(const ScanDesc& d, const uint32_t* tile_off, int stride, const u64* total, u64* per_batch) {
  for (int b = threadIdx.x; b < d.nb; b += blockDim.x) {
    const i64 a = d.tile0[b], z = d.tile0[b + 1];
    const u64 oa = a < d.ntiles ? tile_off[a * stride] : *total, oz = z < d.ntiles ? tile_off[z * stride] : *total;
    per_batch[b] = oz - oa;
  }
}

// Exclusive prefix of the per-tile counts, in place, by one block; *total gets their sum. Each
// thread takes a run of consecutive counts, so one block scan covers the whole array.
__device__ void scan_counts_body(uint32_t* cnt, i64 n, u64* total) {
  __shared__ uint32_t s_w[16];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  const i64 K = (n + blockDim.x - 1) / blockDim.x;
  const i64 i0 = threadIdx.x * K, i1 = i0 + K < n ? i0 + K : n;
  uint32_t run = 0;
  for (i64 i = i0; i < i1; ++i) run += cnt[i];
  uint32_t incl = run;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += y;
  }
  if (lane == kWave - 1) s_w[wave] = incl;
  __syncthreads();
  uint32_t pre = 0, sum = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wave) pre += s_w[w];
    sum += s_w[w];
  }
  pre += incl - run;
  for (i64 i = i0; i < i1; ++i) {
    const uint32_t c = cnt[i];
    cnt[i] = pre;
    pre += c;
  }
  if (threadIdx.x == 0) *total = sum;
}

// Bit i of x (x < 2^16) to bit 4 i.
__device__ __forceinline__ u64 spread4(u64 x) {
  x = (x | (x << 24)) & 0x000000FF000000FFull;
  x = (x | (x << 12)) & 0x000F000F000F000Full;
  x = (x | (x << 6)) & 0x0303030303030303ull;
  x = (x | (x << 3)) & 0x1111111111111111ull;
  return x;
}
// Expands select_mask_body's ballots of tile t = blockIdx.x (one thread per 64-row word,
// blockDim.x == tile_rows / 64 <= kSelExpandThreads). tile_row[t] is the tile's first global row
// times two, plus 1 when its words are raw ballots of four rows per lane (word 4 k + s, bit L =
// row 256 k + 4 L + s; row-ordered word 4 k + m is then bit 4 i + s = bit 16 m + i of word 4 k + s).
// The tile's output offset comes from the scanned counts (`stride` per tile), each thread's from a
// block scan of the words' bit counts. A sparse tile's threads write their rows directly; a dense
// tile's selected rows go to LDS in order first, and the block then writes its run of global
// indices with coalesced stores (each thread's own run would be one store per row, strided across
// the wave). Every load the block needs is issued before the first result is used.
constexpr int kSelExpandThreads = 128;
__device__ void select_expand_body(const u64* masks, const i64* tile_row, const uint32_t* tile_off, int stride,
                                   int32_t* out_idx) {
  __shared__ uint32_t s_w[kSelExpandThreads / kWave];
  __shared__ uint16_t s_pos[kSelExpandThreads * 64];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  const i64 t = blockIdx.x;
  const u64* b4 = masks + t * (i64)blockDim.x + (threadIdx.x & ~3);
  const u64 w0 = b4[0], w1 = b4[1], w2 = b4[2], w3 = b4[3];
  const i64 info = tile_row[t];
  const u64 base = tile_off[t * stride];
  const int q = threadIdx.x & 3;
  u64 m;
  if (info & 1) {  // (block-uniform)
    const int sh = 16 * q;
    m = spread4((w0 >> sh) & 0xFFFFull) | spread4((w1 >> sh) & 0xFFFFull) << 1 | spread4((w2 >> sh) & 0xFFFFull) << 2 |
        spread4((w3 >> sh) & 0xFFFFull) << 3;
  } else {
    m = q == 0 ? w0 : q == 1 ? w1 : q == 2 ? w2 : w3;
  }
  const uint32_t c = (uint32_t)__popcll(m);
  uint32_t incl = c;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += y;
  }
  if (lane == kWave - 1) s_w[wave] = incl;
  __syncthreads();
  uint32_t wbase = 0, count = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wave) wbase += s_w[w];
    count += s_w[w];
  }
  uint32_t k = wbase + incl - c;
  const uint32_t r0 = threadIdx.x * 64;
  const i64 g = info >> 1;
  int32_t* o = out_idx + base;
  if (count <= 2 * blockDim.x) {  // (block-uniform)
    while (m) {
      o[k++] = (int32_t)(g + r0 + (uint32_t)(__ffsll((unsigned long long)m) - 1));
      m &= m - 1;
    }
    return;
  }
  while (m) {
    s_pos[k++] = (uint16_t)(r0 + (uint32_t)(__ffsll((unsigned long long)m) - 1));
    m &= m - 1;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) o[i] = (int32_t)(g + s_pos[i]);
}

}  // namespace ws

// ------------------------------------------------------------------ kernel entry points
// The AOT build instantiates every kernel with InterpEval; a compiled plan instantiates only the one
// it needs with its generated evaluator (extern "C" so the module symbol has a fixed name).
#define WS_KERNEL_BATCH_DENSE(EV, NAME)                                                                     \
  extern "C" __global__ void __launch_bounds__(256) NAME(const ws::BatchQuery* qs, const ws::BatchTile* tiles,  \
                                                         long long ntiles, int lds_slots) {                  \
    ws::batch_dense_body<EV>(qs, tiles, ntiles, lds_slots);                                                 \
  }
#define WS_KERNEL_NOKEYS(EV, NAME)                                                                          \
  extern "C" __global__ void __launch_bounds__(256) NAME(const ws::ScanDesc* d, double* out,               \
                                                         ws::NoKeysPartial* parts, unsigned* done) {        \
    ws::nokeys_body<EV>(*d, out, parts, done);                                                               \
  }
#define WS_KERNEL_DENSE(EV, NAME)                                                                           \
  extern "C" __global__ void __launch_bounds__(256) NAME(const ws::ScanDesc* d, double* gacc, long long* parts, \
                                                         int* flags) {                                       \
    ws::dense_body<EV>(*d, gacc, parts, flags);                                                              \
  }                                                                                                          \
  WS_KERNEL_BATCH_DENSE(EV, NAME##_batch)
// Compiled plans only (the evaluator stages its sets): dense GROUP BY with the sets in LDS.
#define WS_KERNEL_DENSE_JIT(EV, NAME)                                                                       \
  WS_KERNEL_DENSE(EV, NAME)                                                                                  \
  extern "C" __global__ void __launch_bounds__(1024) NAME##_staged(const ws::ScanDesc* d, long long* parts,   \
                                                                   int* flags) {                             \
    ws::dense_staged_body<EV>(*d, parts, flags);                                                             \
  }
#define WS_KERNEL_KEY_BITS(EV, NAME)                                                                        \
  extern "C" __global__ void __launch_bounds__(256) NAME(const ws::ScanDesc* d, long long lo, long long nbits,   \
                                                         unsigned* bits, unsigned long long* parts) {         \
    ws::key_bits_body<EV>(*d, lo, nbits, bits, parts);                                                       \
  }
#define WS_KERNEL_GROUPED(EV, NAME)                                                                         \
  extern "C" __global__ void __launch_bounds__(256) NAME(const ws::ScanDesc* d, ws::GTable gt) {            \
    ws::grouped_body<EV>(*d, gt);                                                                            \
  }                                                                                                          \
  extern "C" __global__ void __launch_bounds__(1024) NAME##_pcount(const ws::ScanDesc* d, ws::GTable) {     \
    ws::part_count_body<EV>(*d);                                                                             \
  }                                                                                                          \
  extern "C" __global__ void __launch_bounds__(1024) NAME##_pscatter(const ws::ScanDesc* d, ws::GTable) {   \
    ws::part_scatter_body<EV>(*d);                                                                           \
  }
#define WS_KERNEL_SELECT(EV, NAME)                                                                          \
  extern "C" __global__ void __launch_bounds__(256) NAME(const ws::ScanDesc* d, unsigned long long* state,  \
                                                         unsigned* tile_ctr, int* out_idx,                  \
                                                         unsigned long long* per_batch,                     \
                                                         unsigned long long* total, int* err) {             \
    ws::select_body<EV>(*d, state, tile_ctr, out_idx, per_batch, total, err);                               \
  }                                                                                                         \
  extern "C" __global__ void __launch_bounds__(256) NAME##_chunk(const ws::ScanDesc* d, unsigned long long* state, \
                                                                 unsigned* chunk_ctr, int* out_idx,         \
                                                                 unsigned long long* per_batch,             \
                                                                 unsigned long long* total, int* err) {     \
    ws::select_chunk_body<EV>(*d, state, chunk_ctr, out_idx, per_batch, total, err);                         \
  }
#define WS_KERNEL_SELECT_JIT(EV, NAME) \
  WS_KERNEL_SELECT(EV, NAME) WS_KERNEL_SELECT_STAGED(EV, NAME) WS_KERNEL_KEY_BITS(EV, NAME##_keybits)
// Compiled plans only (the evaluator stages its sets): the ballot pass of the staged selection.
#define WS_KERNEL_SELECT_STAGED(EV, NAME)                                                                   \
  extern "C" __global__ void __launch_bounds__(1024) NAME##_staged(const ws::ScanDesc* d, unsigned long long* masks, \
                                                                   unsigned* tile_cnt, long long* tile_row) { \
    ws::select_mask_body<EV>(*d, masks, tile_cnt, tile_row);                                                 \
  }
