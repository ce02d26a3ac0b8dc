// GpuAggregator and row selection over batches (reference src/gpu.cu:393-755), planned for gfx950.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <utility>

#include "ws_codegen.h"
#include "ws_internal.h"

// Ahead-of-time kernels: the shared bodies with the bytecode interpreter.
WS_KERNEL_NOKEYS(ws::InterpEval, ws_i_nokeys)
WS_KERNEL_DENSE(ws::InterpEval, ws_i_dense)  // also ws_i_dense_batch
WS_KERNEL_GROUPED(ws::InterpEval, ws_i_grouped)
WS_KERNEL_SELECT(ws::InterpEval, ws_i_select)
WS_KERNEL_NOKEYS(ws::InterpEvalLean, ws_l_nokeys)
WS_KERNEL_DENSE(ws::InterpEvalLean, ws_l_dense)
WS_KERNEL_GROUPED(ws::InterpEvalLean, ws_l_grouped)
WS_KERNEL_SELECT(ws::InterpEvalLean, ws_l_select)
WS_KERNEL_NOKEYS(ws::InterpEvalCols, ws_c_nokeys)
WS_KERNEL_DENSE(ws::InterpEvalCols, ws_c_dense)
WS_KERNEL_GROUPED(ws::InterpEvalCols, ws_c_grouped)
WS_KERNEL_SELECT(ws::InterpEvalCols, ws_c_select)
WS_KERNEL_KEY_BITS(ws::InterpEval, ws_i_key_bits)
WS_KERNEL_KEY_BITS(ws::InterpEvalLean, ws_l_key_bits)
WS_KERNEL_KEY_BITS(ws::InterpEvalCols, ws_c_key_bits)

// Interpreter variant of a plan: the full VM, the lean VM (no scalar-function opcodes), or the
// column-only evaluator (every program one column or literal). IV_PICK names the kernel.
enum { kIvFull = 0, kIvLean = 1, kIvCols = 2 };
#define IV_PICK(iv, stem) ((iv) == kIvCols ? ws_c_##stem : (iv) == kIvLean ? ws_l_##stem : ws_i_##stem)

namespace ws {

constexpr int kBlock = 256;
constexpr int kAggTileRows = 4096;
constexpr int kSelTileRows = kSelRows * kBlock;
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr size_t kStaticLds = 2 * 1024 + 512;  // staged column views + small shared state

// ------------------------------------------------------------------ small kernels

__global__ void __launch_bounds__(1024) scan_counts_k(uint32_t* cnt, int64_t n, u64* total) { scan_counts_body(cnt, n, total); }
__global__ void batch_counts_k(const ScanDesc* d, const uint32_t* tile_off, int stride, const u64* total, u64* per_batch) {
  batch_counts_body(*d, tile_off, stride, total, per_batch);
}
__global__ void __launch_bounds__(kSelExpandThreads) select_expand_k(const u64* masks, const int64_t* tile_row,
                                                                      const uint32_t* tile_off, int stride, int32_t* out_idx) {
  select_expand_body(masks, (const i64*)tile_row, tile_off, stride, out_idx);
}
// *out = sum of a[0, n): one block.
__global__ void __launch_bounds__(1024) sum_u64_k(const u64* a, int64_t n, u64* out) {
  __shared__ u64 s_w[16];
  u64 t = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) t += a[i];
  for (int o = kWave / 2; o > 0; o >>= 1) t += __shfl_down(t, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 s = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += s_w[w];
    *out = s;
  }
}
__global__ void init_table_k(GTable t, int nk, int na, ws_agg_plan plan) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= (int64_t)t.mask; i += (int64_t)gridDim.x * blockDim.x) {
    t.keys[i] = kEmpty;
    for (int a = 0; a < na; ++a) t.acc[i * na + a] = slot_identity(plan.agg_bits[a] != 0, plan.ops[a]);
  }
}
__global__ void __launch_bounds__(256) batch_init_k(const BatchQuery* qs) { batch_init_body(qs); }
// An aggregator's initial values in one launch: flags and group counter cleared, dense
// accumulators (acc may be null) at their slots' identities.
__global__ void agg_prepare_k(uint32_t* flags, u64* count, double* acc, int64_t G, int na, ws_agg_plan plan) {
  const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i0 < 4) flags[i0] = 0;
  if (i0 == 0) *count = 0;
  if (!acc) return;
  for (int64_t i = i0; i < G * na; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(i % na);
    acc[i] = slot_identity(plan.agg_bits[a] != 0, plan.ops[a]);
  }
}
__global__ void __launch_bounds__(kReduceWaves * kWave) dense_reduce_cols_k(ws_agg_plan plan, ws_dense_spec ds, const i64* parts,
                                                                         int nparts, double* gacc) {
  dense_reduce_cols_body(plan, ds, parts, nparts, gacc);
}
__global__ void __launch_bounds__(256) dense_reduce_slot_k(ws_agg_plan plan, ws_dense_spec ds, const i64* parts, int nparts,
                                                           double* gacc) {
  dense_reduce_slot_body(plan, ds, parts, nparts, gacc);
}
// Folds nparts block partials ([part][G * NA]) into the accumulators, in block order: a block per
// slot when the slots are few (all of a slot's partials load at once), else (slot, partial) tiles.
static void dense_reduce(const ws_agg_plan& plan, const ws_dense_spec& ds, const i64* parts, int64_t nparts, double* gacc) {
  const int64_t S = ds.G * plan.naggs;
  if (S <= 0 || nparts <= 0) return;
  if (S <= 2048 && nparts >= 64)
    hipLaunchKernelGGL(dense_reduce_slot_k, dim3((unsigned)S), dim3(256), 0, stream(), plan, ds, parts, (int)nparts, gacc);
  else
    hipLaunchKernelGGL(dense_reduce_cols_k, dim3((unsigned)((S + kWave - 1) / kWave)), dim3(kReduceWaves * kWave), 0, stream(),
                       plan, ds, parts, (int)nparts, gacc);
  check_launch("agg_dense_reduce");
}
// Compacted groups, row-major (fetch) or column-major (compact): rows [off, off + g) of a result
// of G groups, taken from slots idx[i] of a row-major table (idx null: slot i).
__global__ void gather_groups_k(const double* kv, int nk, const double* acc, int na, const int32_t* idx, int64_t g,
                                double* keys_out, double* accs_out, int colmajor, int64_t off, int64_t G) {
  const int64_t total = g * (nk + na);
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = j / (nk + na);
    const int c = (int)(j - i * (nk + na));
    const int64_t s = idx ? (int64_t)idx[i] : i;
    const int64_t o = off + i;
    if (c < nk) {
      const double v = kv[s * nk + c];
      if (colmajor) keys_out[(int64_t)c * G + o] = v;
      else keys_out[o * nk + c] = v;
    } else {
      const int a = c - nk;
      const double v = acc[s * na + a];
      if (colmajor) accs_out[(int64_t)a * G + o] = v;
      else accs_out[o * na + a] = v;
    }
  }
}

// ------------------------------------------------------------------ batched dense fetch

struct FetchQuery {
  const double* acc;    // [G][NA]
  ws_dense_spec ds;     // keys are rebuilt from the group index
  int64_t G;
};
// A batched fetch works in chunks of kFetchChunk groups (many blocks per large query, so the
// largest query does not set the time): chunk j covers groups [g0, g0 + kFetchChunk) of query q.
constexpr int64_t kFetchChunk = 4096;
struct FetchChunk {
  int32_t q, pad;
  int64_t g0;
};
// Chunks of queries with G[0..n) groups, query after query; cb[q] = first chunk of query q.
static std::vector<FetchChunk> fetch_chunks(const std::vector<FetchQuery>& fq, std::vector<int64_t>& cb) {
  std::vector<FetchChunk> ch;
  cb.assign(fq.size() + 1, 0);
  for (size_t q = 0; q < fq.size(); ++q) {
    cb[q] = (int64_t)ch.size();
    for (int64_t g0 = 0; g0 < fq[q].G; g0 += kFetchChunk) ch.push_back(FetchChunk{(int32_t)q, 0, g0});
  }
  cb[fq.size()] = (int64_t)ch.size();
  return ch;
}
// Block j: number of groups of chunk j with rows (slot 0 = COUNT(*) > 0).
__global__ void __launch_bounds__(256) dense_count_many_k(const FetchQuery* qs, const FetchChunk* ch, int na, uint32_t* counts) {
  __shared__ uint32_t s_w[4];
  const FetchChunk c = ch[blockIdx.x];
  const FetchQuery& Q = qs[c.q];
  const int64_t g1 = c.g0 + kFetchChunk < Q.G ? c.g0 + kFetchChunk : Q.G;
  uint32_t n = 0;
#pragma unroll 16
  for (int64_t g = c.g0 + threadIdx.x; g < g1; g += blockDim.x) n += Q.acc[g * na] > 0;
  for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}
// Block j: chunk j's groups with rows, ascending group index, row-major keys | accs from row
// offs[j]. Every thread loads the count slots of its groups of the chunk first (their loads in
// flight together), then the chunk is ranked 256 groups at a time.
__global__ void __launch_bounds__(256) dense_pack_many_k(const FetchQuery* qs, const FetchChunk* ch, ws_agg_plan plan, int nk,
                                                         int na, const uint32_t* offs, double* out_k, double* out_a) {
  constexpr int kPer = (int)(kFetchChunk / 256);
  __shared__ uint32_t s_w[4];
  const FetchChunk c = ch[blockIdx.x];
  const FetchQuery& Q = qs[c.q];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  bool has[kPer];
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int64_t g = c.g0 + it * 256 + threadIdx.x;
    has[it] = g < Q.G && Q.acc[g * na] > 0;
  }
  uint32_t base = offs[blockIdx.x];
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int64_t g = c.g0 + it * 256 + threadIdx.x;
    const u64 m = __ballot(has[it]);
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1));
    __syncthreads();  // the previous step's s_w reads are done
    if (lane == 0) s_w[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int j = 0; j < 4; ++j) {
      if (j < wv) before += s_w[j];
      total += s_w[j];
    }
    if (has[it]) {
      const int64_t i = (int64_t)(base + before + rank);
      for (int k = 0; k < nk; ++k) {
        const i64 v = Q.ds.kmin[k] + (g / Q.ds.stride[k]) % Q.ds.range[k];
        out_k[i * nk + k] = plan.key_int[k] ? bitsd(v) : (double)v;
      }
      for (int a = 0; a < na; ++a) out_a[i * na + a] = Q.acc[g * na + a];
    }
    base += total;
  }
}

// ------------------------------------------------------------------ partitioned GROUP BY (host-side passes)

// part_off[c][p] = sum of part_cnt[c'][p] over c' < c, ptot[p] = sum over all chunks. A block covers
// 64 partitions (one lane each) with 16 chunk ranges (one wave each); reads are coalesced over p.
__global__ void __launch_bounds__(1024) part_offsets_k(const uint32_t* cnt, uint32_t* off, uint32_t* ptot, int64_t nchunks,
                                                       int P) {
  __shared__ uint32_t s_part[16][64];
  const int pl = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + pl;
  const int64_t per = (nchunks + 15) / 16;
  const int64_t c0 = g * per, c1 = c0 + per < nchunks ? c0 + per : nchunks;
  uint32_t sum = 0;
  if (p < P)
    for (int64_t c = c0; c < c1; ++c) sum += cnt[c * P + p];
  s_part[g][pl] = sum;
  __syncthreads();
  uint32_t base = 0;
  for (int j = 0; j < g; ++j) base += s_part[j][pl];
  if (p < P) {
    for (int64_t c = c0; c < c1; ++c) {
      off[c * P + p] = base;
      base += cnt[c * P + p];
    }
    if (g == 15) ptot[p] = base;
  }
}

// Chooses the packed record layout from the integer keys' ranges (thread 0 of part_start_k;
// range = [min, max] per key over the whole input).
__device__ void pack_plan(const long long* range, int nk, PackPlan* pk) {
  PackPlan q{};
  int bits = 0;
  bool fits = nk <= kPackKeys;
  for (int k = 0; k < nk && fits; ++k) {
    const i64 lo = range[k], hi = range[kPackKeys + k];
    int w = 0;
    if (lo <= hi) {
      const u64 span = (u64)hi - (u64)lo;  // hi - lo without overflow
      w = span ? 64 - __clzll((long long)span) : 0;
    }
    q.shift[k] = bits;
    q.mask[k] = w >= 64 ? ~0ull : (1ull << w) - 1;
    q.kmin[k] = lo <= hi ? lo : 0;
    bits += w;
    fits = bits <= 63;
  }
  q.ok = fits ? 1 : 0;
  *pk = q;
}

// pstart[p] = exclusive prefix of ptot over partitions, pstart[P] = total (one block, P <= 4096).
// With a pack plan to fill, the block also folds the count pass's per-chunk key ranges
// ([nchunks][2 * nk]) and thread 0 chooses the record layout (the scatter pass, next, reads it).
__global__ void __launch_bounds__(1024) part_start_k(const uint32_t* ptot, uint32_t* pstart, int P,
                                                     const int64_t* key_range = nullptr, int64_t nchunks = 0, int nk = 0,
                                                     PackPlan* pack = nullptr) {
  __shared__ long long s_rng[2 * kPackKeys];
  if (pack) {
    if (threadIdx.x < 2 * kPackKeys) s_rng[threadIdx.x] = threadIdx.x < kPackKeys ? kI64Max : kI64Min;
    __syncthreads();
    for (int k = 0; k < nk && k < kPackKeys; ++k) {
      i64 lo = kI64Max, hi = kI64Min;
      for (int64_t c = threadIdx.x; c < nchunks; c += blockDim.x) {
        const i64 a = key_range[c * 2 * nk + k], b = key_range[c * 2 * nk + nk + k];
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
      }
      for (int o = 32; o > 0; o >>= 1) {
        const i64 a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
      }
      if ((threadIdx.x & 63) == 0) {
        atomicMin(&s_rng[k], (long long)lo);
        atomicMax(&s_rng[kPackKeys + k], (long long)hi);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) pack_plan(s_rng, nk, pack);
  }
  __shared__ uint32_t s_w[16];
  const int per = (P + 1023) / 1024;
  const int i0 = threadIdx.x * per;
  uint32_t n = 0;
  for (int i = i0; i < i0 + per && i < P; ++i) n += ptot[i];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t incl = n;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) s_w[wv] = incl;
  __syncthreads();
  uint32_t base = 0, total = 0;
  for (int j = 0; j < 16; ++j) {
    if (j < wv) base += s_w[j];
    total += s_w[j];
  }
  uint32_t run = base + incl - n;
  for (int i = i0; i < i0 + per && i < P; ++i) {
    pstart[i] = run;
    run += ptot[i];
  }
  if (threadIdx.x == 0) pstart[P] = total;
}

__device__ __forceinline__ double wave_combine_slot(bool bits, int op, double v) {
  return bits ? bitsd(wave_combine_i(op, dbits(v))) : wave_combine_d(op, v);
}

struct PartMerge {
  const u64* recs;          // [records][W]
  const uint32_t* pstart;   // [P + 1]
  int W, NK, NA, S;         // record words (full layout), keys, slots, LDS table slots
  u64 kinds;                // slot kinds (slot_kind)
  double* okv;              // [records][NK] finished groups: partition p's at rows pstart[p] + rank
  double* oacc;             // [records][NA]
  uint32_t* gcnt;           // [P] groups finished per partition
  const PackPlan* pack;     // the layout the scatter pass used (null: full records)
};

// Row of the partitioned result that holds group i of the compacted order (groups of partition p
// are i in [goff[p], goff[p + 1]), stored from row pstart[p]).
__device__ __forceinline__ int64_t part_row(const uint32_t* goff, const uint32_t* pstart, int P, int64_t i) {
  int lo = 0, hi = P - 1;  // last p with goff[p] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)goff[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return (int64_t)pstart[lo] + (i - (int64_t)goff[lo]);
}

// Compacted groups of a partitioned result: rows [0, g) of a result of G groups. Block b copies
// partitions b, b + gridDim.x, ...: partition p's groups are rows pstart[p] + [0, goff[p + 1] -
// goff[p]) and become compacted groups goff[p] + rank (contiguous reads and writes, no search).
__global__ void gather_part_k(const double* kv, int nk, const double* acc, int na, const uint32_t* goff,
                              const uint32_t* pstart, int P, int64_t g, double* keys_out, double* accs_out, int colmajor,
                              int64_t G) {
  for (int p = blockIdx.x; p < P; p += gridDim.x) {
    const int64_t o0 = goff[p], cnt = (int64_t)goff[p + 1] - o0, s0 = pstart[p];
    for (int64_t j = threadIdx.x; j < cnt; j += blockDim.x) {
      const int64_t s = s0 + j, i = o0 + j;
      for (int c = 0; c < nk; ++c) {
        const double v = kv[s * nk + c];
        if (colmajor) keys_out[(int64_t)c * G + i] = v;
        else keys_out[i * nk + c] = v;
      }
      for (int a = 0; a < na; ++a) {
        const double v = acc[s * na + a];
        if (colmajor) accs_out[(int64_t)a * G + i] = v;
        else accs_out[i * na + a] = v;
      }
    }
  }
}

// Slot kinds of a plan packed 3 bits per slot (op in bits 0-1, int64-bits flag in bit 2), so
// kernels branch on a uniform scalar instead of loading the plan per record.
__host__ __device__ __forceinline__ int slot_kind(u64 kinds, int a) { return (int)((kinds >> (3 * a)) & 7); }
__device__ __forceinline__ double kind_identity(int k) { return slot_identity((k & 4) != 0, k & 3); }
template <int SCOPE>
__device__ __forceinline__ void atomic_combine_kind(int k, double* addr, double v) {
  atomic_combine_slot<SCOPE>((k & 4) != 0, k & 3, addr, v);
}
__device__ __forceinline__ double wave_combine_kind(int k, double v) {
  return (k & 4) ? bitsd(wave_combine_i(k & 3, dbits(v))) : wave_combine_d(k & 3, v);
}
// Additive slots only (every op COUNT or SUM, bit 1 of each kind clear): the combine is an add of
// int64 bits or a NULL-skipping add of doubles, so the merge's per-record code needs no dispatch
// on the op.
__host__ __device__ __forceinline__ bool kinds_additive(u64 kinds, int na) {
  for (int a = 0; a < na; ++a)
    if (slot_kind(kinds, a) & 2) return false;
  return true;
}
template <bool AS>
__device__ __forceinline__ double kind_identity_t(int k) { return AS ? 0.0 : kind_identity(k); }
template <bool AS, int SCOPE>
__device__ __forceinline__ void atomic_combine_kind_t(int k, double* addr, double v) {
  if (!AS) return atomic_combine_kind<SCOPE>(k, addr, v);
  if (k & 4) __hip_atomic_fetch_add((u64*)addr, (u64)dbits(v), __ATOMIC_RELAXED, SCOPE);
  else if (v == v) __hip_atomic_fetch_add(addr, v, __ATOMIC_RELAXED, SCOPE);
}
template <bool AS>
__device__ __forceinline__ double wave_combine_kind_t(int k, double v) {
  if (!AS) return wave_combine_kind(k, v);
  return (k & 4) ? bitsd(wave_sum_i64(dbits(v))) : wave_combine_d(WS_AGG_SUM, v);
}

// Inserts one group (hash, key bits, slot values in registers) into the spill table; false when
// the table overflowed.
template <int NA_MAX>
__device__ __forceinline__ bool spill_insert(const GTable& t, u64 kinds, int NK, int NA, u64 h, const u64* key,
                                             const u64 (&val)[NA_MAX], unsigned* new_groups) {
  u64 s = h & t.mask;
  for (int probe = 0; probe < 256; ++probe) {
    u64 prev = kEmpty;
    const bool claimed = __hip_atomic_compare_exchange_strong(&t.keys[s], &prev, h, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
    if (claimed) {
      for (int k = 0; k < NK; ++k) t.kv[s * NK + k] = bitsd((i64)key[k]);
      __hip_atomic_fetch_add(new_groups, 1u, __ATOMIC_RELAXED, WS_LDS);
    }
    if (claimed || prev == h) {
#pragma unroll
      for (int a = 0; a < NA_MAX; ++a)
        if (a < NA) atomic_combine_kind<WS_DEV>(slot_kind(kinds, a), &t.acc[s * NA + a], bitsd((i64)val[a]));
      return true;
    }
    s = (s + 1) & t.mask;
  }
  t.flags[0] = 1;
  return false;
}

// Key hash of the unpacked keys of a packed record: the formula of row_keys over integer keys.
__device__ __forceinline__ u64 packed_keys(const PackPlan& pk, int NK, u64 w, u64* key) {
  u64 h = 0x9e3779b97f4a7c15ull;
#pragma unroll
  for (int k = 0; k < kPackKeys; ++k) {
    if (k >= NK) break;
    key[k] = (u64)((i64)((w >> pk.shift[k]) & pk.mask[k]) + pk.kmin[k]);
    h = mix64(h ^ key[k]) + (u64)k;
  }
  return h == kEmpty ? h ^ 1 : h;
}

// Merge pass, for records of W words: block p combines partition p's records in an LDS table of
// m.S slots (linear probing from a multiply-shift index; the partition is the hash's high bits).
// Full records are [hash, key values, slot values] and the table keeps the hash and the key
// values. Packed ones (PK) are [packed keys, slot values]: the packed word identifies the group
// exactly, so the table keeps only that word, indexes by a multiplicative hash of it, and the
// canonical hash (the plain path's, from the unpacked keys) is computed only for a record that
// spills to the global table. Each thread loads R records into registers before combining them
// (the step is unrolled), so a block keeps many loads in flight.
// Lanes of a wave holding the same group as the wave's first live lane fold into that lane first:
// a hot group costs one LDS update per wave instead of one per lane. A record that finds no LDS
// slot in 64 probes goes to the spill table (device-scope atomics, as the plain path). At the end
// the block writes its groups inside the partition's own record range and its group count to
// gcnt[p]; no block waits on another or on a shared counter. Launched for both layouts; the one
// that does not match the scatter pass's choice (m.pack) returns at once.
template <int W, bool PK, bool AS = false>
__global__ void __launch_bounds__(512) part_merge_k(PartMerge m, GTable spill) {
  constexpr int R = W <= 6 ? 4 : W <= 10 ? 2 : 1;  // records per thread per step (VGPRs: 3 blocks per CU)
  const bool packed = m.pack && m.pack->ok;
  if (packed != PK) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char ws_smem[];
  __shared__ unsigned s_new;
  __shared__ uint32_t s_w[8];
  const int NK = m.NK, S = m.S;
  const int KW = PK ? 1 : NK;             // key words per record / LDS slot
  const int NA = W - 1 - (PK ? 0 : NK);   // slot values per record
  constexpr int VA = PK ? W - 1 : W - 2;  // register bound on slot values (NK >= 1)
  constexpr int KB = PK ? kPackKeys : (W > 1 ? W - 1 : 1);  // register bound on key words
  PackPlan pk{};
  if (PK) {
    pk = *m.pack;
#pragma unroll
    for (int k = 0; k < kPackKeys; ++k) {  // the same for every lane: scalar registers
      pk.shift[k] = uni(pk.shift[k]);
      pk.mask[k] = (u64)uni((i64)pk.mask[k]);
      pk.kmin[k] = uni((i64)pk.kmin[k]);
    }
  }
  const u64 kinds = m.kinds;
  u64* lk = (u64*)ws_smem;                                   // [S]: group identity
  u64* lkey = lk + S;                                        // [S][NK]: key bits (full records)
  double* lacc = (double*)(lkey + (PK ? 0 : (int64_t)S * KW));  // [S][NA]
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    lk[i] = kEmpty;
    for (int a = 0; a < NA; ++a) lacc[(int64_t)i * NA + a] = kind_identity_t<AS>(slot_kind(kinds, a));
  }
  if (threadIdx.x == 0) s_new = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t beg = uni((i64)m.pstart[blockIdx.x]), end = uni((i64)m.pstart[blockIdx.x + 1]);
  const u64 WS_GLOBAL* recs = (const u64 WS_GLOBAL*)m.recs;
  for (int64_t i0 = beg; i0 < end; i0 += (int64_t)blockDim.x * R) {
    u64 q[R][W];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      // A slice of the step that starts past the partition end loads nothing (block-uniform test);
      // inside a slice the loads are unguarded (the record buffer is padded past its last record).
      const int64_t i = i0 + (int64_t)r * blockDim.x + threadIdx.x;
      const bool slice = i0 + (int64_t)r * blockDim.x < end;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        u64 v = kEmpty;
        if (slice) v = recs[i * W + j];
        q[r][j] = i < end ? v : kEmpty;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (i0 + (int64_t)r * blockDim.x >= end) break;  // (block-uniform) the rest of the step is past the partition
      u64 rec[W];
#pragma unroll
      for (int j = 0; j < W; ++j) rec[j] = q[r][j];
      bool live = rec[0] != kEmpty;
      u64 key[KB] = {};  // key bits (copied for a spill)
      // Group identity: the packed key word (exact, below 2^63), or the hash of a full record.
      const u64 h = rec[0];
      u64 val[VA > 0 ? VA : 1];
#pragma unroll
      for (int a = 0; a < (VA > 0 ? VA : 1); ++a) {
        const int j = 1 + (PK ? 0 : NK) + a;
        u64 v = 0;
#pragma unroll
        for (int jj = 1; jj < W; ++jj)
          if (jj == j) v = rec[jj];
        val[a] = v;
      }
      const u64 lv = __ballot(live);
      if (lv) {
        const int L = __ffsll((unsigned long long)lv) - 1;
        const u64 hl = (u64)__shfl((long long)h, L, 64);
        const u64 same = __ballot(live && h == hl);
        if (__popcll(same) > 1) {
          const bool in = (same >> lane) & 1;
#pragma unroll
          for (int a = 0; a < (VA > 0 ? VA : 1); ++a) {
            if (a >= NA) break;
            const int k = slot_kind(kinds, a);
            const double v = wave_combine_kind_t<AS>(k, in ? bitsd((i64)val[a]) : kind_identity_t<AS>(k));
            if (lane == L) val[a] = (u64)dbits(v);
          }
          if (in && lane != L) live = false;
        }
      }
      if (!live) continue;
      // Slot: multiply-shift of the hash's low word; a packed word is first folded and multiplied
      // by a 32-bit odd constant (its high bits then spread the group over the table).
      const unsigned hx = PK ? ((unsigned)h ^ (unsigned)(h >> 32)) * 0x9e3779b1u : (unsigned)h;
      unsigned s = (unsigned)(((u64)hx * (u64)S) >> 32);
      int slot = -1;
      for (int probe = 0; probe < 64; ++probe) {
        u64 prev = kEmpty;
        const bool claimed = __hip_atomic_compare_exchange_strong(&lk[s], &prev, h, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
        if (claimed && !PK) {
#pragma unroll
          for (int j = 1; j < W; ++j)
            if (j <= NK) lkey[(int64_t)s * NK + (j - 1)] = rec[j];
        }
        if (claimed || prev == h) {
          slot = (int)s;
          break;
        }
        s = s + 1 == (unsigned)S ? 0 : s + 1;
      }
      if (slot >= 0) {
#pragma unroll
        for (int a = 0; a < (VA > 0 ? VA : 1); ++a)
          if (a < NA) atomic_combine_kind_t<AS, WS_LDS>(slot_kind(kinds, a), &lacc[(int64_t)slot * NA + a], bitsd((i64)val[a]));
      } else if (PK) {
        spill_insert(spill, kinds, NK, NA, packed_keys(pk, NK, h, key), key, val, &s_new);
      } else {
#pragma unroll
        for (int k = 0; k < KB; ++k)
          if (k < NK) key[k] = rec[1 + k < W ? 1 + k : 0];
        spill_insert(spill, kinds, NK, NA, h, key, val, &s_new);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_new) {
    const u64 total = atomicAdd((unsigned long long*)spill.count, (unsigned long long)s_new) + s_new;
    if (total * 10 > (spill.mask + 1) * 7) spill.flags[0] = 1;  // spill load above 0.7: retry bigger
  }
  const int K = (S + (int)blockDim.x - 1) / (int)blockDim.x;
  const int s0 = threadIdx.x * K, s1 = s0 + K < S ? s0 + K : S;
  uint32_t n = 0;
  for (int qi = s0; qi < s1; ++qi) n += lk[qi] != kEmpty;
  uint32_t incl = n;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 63) s_w[wv] = incl;
  __syncthreads();
  uint32_t wbase = 0, total = 0;
  for (int j = 0; j < nw; ++j) {
    if (j < wv) wbase += s_w[j];
    total += s_w[j];
  }
  // Groups stay inside the partition's own record range (a partition has at least as many
  // records as groups), so no block waits on another or on a shared counter.
  if (threadIdx.x == 0) m.gcnt[blockIdx.x] = total;
  u64 o = (u64)m.pstart[blockIdx.x] + wbase + incl - n;
  for (int qi = s0; qi < s1; ++qi) {
    if (lk[qi] == kEmpty) continue;
    if (PK) {
      u64 kb[kPackKeys];
      packed_keys(pk, NK, lk[qi], kb);
      for (int k = 0; k < NK; ++k) m.okv[o * NK + k] = bitsd((i64)kb[k]);
    } else {
      for (int k = 0; k < NK; ++k) m.okv[o * NK + k] = bitsd((i64)lkey[(int64_t)qi * NK + k]);
    }
    for (int a = 0; a < NA; ++a) m.oacc[o * NA + a] = lacc[(int64_t)qi * NA + a];
    ++o;
  }
}

// The merge kernels for records of W words: full layout (3..16 words) and packed (1..16).
typedef void (*MergeKernel)(PartMerge, GTable);
template <bool PK, bool AS, int Base, int... Ws>
static MergeKernel merge_pick(int W, std::integer_sequence<int, Ws...>) {
  MergeKernel k = nullptr;
  ((k = W == Base + Ws ? part_merge_k<Base + Ws, PK, AS> : k), ...);
  return k;
}
// Packed records with additive slots only (the common COUNT / SUM plans) get the variant without
// per-record op dispatch.
static MergeKernel merge_kernel(int W, bool packed, bool additive) {
  if (packed)
    return additive ? merge_pick<true, true, 1>(W, std::make_integer_sequence<int, 16>{})
                    : merge_pick<true, false, 1>(W, std::make_integer_sequence<int, 16>{});
  return merge_pick<false, false, 3>(W, std::make_integer_sequence<int, 14>{});
}

// Moves finished groups into a global table: rows of a partitioned result (goff set), or the
// occupied slots of a table (keys set). A partitioned result becomes a plain table this way when
// more batches follow.
__global__ void table_absorb_k(GTable dst, ws_agg_plan plan, const double* kv, const double* acc, const u64* keys,
                               const uint32_t* goff, const uint32_t* pstart, int P, int64_t n) {
  __shared__ unsigned s_new;
  if (threadIdx.x == 0) s_new = 0;
  __syncthreads();
  const int NK = plan.nkeys, NA = plan.naggs;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = goff ? part_row(goff, pstart, P, j) : j;
    if (keys && keys[i] == kEmpty) continue;
    u64 h = 0x9e3779b97f4a7c15ull;
    for (int k = 0; k < NK; ++k) h = mix64(h ^ (u64)dbits(kv[i * NK + k])) + (u64)k;
    if (h == kEmpty) h ^= 1;
    global_insert(dst, plan, NK, NA, keys ? keys[i] : h, &kv[i * NK], &acc[i * NA], &s_new);
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_new) {
    const u64 total = atomicAdd((unsigned long long*)dst.count, (unsigned long long)s_new) + s_new;
    if (total * 10 > (dst.mask + 1) * 7) dst.flags[0] = 1;
  }
}
// Dense groups: keys rebuilt from the group index (reference denseFinalize, gpu.cu:379-391).
__global__ void gather_dense_k(ws_dense_spec ds, ws_agg_plan plan, const double* acc, int nk, int na,
                               const int32_t* idx, int64_t g, double* keys_out, double* accs_out, int colmajor) {
  const int64_t total = g * (nk + na);
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = j / (nk + na);
    const int c = (int)(j - i * (nk + na));
    const int64_t s = idx[i];
    if (c < nk) {
      const i64 v = ds.kmin[c] + (s / ds.stride[c]) % ds.range[c];
      const double kvv = plan.key_int[c] ? bitsd(v) : (double)v;
      if (colmajor) keys_out[(int64_t)c * g + i] = kvv;
      else keys_out[i * nk + c] = kvv;
    } else {
      const int a = c - nk;
      const double v = acc[s * na + a];
      if (colmajor) accs_out[(int64_t)a * g + i] = v;
      else accs_out[i * na + a] = v;
    }
  }
}

// ------------------------------------------------------------------ slot compaction

struct PredOccupied {
  const u64* keys;
  __device__ bool operator()(i64 i) const { return keys[i] != kEmpty; }
};
struct PredDense {
  const double* acc;
  int na;
  __device__ bool operator()(i64 i) const { return acc[i * na] > 0; }
};
struct PredFlag {
  const uint8_t* f;
  __device__ bool operator()(i64 i) const { return f[i] != 0; }
};
template <class P>
__global__ void __launch_bounds__(256) compact_k(P pred, int64_t n, u64* state, unsigned* ctr, int32_t* out, u64* total,
                                                 int* err) {
  compact_body(pred, n, state, ctr, out, total, err);
}

template <class P>
static int64_t compact_run(const P& pred, int64_t n, int32_t* out_idx, const char* name) {
  if (n <= 0) return 0;
  const int64_t tile = kSelTileRows;
  const int64_t ntiles = (n + tile - 1) / tile;
  DevBuf<u64> st(ntiles + 4);  // tile states, then [counter, total, err]
  fill_u64(st.p, ntiles + 4, 0);
  {
    ProfScope ps(name);
    hipLaunchKernelGGL(compact_k<P>, dim3((unsigned)ntiles), dim3(kBlock), 0, stream(), pred, n, st.p, (unsigned*)(st.p + ntiles),
                       out_idx, st.p + ntiles + 1, (int*)(st.p + ntiles + 2));
    check_launch(name);
  }
  u64 res[2];
  d2h_sync(res, st.p + ntiles + 1, sizeof res);
  if (res[1]) throw Error("compaction look-back did not complete", WS_ERR_INTERNAL);
  return (int64_t)res[0];
}
int64_t compact_occupied_hash(const u64* keys, int64_t n, int32_t* out_idx) {
  return compact_run(PredOccupied{keys}, n, out_idx, "compact_groups");
}
int64_t compact_dense(const double* acc, int naggs, int64_t G, int32_t* out_idx) {
  return compact_run(PredDense{acc, naggs}, G, out_idx, "compact_groups");
}
int64_t compact_flags(const uint8_t* flags, int64_t n, int32_t* out_idx) {
  return compact_run(PredFlag{flags}, n, out_idx, "compact_flags");
}

// ------------------------------------------------------------------ launch descriptors

// Uploads the ScanDesc with its tables (and copies of plan / dense spec / env) in one blob.
struct DescUpload {
  uint8_t* dev = nullptr;
  ScanDesc* ddesc = nullptr;
  ~DescUpload() {
    if (dev) tmp_free(dev);
  }
};
static void build_desc(DescUpload& up, ScanDesc d, const ScanInput& in, const ws_agg_plan& plan,
                       const ws_dense_spec* ds, const ws_expr_env& env, bool cols_from_env, size_t extra_zero_bytes,
                       size_t* extra_off) {
  Blob b;
  const size_t o_desc = b.reserve(sizeof(ScanDesc));
  const size_t o_plan = b.put(&plan, sizeof plan);
  const size_t o_ds = ds ? b.put(ds, sizeof *ds) : 0;
  const size_t o_env = b.put(&env, sizeof env);
  const size_t o_n = b.put(in.n.data(), 8 * in.n.size());
  const size_t o_del = b.put(in.del.data(), 8 * in.del.size());
  std::vector<int64_t> tile0(in.nb + 1, 0), row0(in.nb, 0);
  for (int i = 0; i < in.nb; ++i) {
    tile0[i + 1] = tile0[i] + (in.n[i] + d.tile_rows - 1) / d.tile_rows;
    if (i + 1 < in.nb) row0[i + 1] = row0[i] + in.n[i];
  }
  const size_t o_t0 = b.put(tile0.data(), 8 * tile0.size());
  const size_t o_r0 = b.put(row0.data(), 8 * row0.size());
  const size_t o_cols = cols_from_env ? 0 : b.put(in.cols.data(), sizeof(ws_colview) * in.cols.size());
  const size_t o_extra = extra_zero_bytes ? b.reserve(extra_zero_bytes) : 0;
  up.dev = (uint8_t*)tmp_alloc(b.host.size());
  d.nb = in.nb;
  d.ncols = in.ncols;
  d.ntiles = tile0[in.nb];
  d.n = (const int64_t*)(up.dev + o_n);
  d.del = (const uint32_t* const*)(up.dev + o_del);
  d.tile0 = (const int64_t*)(up.dev + o_t0);
  d.row0 = (const int64_t*)(up.dev + o_r0);
  d.env = (const ws_expr_env*)(up.dev + o_env);
  d.plan = (const ws_agg_plan*)(up.dev + o_plan);
  d.ds = ds ? (const ws_dense_spec*)(up.dev + o_ds) : nullptr;
  d.cols = cols_from_env ? (const ws_colview*)(up.dev + o_env + offsetof(ws_expr_env, cols))
                         : (const ws_colview*)(up.dev + o_cols);
  *b.at<ScanDesc>(o_desc) = d;
  up.ddesc = (ScanDesc*)(up.dev + o_desc);
  if (extra_off) *extra_off = o_extra;
  h2d_async_staged(up.dev, b.host.data(), b.host.size());
}

// ------------------------------------------------------------------ aggregator

struct DensePlan {
  int mode = 0;  // 0 global, 1 private, 2 LDS copies
  int copies = 0;
  int copy_stride = 0;
  size_t lds = 0;
  int32_t slot_off[WS_MAX_AGGS] = {};
};

// Physical layout of a dense aggregation on gfx950 (160 KiB LDS per CU). The engine's DenseSpec
// fixes the groups and accumulator kinds; copies / private mode are re-derived here for this LDS
// size (they only change where partial sums live, not the results beyond FP summation order).
static DensePlan plan_dense(const ws_dense_spec& ds, int na) {
  DensePlan p;
  const int64_t S = ds.G * na;
  if (S * 8 * kBlock <= 32 * 1024) {  // tiny: every thread owns all slots, no atomics per row
    p.mode = 1;
    p.lds = (size_t)S * 8 * kBlock;
    return p;
  }
  size_t off = 0;
  for (int a = 0; a < na; ++a) {
    p.slot_off[a] = (int32_t)off;
    const size_t bytes = (size_t)ds.G * (ds.kind[a] == WS_DK_COUNT ? 4 : 8);
    off += (bytes + 15) / 16 * 16;
  }
  if (off <= 60 * 1024) {
    // One copy per wave when four fit in 48 KiB, else as many as fit (>= 1).
    int c = (int)std::min<size_t>(4, std::max<size_t>(1, (48 * 1024) / off));
    if (off * c > 60 * 1024) c = 1;
    p.mode = 2;
    p.copies = c;
    p.copy_stride = (int)off;
    p.lds = off * c;
    return p;
  }
  p.mode = 0;
  return p;
}

}  // namespace ws

using namespace ws;

struct ws_agg {
  ws_agg_plan plan{};
  ws_dense_spec dense{};
  bool is_dense = false;
  int64_t cap = 1;             // capacity the caller sized (slots of the plain table)
  int64_t tcap = 0;            // slots of the allocated hash table (cap, or the spill table)
  u64* keys = nullptr;         // hash table, allocated at the first run that needs it
  double* kv = nullptr;
  double* acc = nullptr;       // hash table accumulators, or the dense groups
  int* flags = nullptr;        // [0] overflow, [1] dense key out of range, [2] partition run overflow
  u64* count = nullptr;        // groups inserted into the hash table
  double* nk_out = nullptr;    // no GROUP BY: [2 * WS_MAX_AGGS] running result
  std::vector<int64_t> hi;
  int ncols = 0;
  std::vector<int64_t> bn;
  std::vector<const uint32_t*> bdel;
  std::vector<ws_colview> bcols;
  int64_t rows_done = 0;
  int batches_done = 0, windows = 0, runs = 0;
  // strategy: WS_STRAT_AUTO / WS_STRAT_TABLE / WS_STRAT_PARTITIONED, and the expected group count
  int strategy = WS_STRAT_AUTO;
  int64_t expected_groups = 0;
  float features[WS_SIZER_INPUTS] = {};     // sizing features; logged with the exact group count
  bool has_features = false, logged = false;
  ws_expr_env last_env{};                  // the last run's environment and plan (for a rerun)
  const ws_jit_program* last_jit = nullptr;
  // partitioned result: finished groups row-major in okv / oacc (plus any spilled to the table)
  bool part = false, settled = true;
  int64_t part_groups = 0;
  int P = 0;                    // partitions of the partitioned result
  double* okv = nullptr;        // [records][nk] groups of partition p from row pstart[p]
  double* oacc = nullptr;       // [records][na]
  uint32_t* pstart = nullptr;   // [P + 1] first record row of each partition
  uint32_t* goff = nullptr;     // [P + 1] first compacted group of each partition; goff[P] = groups
  // flags, table count and partitioned group count as last read back (one round trip); valid
  // until the next launch that can change them
  struct Status {
    int flags[4];
    u64 count;
    uint32_t groups;
  } st{};
  bool st_valid = false;
  // small dense aggregations: the accumulators read back with the status (host_acc_ok while
  // st_valid), so a fetch needs no further round trip
  std::vector<double> host_acc;
  bool host_acc_ok = false;
  // fetch state
  int64_t fetched = -1;
  DevBuf<double> fetch_buf;
  std::vector<double> host_keys, host_accs;  // a fetch served from host_acc
  bool fetched_host = false;
  // the flags, counter and dense accumulators still need their initial values (deferred from
  // creation so a batch of aggregations is initialized by one launch)
  bool pending_init = false;
  void free_part() {
    for (void* p : {(void*)okv, (void*)oacc, (void*)pstart, (void*)goff}) tmp_free(p);
    okv = oacc = nullptr;
    pstart = goff = nullptr;
    P = 0;
    part = false;
    settled = true;
    part_groups = 0;
  }
  void free_table() {
    for (void* p : {(void*)keys, (void*)kv}) tmp_free(p);
    if (!is_dense) tmp_free(acc), acc = nullptr;
    keys = nullptr, kv = nullptr, tcap = 0;
  }
  ~ws_agg() {
    free_part();
    free_table();
    for (void* p : {(void*)acc, (void*)flags, (void*)count, (void*)nk_out}) tmp_free(p);
  }
};

// Allocates and clears the hash table (slots a power of two) and its group counter.
static void ensure_table(ws_agg* A, int64_t slots) {
  if (A->keys) return;
  A->st_valid = false;
  const int nk = A->plan.nkeys, na = std::max(A->plan.naggs, 1);
  A->tcap = slots;
  A->keys = (u64*)tmp_alloc(sizeof(u64) * (size_t)slots);
  A->kv = (double*)tmp_alloc(sizeof(double) * (size_t)slots * nk);
  A->acc = (double*)tmp_alloc(sizeof(double) * (size_t)slots * na);
  fill_u64(A->count, 1, 0);
  GTable t{A->keys, A->kv, A->acc, (u64)(slots - 1), A->flags, A->count};
  hipLaunchKernelGGL(init_table_k, dim3(grid_for(slots)), dim3(256), 0, stream(), t, nk, A->plan.naggs, A->plan);
  check_launch("init_table");
}
// Gives the flags, counter and dense accumulators their initial values if creation deferred it.
static void prepare(ws_agg* A) {
  if (!A->pending_init) return;
  A->pending_init = false;
  const int64_t n = A->is_dense ? A->cap * std::max(A->plan.naggs, 1) : 4;
  hipLaunchKernelGGL(agg_prepare_k, dim3(grid_for(std::max<int64_t>(n, 4))), dim3(256), 0, stream(), (uint32_t*)A->flags, A->count,
                     A->is_dense ? A->acc : nullptr, A->cap, A->plan.naggs, A->plan);
  check_launch("agg_prepare");
}
// Reads the aggregation's small results in one round trip (cached until the next launch).
// Dense accumulators up to this size come back with the status read (one round trip per fetch).
constexpr size_t kSmallDenseFetchBytes = 56 * 1024;
static const ws_agg::Status& status(ws_agg* A) {
  prepare(A);
  if (A->st_valid) return A->st;
  A->st = ws_agg::Status{};
  D2HPiece pc[3];
  int n = 0;
  if (A->flags) pc[n++] = D2HPiece{A->st.flags, A->flags, sizeof A->st.flags};
  if (A->count) pc[n++] = D2HPiece{&A->st.count, A->count, sizeof(u64)};
  if (A->part && A->goff) pc[n++] = D2HPiece{&A->st.groups, A->goff + A->P, sizeof(uint32_t)};
  const size_t acc_bytes = A->is_dense && !A->part && A->acc ? sizeof(double) * (size_t)A->dense.G * A->plan.naggs : 0;
  A->host_acc_ok = acc_bytes > 0 && acc_bytes <= kSmallDenseFetchBytes;
  if (A->host_acc_ok) {
    A->host_acc.resize(acc_bytes / sizeof(double));
    pc[n++] = D2HPiece{A->host_acc.data(), A->acc, acc_bytes};
  }
  d2h_gather_sync(pc, n);
  A->st_valid = true;
  return A->st;
}
static GTable table_of(const ws_agg* A) { return GTable{A->keys, A->kv, A->acc, (u64)(A->tcap - 1), A->flags, A->count}; }

// The lean interpreter serves plans whose programs use no scalar-function opcodes.
static bool plan_is_simple(const ws_expr_env& env, const ws_agg_plan& plan) {
  if (plan.has_where && !prog_is_simple(env, plan.where)) return false;
  for (int k = 0; k < plan.nkeys; ++k)
    if (!prog_is_simple(env, plan.keys[k])) return false;
  for (int a = 0; a < plan.naggs; ++a)
    if (plan.ops[a] != WS_AGG_COUNT && !prog_is_simple(env, plan.aggs[a])) return false;
  return true;
}
static bool prog_is_leaf(const ws_expr_env& env, ws_prog p) {
  return p.len == 1 && p.start >= 0 && p.start < WS_MAX_CODE &&
         (env.code[p.start].op == WS_OP_COL || env.code[p.start].op == WS_OP_CONST);
}
static int interp_of(const ws_expr_env& env, const ws_agg_plan& plan) {
  bool leaf = !plan.has_where || prog_is_leaf(env, plan.where);
  for (int k = 0; k < plan.nkeys; ++k) leaf = leaf && prog_is_leaf(env, plan.keys[k]);
  for (int a = 0; a < plan.naggs; ++a) leaf = leaf && (plan.ops[a] == WS_AGG_COUNT || prog_is_leaf(env, plan.aggs[a]));
  return leaf ? kIvCols : plan_is_simple(env, plan) ? kIvLean : kIvFull;
}

static double host_identity(uint8_t op) { return op == WS_AGG_MIN ? INFINITY : op == WS_AGG_MAX ? -INFINITY : 0.0; }

// Launches the grouped kernel (compiled, or the interpreter) of a plan with one descriptor.
// pass: 0 the table kernel, 1 partition count, 2 partition scatter; jf: the compiled kernel of
// that pass, or null for the interpreter.
static void launch_grouped(void* jf, int pass, int iv, ScanDesc* ddesc, GTable t, unsigned grid, size_t lds,
                           const char* name, const char* jit_name, int block = kBlock) {
  if (jf) {
    void* args[] = {&ddesc, &t};
    jit_launch_raw(jf, grid, (unsigned)block, lds, args, jit_name);
    return;
  }
  ProfScope ps(name);
  if (pass == 0)
    hipLaunchKernelGGL(IV_PICK(iv, grouped), dim3(grid), dim3(block), lds, stream(), ddesc, t);
  else if (pass == 1)
    hipLaunchKernelGGL(IV_PICK(iv, grouped_pcount), dim3(grid), dim3(block), lds, stream(), ddesc, t);
  else
    hipLaunchKernelGGL(IV_PICK(iv, grouped_pscatter), dim3(grid), dim3(block), lds, stream(), ddesc, t);
  check_launch(name);
}

// ------------------------------------------------------------------ partitioned GROUP BY (host)

constexpr int kPartMergeBlock = 512;
constexpr int kPartScanBlock = 1024;          // count / scatter blocks (one chunk each)
constexpr int kPartMaxBits = 12;

// Tuning knobs of the partitioned plan (set through ws_tune_set; the autotuner's interface).
struct TuneKnob {
  const char* name;
  std::atomic<int64_t> value;
  int64_t lo, hi;
};
TuneKnob g_tune[] = {
    {"part_merge_lds_kib", {60}, 8, 152},     // merge table per block (KiB) for full records; sets the slot count, packed tables take less
    {"part_scatter_lds_kib", {120}, 16, 152}, // cursors + pre-aggregation table per scatter block (KiB)
    {"part_chunks_per_cu", {1}, 1, 64},       // count / scatter blocks (chunks) per CU
    {"part_auto_rows", {1 << 20}, 0, (int64_t)1 << 40},   // automatic choice: at least this many rows
    {"part_auto_groups", {1 << 16}, 0, (int64_t)1 << 40}, // and this many expected groups
    {"batch_tile_rows", {4096}, 256, 1 << 20},  // batched dense run: rows per tile
    {"batch_lds_slots", {1024}, 64, 8192},      // its LDS table slots per block
    {"batch_blocks_per_cu", {4}, 1, 32},        // its grid, per CU
    {"part_pack", {1}, 0, 1},                   // packed partial-group records allowed (integer keys)
    {"select_chunks_per_cu", {8}, 1, 64},       // chunked selection: chunks of tiles per CU
    {"select_mask_min_rows", {1 << 16}, 0, (int64_t)1 << 40},  // compiled selection: ballot pass from this many rows
    {"dense_staged_min_set_words", {1024}, 0, (int64_t)1 << 40},  // compiled dense GROUP BY: sets staged in LDS from this size
    {"batch_key_ranges", {0}, 0, 64},           // batched dense run: most key ranges per query beyond the table (0: none)
    {"topk_steps", {8}, 1, 64},                 // top-k selection, first pass: steps of 256 rows per block
    {"topk_blocks_per_cu", {1}, 1, 16},         // and at most this many blocks per CU
};
static int64_t tune(int i) { return g_tune[i].value.load(std::memory_order_relaxed); }
int64_t ws::tune_knob(const char* name) {
  for (const TuneKnob& k : g_tune)
    if (std::strcmp(k.name, name) == 0) return k.value.load(std::memory_order_relaxed);
  throw Error(std::string("tune: unknown knob ") + name, WS_ERR_INTERNAL);
}
enum { kTuneMergeLds, kTuneScatterLds, kTuneChunks, kTuneAutoRows, kTuneAutoGroups, kTuneBatchTile, kTuneBatchSlots,
       kTuneBatchBlocksPerCu, kTunePack, kTuneSelectChunks, kTuneMaskMinRows, kTuneDenseStagedWords, kTuneBatchRanges };

// Dense plans whose WHERE reads set bitmaps of at least dense_staged_min_set_words words, over at
// least select_mask_min_rows rows: the staged kernel (one 1024-thread block per CU, up to four
// LDS copies of the accumulators, the sets' leading words in the rest of the LDS; tiles of
// kSelMaskRows rows per thread), then the usual fold of the block partials. False when the plan
// does not qualify or its copies would leave the sets less than half the LDS.
static bool run_dense_staged(ws_agg* A, const ws_expr_env& env, const ws_jit_program* jit, const ScanInput& in, ScanDesc d) {
  const ws_agg_plan& plan = A->plan;
  const ws_dense_spec& ds = A->dense;
  void* jf = plan.has_where ? jit_function(jit, kJitDenseStaged) : nullptr;
  if (!jf) return false;
  int64_t set_words = 0, rows = 0;
  for (int k = 0; k < WS_MAX_SETS; ++k)
    if (env.sets[k] && env.set_size[k] > 0) set_words += ((int64_t)env.set_size[k] + 31) / 32;
  for (int64_t n : in.n) rows += n;
  if (set_words == 0 || set_words < tune(kTuneDenseStagedWords) || rows < tune(kTuneMaskMinRows)) return false;
  size_t off = 0;
  for (int a = 0; a < plan.naggs; ++a) {
    d.slot_off[a] = (int32_t)off;
    const size_t bytes = (size_t)ds.G * (ds.kind[a] == WS_DK_COUNT ? 4 : 8);
    off += (bytes + 15) / 16 * 16;
  }
  const int C = (int)std::min<size_t>(4, std::max<size_t>(1, (16 * 1024) / std::max<size_t>(off, 1)));
  const size_t copies = off * C;
  const size_t stat = (size_t)jit_staged_static_lds(jit);
  if (copies + stat > kLdsPerCu / 2) return false;
  constexpr int kStagedBlock = 1024;
  d.mode = 2;
  d.lds_slots = C;
  d.copy_stride = (int32_t)off;
  d.tile_rows = kSelMaskRows * kStagedBlock;
  int64_t ntiles = 0;
  for (int64_t n : in.n) ntiles += (n + d.tile_rows - 1) / d.tile_rows;
  d.set_words = (int32_t)std::min<int64_t>(set_words, (int64_t)(kLdsPerCu - stat - copies) / 4);
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(ntiles, cu_count()));
  const int64_t S = ds.G * plan.naggs;
  DescUpload up;
  build_desc(up, d, in, plan, &ds, env, false, 0, nullptr);
  DevBuf<i64> parts(grid * S);
  int* flags = A->flags;
  i64* pp = parts.p;
  void* args[] = {&up.ddesc, &pp, &flags};
  jit_launch_raw(jf, (unsigned)grid, kStagedBlock, copies + 4 * (size_t)d.set_words, args, jit_prof_name(jit, "agg_dense_staged"));
  ProfScope ps("agg_dense_reduce");
  dense_reduce(plan, ds, (const i64*)pp, grid, A->acc);
  return true;
}

struct PartPlan {
  int pbits = 0;  // 0: plain table path
  int W = 0, S = 0, pre = 0, chunk_tiles = 0;
  int64_t nchunks = 0, rows = 0;
};

// Partitioned aggregation pays when groups are many (a shared table then takes one device-scope
// atomic per row and slot); the partition count keeps each partition's expected groups at about
// half of the merge block's LDS table. Only a first run over all batches is partitioned.
static PartPlan plan_partitions(const ws_agg* A, int64_t rows, int64_t ntiles) {
  PartPlan pp;
  const ws_agg_plan& plan = A->plan;
  if (A->strategy == WS_STRAT_TABLE || A->is_dense || plan.nkeys == 0 || A->runs > 0 || A->keys || A->part) return pp;
  if (rows <= 0 || rows >= ((int64_t)1 << 31)) return pp;
  const int W = 1 + plan.nkeys + plan.naggs;
  if (W < 3 || W > 16) return pp;  // the merge kernels' record widths
  const int64_t expect = A->expected_groups > 0 ? A->expected_groups : std::max<int64_t>(A->cap / 2, 1);
  if (A->strategy == WS_STRAT_AUTO && (rows < tune(kTuneAutoRows) || expect < tune(kTuneAutoGroups))) return pp;
  pp.W = W;
  pp.S = (int)((size_t)tune(kTuneMergeLds) * 1024 / (8 * (size_t)W));
  int bits = 1;
  while (bits < kPartMaxBits && ((int64_t)1 << bits) * (pp.S / 2) < expect) ++bits;
  pp.pbits = bits;
  const int P = 1 << bits;
  int pre = 8192;
  while (pre > 64 && 8 * (size_t)P + (size_t)pre * 8 * W > (size_t)tune(kTuneScatterLds) * 1024) pre >>= 1;
  pp.pre = pre;
  const int64_t want = std::max<int64_t>(1, (int64_t)cu_count() * tune(kTuneChunks));
  pp.chunk_tiles = (int)std::max<int64_t>(1, (ntiles + want - 1) / want);
  pp.nchunks = (ntiles + pp.chunk_tiles - 1) / pp.chunk_tiles;
  pp.rows = rows;
  return pp;
}

static void agg_run_part(ws_agg* A, const ws_expr_env& env, const ws_jit_program* jit, const ScanInput& in, ScanDesc d,
                         const PartPlan& pp) {
  const ws_agg_plan& plan = A->plan;
  // Both scan passes must evaluate rows identically: compiled both, or interpreted both.
  void* jf_count = jit ? jit_function(jit, kJitPartCount) : nullptr;
  void* jf_scatter = jit ? jit_function(jit, kJitPartScatter) : nullptr;
  if (!jf_count || !jf_scatter) jf_count = jf_scatter = nullptr;
  const int iv = interp_of(env, plan);
  const int NK = plan.nkeys, NA = plan.naggs, P = 1 << pp.pbits;
  DevBuf<uint32_t> cnt(pp.nchunks * P), off(pp.nchunks * P), ptot(P), gcnt(P);
  DevBuf<u64> recs((pp.rows + 8 * (int64_t)kPartMergeBlock) * pp.W);  // padded for the merge's unguarded loads
  // Integer keys (at most kPackKeys): the count pass measures their ranges and, when they fit 63
  // bits together, the scatter pass writes packed records (fewer bytes to write and to merge).
  bool packable = NK <= kPackKeys && tune(kTunePack) != 0;
  for (int k = 0; k < NK; ++k) packable = packable && plan.key_int[k];
  DevBuf<int64_t> krange(packable ? pp.nchunks * 2 * NK : 0);  // every chunk writes its partial
  DevBuf<PackPlan> pack(packable ? 1 : 0);
  A->okv = (double*)tmp_alloc(sizeof(double) * (size_t)pp.rows * NK);
  A->oacc = (double*)tmp_alloc(sizeof(double) * (size_t)pp.rows * std::max(NA, 1));
  A->pstart = (uint32_t*)tmp_alloc(sizeof(uint32_t) * (size_t)(P + 1));
  A->goff = (uint32_t*)tmp_alloc(sizeof(uint32_t) * (size_t)(P + 1));
  A->P = P;
  A->part = true;
  A->settled = false;
  uint32_t* pstart = A->pstart;
  // spill table for partitions whose groups outgrow the merge block's LDS table
  // (grows with the caller's capacity, so overflow retries converge even under a bad estimate)
  int64_t spill = 4096;
  const int64_t expect = A->expected_groups > 0 ? A->expected_groups : std::max<int64_t>(A->cap / 2, 1);
  while (spill < A->cap && (spill < expect / 4 || spill < A->cap / 8)) spill <<= 1;
  ensure_table(A, spill);
  d.pbits = pp.pbits;
  d.chunk_tiles = pp.chunk_tiles;
  d.pre_slots = pp.pre;
  d.rec_words = pp.W;
  d.part_cnt = cnt.p;
  d.part_off = off.p;
  d.part_start = pstart;
  d.recs = recs.p;
  d.part_err = A->flags + 2;
  d.key_range = packable ? krange.p : nullptr;
  d.pack = packable ? pack.p : nullptr;
  GTable t = table_of(A);
  {
    DescUpload up;
    build_desc(up, d, in, plan, nullptr, env, false, 0, nullptr);
    launch_grouped(jf_count, 1, iv, up.ddesc, t, (unsigned)pp.nchunks, (size_t)P * 4, "agg_part_count",
                   jit_prof_name(jit, "agg_part_count_jit"), kPartScanBlock);
  }
  {
    ProfScope ps("agg_part_offsets");
    hipLaunchKernelGGL(part_offsets_k, dim3((unsigned)((P + 63) / 64)), dim3(1024), 0, stream(), (const uint32_t*)cnt.p, off.p,
                       ptot.p, pp.nchunks, P);
    hipLaunchKernelGGL(part_start_k, dim3(1), dim3(1024), 0, stream(), (const uint32_t*)ptot.p, pstart, P,
                       (const int64_t*)krange.p, pp.nchunks, NK, packable ? pack.p : nullptr);
    check_launch("agg_part_offsets");
  }
  {
    DescUpload up;
    build_desc(up, d, in, plan, nullptr, env, false, 0, nullptr);
    const size_t lds = (size_t)P * 8 + (size_t)pp.pre * 8 * pp.W;
    launch_grouped(jf_scatter, 2, iv, up.ddesc, t, (unsigned)pp.nchunks, lds, "agg_part_scatter", jit_prof_name(jit, "agg_part_scatter_jit"),
                   kPartScanBlock);
  }
  {
    u64 kinds = 0;
    for (int a = 0; a < NA; ++a) kinds |= (u64)((plan.ops[a] & 3) | (plan.agg_bits[a] ? 4 : 0)) << (3 * a);
    PartMerge m{recs.p, pstart, pp.W, NK, NA, pp.S, kinds, A->okv, A->oacc, gcnt.p, packable ? pack.p : nullptr};
    ProfScope ps("agg_part_merge");
    // both layouts are launched; the one the scatter pass did not choose returns at once
    hipLaunchKernelGGL(merge_kernel(pp.W, false, false), dim3((unsigned)P), dim3(kPartMergeBlock), (size_t)pp.S * 8 * pp.W, stream(), m,
                       t);
    if (packable)  // a packed slot is 1 + NA words: less LDS per block, more blocks per CU
      hipLaunchKernelGGL(merge_kernel(1 + NA, true, kinds_additive(kinds, NA)), dim3((unsigned)P), dim3(kPartMergeBlock), (size_t)pp.S * 8 * (1 + NA), stream(),
                         m, t);
    hipLaunchKernelGGL(part_start_k, dim3(1), dim3(1024), 0, stream(), (const uint32_t*)gcnt.p, A->goff, P);
    check_launch("agg_part_merge");
  }
}

static void agg_run(ws_agg* A, const ws_expr_env& env, const ws_jit_program* jit, bool allow_part);

// Waits for a partitioned run and reads its group count. A chunk whose second pass found more rows
// than its first (which the passes' identical evaluation rules out) is answered by redoing the
// aggregation on the plain path, so the result never depends on it.
static void settle(ws_agg* A) {
  if (!A->part || A->settled) return;
  const ws_agg::Status& s = status(A);  // flags, spilled groups and finished groups together
  A->settled = true;
  A->part_groups = (int64_t)s.groups;
  if (s.flags[2]) {
    A->free_part();
    A->free_table();
    fill_u32((uint32_t*)A->flags, 4, 0);
    A->st_valid = false;
    A->strategy = WS_STRAT_TABLE;
    agg_run(A, A->last_env, A->last_jit, false);
  }
}

// Turns a partitioned result into a plain table of the caller's capacity (more batches follow).
static void part_to_table(ws_agg* A) {
  settle(A);
  if (!A->part) return;
  const u64 spilled = status(A).count;
  A->st_valid = false;
  u64* okeys = A->keys;
  double *okv = A->kv, *oacc = A->acc;
  const int64_t ocap = A->tcap;
  A->keys = nullptr, A->kv = nullptr, A->acc = nullptr, A->tcap = 0;
  ensure_table(A, A->cap);
  const GTable t = table_of(A);
  if (A->part_groups)
    hipLaunchKernelGGL(table_absorb_k, dim3(grid_for(A->part_groups)), dim3(256), 0, stream(), t, A->plan, (const double*)A->okv,
                       (const double*)A->oacc, (const u64*)nullptr, (const uint32_t*)A->goff, (const uint32_t*)A->pstart, A->P,
                       A->part_groups);
  if (spilled)
    hipLaunchKernelGGL(table_absorb_k, dim3(grid_for(ocap)), dim3(256), 0, stream(), t, A->plan, (const double*)okv,
                       (const double*)oacc, (const u64*)okeys, (const uint32_t*)nullptr, (const uint32_t*)nullptr, 0, ocap);
  check_launch("table_absorb");
  for (void* p : {(void*)okeys, (void*)okv, (void*)oacc}) tmp_free(p);
  A->free_part();
}

static void agg_run(ws_agg* A, const ws_expr_env& env, const ws_jit_program* jit, bool allow_part) {
  prepare(A);
  const ws_agg_plan& plan = A->plan;
  if (&env != &A->last_env) A->last_env = env;
  A->last_jit = jit;
  A->st_valid = false;
  if (!jit && (plan.naggs > WS_INTERP_AGGS || plan.nkeys > WS_INTERP_KEYS))
    throw Error("this aggregation needs a compiled plan (more than 16 aggregates or 4 GROUP BY keys); codegen failed or is off",
                WS_ERR_NEEDS_PLAN);
  const int nb = (int)A->bn.size();
  if (plan.nkeys == 0 && !A->nk_out) {  // no GROUP BY: the running result, initialized once
    std::vector<double> init(2 * WS_MAX_AGGS, 0.0);
    for (int a = 0; a < WS_MAX_AGGS; ++a) {
      double v = a < plan.naggs ? host_identity(plan.ops[a]) : 0.0;
      if (a < plan.naggs && plan.agg_int[a] && (plan.ops[a] == WS_AGG_MIN || plan.ops[a] == WS_AGG_MAX)) {
        const long long id = plan.ops[a] == WS_AGG_MIN ? LLONG_MAX : LLONG_MIN;
        std::memcpy(&v, &id, 8);
      }
      init[a] = v;
    }
    A->nk_out = (double*)tmp_alloc(sizeof(double) * 2 * WS_MAX_AGGS);
    h2d_async_staged(A->nk_out, init.data(), sizeof(double) * init.size());
  }
  if (nb == 0) return;
  ScanInput in;
  in.nb = nb;
  in.ncols = A->ncols;
  in.n = A->bn;
  in.del = A->bdel;
  in.cols = A->bcols;
  ScanDesc d{};
  d.tile_rows = kAggTileRows;
  int64_t ntiles = 0;
  for (int64_t n : in.n) ntiles += (n + d.tile_rows - 1) / d.tile_rows;
  const int fn = plan.nkeys == 0 ? WS_JF_NOKEYS : A->is_dense ? WS_JF_DENSE : WS_JF_GROUPED;
  void* jf = jit ? jit_function(jit, fn) : nullptr;
  const int iv = interp_of(env, plan);
  const int cus = cu_count();
  if (plan.nkeys == 0) {
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)cus * 8));
    DescUpload up;
    size_t o_done = 0;
    build_desc(up, d, in, plan, nullptr, env, false, 16, &o_done);
    DevBuf<NoKeysPartial> parts(grid);
    unsigned* done = (unsigned*)(up.dev + o_done);
    double* out = A->nk_out;
    if (jf) {
      void* args[] = {&up.ddesc, &out, &parts.p, &done};
      jit_launch_raw(jf, grid, kBlock, 0, args, "agg_nokeys_jit");
    } else {
      ProfScope ps("agg_nokeys");
      hipLaunchKernelGGL(IV_PICK(iv, nokeys), dim3(grid), dim3(kBlock), 0, stream(), up.ddesc, out, parts.p, done);
      check_launch("agg_nokeys");
    }
    return;
  }
  if (A->is_dense) {
    const ws_dense_spec& ds = A->dense;
    if (jf && run_dense_staged(A, env, jit, in, d)) return;
    DensePlan dp = plan_dense(ds, plan.naggs);
    d.mode = dp.mode;
    d.lds_slots = dp.copies;
    d.copy_stride = dp.copy_stride;
    std::memcpy(d.slot_off, dp.slot_off, sizeof d.slot_off);
    const size_t lds_block = dp.lds + kStaticLds;
    int64_t bpc = std::max<int64_t>(1, std::min<int64_t>(8, (int64_t)(kLdsPerCu / lds_block)));
    int64_t grid = std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)cus * bpc));
    const int64_t S = ds.G * plan.naggs;
    if (dp.mode != 0) grid = std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t)(64ll << 20) / (S * 8)));
    DescUpload up;
    build_desc(up, d, in, plan, &ds, env, false, 0, nullptr);
    DevBuf<i64> parts(dp.mode != 0 ? grid * S : 0);
    double* gacc = A->acc;
    int* flags = A->flags;
    i64* pp = parts.p;
    if (jf) {
      void* args[] = {&up.ddesc, &gacc, &pp, &flags};
      jit_launch_raw(jf, (unsigned)grid, kBlock, dp.lds, args, "agg_dense_jit");
    } else {
      ProfScope ps("agg_dense");
      hipLaunchKernelGGL(IV_PICK(iv, dense), dim3((unsigned)grid), dim3(kBlock), dp.lds, stream(), up.ddesc, gacc, pp,
                         flags);
      check_launch("agg_dense");
    }
    if (dp.mode != 0) {
      ProfScope ps("agg_dense_reduce");
      dense_reduce(plan, ds, (const i64*)pp, grid, gacc);
    }
    return;
  }
  int64_t rows = 0;
  for (int64_t n : in.n) rows += n;
  const PartPlan pp = allow_part ? plan_partitions(A, rows, ntiles) : PartPlan{};
  if (pp.pbits) {
    agg_run_part(A, env, jit, in, d, pp);
    return;
  }
  if (A->part) part_to_table(A);
  ensure_table(A, A->cap);
  // hash GROUP BY: an LDS table per block, sized to leave room for several blocks per CU
  const size_t per_slot = 8 + 8 * (size_t)plan.nkeys + 8 * (size_t)plan.naggs;
  int S = 64;
  while ((size_t)S * 2 * per_slot <= 32 * 1024) S *= 2;
  d.lds_slots = S;
  const size_t lds = (size_t)S * per_slot;
  const int64_t bpc = std::max<int64_t>(1, std::min<int64_t>(8, (int64_t)(kLdsPerCu / (lds + kStaticLds))));
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)cus * bpc));
  DescUpload up;
  build_desc(up, d, in, plan, nullptr, env, false, 0, nullptr);
  GTable t = table_of(A);
  launch_grouped(jf, 0, iv, up.ddesc, t, grid, lds, "agg_hash", jit_prof_name(jit, "agg_hash_jit"));
}

// Compacts the groups into tmp buffers: keys then accs in `buf`, row-major (fetch) or
// column-major; with `abuf` the accs go there instead (two separately owned buffers).
static int64_t agg_compact_into(ws_agg* A, DevBuf<double>& buf, bool colmajor, DevBuf<double>* abuf = nullptr) {
  const int nk = A->plan.nkeys, na = A->plan.naggs;
  prepare(A);
  settle(A);
  if (!A->is_dense && !A->keys) {  // no run yet: no groups
    buf = DevBuf<double>(1);
    if (abuf) *abuf = DevBuf<double>(1);
    return 0;
  }
  const int64_t slots = A->is_dense ? A->dense.G : A->tcap;
  DevBuf<int32_t> idx(slots);
  int64_t g1 = 0;  // finished groups of a partitioned run come first, then any spilled ones
  bool table_groups = true;
  if (A->part) {
    const u64 spilled = status(A).count;
    g1 = A->part_groups;
    table_groups = spilled != 0;
  }
  const int64_t g2 = !table_groups ? 0
                     : A->is_dense ? compact_dense(A->acc, na, A->dense.G, idx.p)
                                   : compact_occupied_hash(A->keys, A->tcap, idx.p);
  const int64_t g = g1 + g2;
  double *ko, *ao;
  if (abuf) {
    buf = DevBuf<double>(std::max<int64_t>(g * nk, 1));
    *abuf = DevBuf<double>(std::max<int64_t>(g * na, 1));
    ko = buf.p;
    ao = abuf->p;
  } else {
    buf = DevBuf<double>(std::max<int64_t>(g * (nk + na), 1));
    ko = buf.p;
    ao = buf.p + g * nk;
  }
  ProfScope ps("gather_groups");
  if (g1)
    hipLaunchKernelGGL(gather_part_k, dim3((unsigned)A->P), dim3(256), 0, stream(), (const double*)A->okv, nk, (const double*)A->oacc,
                       na, (const uint32_t*)A->goff, (const uint32_t*)A->pstart, A->P, g1, ko, ao, colmajor ? 1 : 0, g);
  if (g2) {
    if (A->is_dense)
      hipLaunchKernelGGL(gather_dense_k, dim3(grid_for(g2 * (nk + na))), dim3(256), 0, stream(), A->dense, A->plan, A->acc, nk,
                         na, idx.p, g2, ko, ao, colmajor ? 1 : 0);
    else
      hipLaunchKernelGGL(gather_groups_k, dim3(grid_for(g2 * (nk + na))), dim3(256), 0, stream(), (const double*)A->kv, nk,
                         (const double*)A->acc, na, (const int32_t*)idx.p, g2, ko, ao, colmajor ? 1 : 0, g1, g);
  }
  check_launch("gather_groups");
  return g;
}

// ------------------------------------------------------------------ batched runs

// A batched run's launch description: the blob [BatchQuery n][BatchTile T] then per query plan,
// dense spec, env, n[], del[], cols[], tile0[], with device addresses resolved against `dev`.
// LDS bytes of the batched kernel's table of `slots` slots, and its slot count: the largest power
// of two up to the batch_lds_slots knob that fits the LDS.
static size_t batch_lds(int slots, int na) {
  return (((size_t)slots * 4 + 15) & ~(size_t)15) + (size_t)slots * std::max(na, 1) * 8;
}
static int batch_slots(int na) {
  int S = 64;
  while (S * 2 <= (int)tune(kTuneBatchSlots) && batch_lds(S * 2, na) + kStaticLds <= kLdsPerCu) S *= 2;
  return S;
}
// A query with more groups than the table, in at most batch_key_ranges table-sized key ranges, reads
// its rows once per range (groups in LDS) instead of sending them to device-scope atomics.

struct BatchBuild {
  int S = 64;  // the kernel's LDS table slots
  Blob blob;
  uint8_t* dev = nullptr;
  const BatchQuery* dq = nullptr;
  const BatchTile* dt = nullptr;
  int64_t ntiles = 0, gmax = 1;
  int pending = 0, iv = kIvCols, na = 0;
  void* jf = nullptr;
};
template <class Alloc>
static BatchBuild build_batch(int n, ws_agg* const* aggs, const ws_expr_env* const* envs, const ws_jit_program* jit,
                              bool init_all, Alloc alloc) {
  BatchBuild r;
  const ws_agg* A0 = aggs[0];
  const int nk = A0->plan.nkeys, na = A0->plan.naggs;
  r.na = na;
  for (int q = 0; q < n; ++q) {
    const ws_agg* A = aggs[q];
    if (!A->is_dense || A->plan.nkeys != nk || A->plan.naggs != na)
      throw Error("batched run: every aggregation must be dense with the same plan shape", WS_ERR_INTERNAL);
    r.iv = std::min(r.iv, interp_of(*envs[q], A->plan));
  }
  if (!jit && (na > WS_INTERP_AGGS || nk > WS_INTERP_KEYS))
    throw Error("this aggregation needs a compiled plan (more than 16 aggregates or 4 GROUP BY keys); codegen failed or is off",
                WS_ERR_NEEDS_PLAN);
  r.jf = jit ? jit_function(jit, kJitBatchDense) : nullptr;
  const int tile_rows = (int)tune(kTuneBatchTile);
  r.S = batch_slots(na);
  std::vector<BatchTile> tiles;
  for (int q = 0; q < n; ++q) {
    const int64_t G = aggs[q]->dense.G;
    const int64_t ranges = G > r.S ? (G + r.S - 1) / r.S : 0;
    const bool split = ranges > 0 && ranges <= tune(kTuneBatchRanges);
    for (int64_t k = 0; k < (split ? ranges : 1); ++k) {
      const int32_t glo = split ? (int32_t)(k * r.S) : -1;
      for (size_t b = 0; b < aggs[q]->bn.size(); ++b)
        for (int64_t f = 0; f < aggs[q]->bn[b]; f += tile_rows)
          tiles.push_back(BatchTile{q, (int32_t)b, (int32_t)std::min<int64_t>(tile_rows, aggs[q]->bn[b] - f), glo, f});
    }
  }
  r.ntiles = (int64_t)tiles.size();
  Blob& blob = r.blob;
  const size_t o_q = blob.reserve(sizeof(BatchQuery) * n);
  const size_t o_t = blob.put(tiles.data(), sizeof(BatchTile) * tiles.size());
  std::vector<size_t> o_plan(n), o_ds(n), o_env(n), o_n(n), o_del(n), o_cols(n), o_t0(n);
  for (int q = 0; q < n; ++q) {
    const ws_agg* A = aggs[q];
    o_plan[q] = blob.put(&A->plan, sizeof A->plan);
    o_ds[q] = blob.put(&A->dense, sizeof A->dense);
    o_env[q] = blob.put(envs[q], sizeof(ws_expr_env));
    o_n[q] = blob.put(A->bn.data(), 8 * A->bn.size());
    o_del[q] = blob.put(A->bdel.data(), 8 * A->bdel.size());
    o_cols[q] = blob.put(A->bcols.data(), sizeof(ws_colview) * A->bcols.size());
    std::vector<int64_t> t0(A->bn.size() + 1, 0);
    o_t0[q] = blob.put(t0.data(), 8 * t0.size());
  }
  uint8_t* dev = alloc(blob.host.size());
  r.dev = dev;
  for (int q = 0; q < n; ++q) r.gmax = std::max<int64_t>(r.gmax, aggs[q]->dense.G * std::max(na, 1));
  for (int q = 0; q < n; ++q) {
    ws_agg* A = aggs[q];
    BatchQuery bq{};
    ScanDesc& d = bq.d;
    d.nb = (int32_t)A->bn.size();
    d.ncols = A->ncols;
    d.tile_rows = tile_rows;
    d.n = (const int64_t*)(dev + o_n[q]);
    d.del = (const uint32_t* const*)(dev + o_del[q]);
    d.cols = (const ws_colview*)(dev + o_cols[q]);
    d.tile0 = (const int64_t*)(dev + o_t0[q]);
    d.row0 = d.tile0;
    d.env = (const ws_expr_env*)(dev + o_env[q]);
    d.plan = (const ws_agg_plan*)(dev + o_plan[q]);
    d.ds = (const ws_dense_spec*)(dev + o_ds[q]);
    bq.gacc = A->acc;
    bq.flags = A->flags;
    bq.count = A->count;
    bq.init = (init_all || A->pending_init) ? 1 : 0;
    r.pending += bq.init;
    *blob.at<BatchQuery>(o_q + sizeof(BatchQuery) * q) = bq;
    A->fetched = -1;
    A->st_valid = false;
    ++A->runs;
  }
  r.dq = (const BatchQuery*)(dev + o_q);
  r.dt = (const BatchTile*)(dev + o_t);
  return r;
}
// The batch's launches: the initial values (queries whose BatchQuery asks for them) and the
// batched kernel. prof = false leaves out the profiling events (inside a graph capture).
static void launch_batch(const BatchBuild& b, int n, bool prof) {
  if (b.pending) {  // the deferred initial values of the whole batch, one launch
    const unsigned gx = (unsigned)std::min<int64_t>(64, (b.gmax + 2047) / 2048);
    hipLaunchKernelGGL(batch_init_k, dim3(gx, (unsigned)n), dim3(256), 0, stream(), b.dq);
    check_launch("batch_init");
  }
  const int S = b.S;
  const size_t lds = batch_lds(S, b.na);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(b.ntiles, (int64_t)cu_count() * tune(kTuneBatchBlocksPerCu)));
  const BatchQuery* dq = b.dq;
  const BatchTile* dt = b.dt;
  long long nt = (long long)b.ntiles;
  int sl = S;
  if (b.jf) {
    void* args[] = {&dq, &dt, &nt, &sl};
    if (prof) jit_launch_raw(b.jf, grid, kBlock, lds, args, "agg_batch_dense_jit");
    else WS_CK(hipModuleLaunchKernel((hipFunction_t)b.jf, grid, 1, 1, kBlock, 1, 1, (unsigned)lds, stream(), args, nullptr));
  } else {
    std::unique_ptr<ProfScope> ps(prof ? new ProfScope("agg_batch_dense") : nullptr);
    const int iv = b.iv;
    hipLaunchKernelGGL(IV_PICK(iv, dense_batch), dim3(grid), dim3(kBlock), lds, stream(), dq, dt, nt, sl);
    check_launch("agg_batch_dense");
  }
}

// A batch of dense aggregations prepared once and replayed: its descriptor blob stays resident,
// and its launches (initial values, the batched kernel, then the fetch's count, scan and pack
// kernels) are recorded in a HIP graph; a replay is one graph launch and two read-backs.
// A prepared batch's stream, events and page-locked offsets buffer, kept for the next plan on the
// same device once the plan is released (a pipeline that prepares a plan per batch would otherwise
// create and destroy them per batch, and page-locking memory is a slow call). Idle ones only: a
// plan's stream is synchronized before its resources come back.
struct PlanRes {
  int dev = -1;
  hipStream_t cs = nullptr;
  hipEvent_t ev_in = nullptr, ev_cnt = nullptr;
  uint32_t* hoffs = nullptr;
  size_t hcap = 0;  // words of hoffs
};
static std::mutex g_plan_res_mu;
static std::vector<PlanRes> g_plan_res;
static PlanRes plan_res_acquire(size_t words) {
  int dev = 0;
  WS_CK(hipGetDevice(&dev));
  PlanRes r;
  {
    std::lock_guard<std::mutex> g(g_plan_res_mu);
    for (size_t i = g_plan_res.size(); i-- > 0;)
      if (g_plan_res[i].dev == dev) {
        r = g_plan_res[i];
        g_plan_res.erase(g_plan_res.begin() + (std::ptrdiff_t)i);
        break;
      }
  }
  r.dev = dev;
  if (!r.cs) WS_CK(hipStreamCreateWithFlags(&r.cs, hipStreamNonBlocking));
  if (!r.ev_in) WS_CK(hipEventCreateWithFlags(&r.ev_in, hipEventDisableTiming));
  if (!r.ev_cnt) WS_CK(hipEventCreateWithFlags(&r.ev_cnt, hipEventDisableTiming));
  if (r.hcap < words) {
    if (r.hoffs) (void)hipHostFree(r.hoffs);
    r.hoffs = nullptr;
    r.hcap = 0;
    WS_CK(hipHostMalloc((void**)&r.hoffs, sizeof(uint32_t) * words, hipHostMallocDefault));
    r.hcap = words;
  }
  return r;
}
static void plan_res_release(const PlanRes& r) {
  std::lock_guard<std::mutex> g(g_plan_res_mu);
  g_plan_res.push_back(r);
}

struct ws_agg_batch_plan {
  int n = 0, nk = 0, na = 0;
  ws_agg_plan plan{};
  BatchBuild b;
  ws::DevBuf<ws::FetchQuery> fq;
  ws::DevBuf<ws::FetchChunk> ch;  // fetch chunks; cb[q] = first chunk of query q
  std::vector<int64_t> cb;
  int64_t nch = 0;
  ws::DevBuf<uint32_t> cnt, offs;
  ws::DevBuf<double> pk, pa;
  int64_t gcap = 0;
  PlanRes res;               // from the pool (returned on release)
  hipStream_t cs = nullptr;  // the plan's stream: its graph and read-backs
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev_in = nullptr, ev_cnt = nullptr;  // the caller's earlier work; the counts read back
  uint32_t* hoffs = nullptr;                     // pinned: the group offsets of the last launch
  bool launched = false;
  ~ws_agg_batch_plan() {
    if (cs) (void)hipStreamSynchronize(cs);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (res.cs) plan_res_release(res);
    if (b.dev) ws::tmp_free(b.dev);
  }
};

extern "C" {

int ws_agg_create(const ws_agg_plan* plan, int64_t capacity, const ws_dense_spec* dense, ws_agg** out) {
  return guard([&] {
    *out = nullptr;
    ensure_init();
    auto A = std::make_unique<ws_agg>();
    A->plan = *plan;
    A->plan.genv = nullptr;
    if (dense && dense->G > 0 && plan->nkeys) {
      A->dense = *dense;
      A->is_dense = true;
    }
    int64_t c = 1;
    while (c < capacity) c <<= 1;
    A->cap = plan->nkeys ? (A->is_dense ? A->dense.G : c) : 1;
    if (plan->nkeys) {
      const int na = std::max(plan->naggs, 1);
      A->flags = (int*)tmp_alloc(sizeof(int) * 4);
      A->count = (u64*)tmp_alloc(sizeof(u64));
      // The hash table is allocated by the first run, which knows whether it is partitioned.
      if (A->is_dense) A->acc = (double*)tmp_alloc(sizeof(double) * (size_t)A->cap * na);
      A->pending_init = true;  // initial values at the first use (prepare, or a batched run)
    }
    *out = A.release();
  });
}
void ws_agg_destroy(ws_agg* a) { delete a; }
int ws_agg_add_batch(ws_agg* a, const ws_expr_env* env, int ncols, int64_t n, const uint32_t* del) {
  return guard([&] {
    if (n <= 0) return;
    a->ncols = ncols;
    a->bn.push_back(n);
    a->bdel.push_back(del);
    for (int i = 0; i < ncols; ++i) a->bcols.push_back(env->cols[i]);
  });
}
int ws_agg_run(ws_agg* a, const ws_expr_env* env, const ws_jit_program* jit) {
  return guard([&] {
    agg_run(a, *env, jit, true);
    ++a->runs;
    a->fetched = -1;
  });
}
int ws_agg_run_window(ws_agg* a, const ws_expr_env* env, const ws_jit_program* jit) {
  return guard([&] {
    if (a->bn.empty()) return;
    agg_run(a, *env, jit, false);  // windows accumulate into the shared table
    ++a->runs;
    ++a->windows;
    for (int64_t n : a->bn) a->rows_done += n;
    a->batches_done += (int)a->bn.size();
    a->bn.clear();
    a->bdel.clear();
    a->bcols.clear();
  });
}
int ws_agg_batches(const ws_agg* a) { return a->batches_done + (int)a->bn.size(); }
int ws_agg_windows(const ws_agg* a) { return a->windows; }
int64_t ws_agg_rows_in(const ws_agg* a) {
  int64_t r = a->rows_done;
  for (int64_t n : a->bn) r += n;
  return r;
}
int ws_agg_dense(const ws_agg* a) { return a->is_dense ? 1 : 0; }
int ws_agg_overflowed(ws_agg* a, int* out) {
  return guard([&] {
    *out = 0;
    if (a->plan.nkeys == 0) return;
    settle(a);
    const ws_agg::Status& s = status(a);
    if (s.flags[1]) throw Error("dense GROUP BY key outside the planned key range", WS_ERR_INTERNAL);
    if (!a->is_dense) *out = s.flags[0] != 0;
    if (!*out && a->has_features && !a->logged && !a->is_dense) {  // the exact count, for the sizing model
      const u64 c = a->keys ? s.count : 0;
      sizing_log(a->features, (int64_t)c + (a->part ? a->part_groups : 0));
      a->logged = true;
    }
  });
}
int ws_agg_set_features(ws_agg* a, const float* features) {
  std::memcpy(a->features, features, sizeof a->features);
  a->has_features = true;
  return WS_OK;
}
int ws_agg_fetch_begin(ws_agg* a, int64_t* groups) {
  return guard([&] {
    if (a->plan.nkeys == 0) {
      *groups = 1;
      a->fetched = 1;
      return;
    }
    a->fetched_host = false;
    if (a->is_dense && !a->part) {
      status(a);  // (small: the accumulators come with it)
      if (a->host_acc_ok) {
        // the groups with rows (COUNT slot > 0, as compact_dense), keys from the group index
        const ws_dense_spec& ds = a->dense;
        const int nk = a->plan.nkeys, na = a->plan.naggs;
        a->host_keys.clear();
        a->host_accs.clear();
        for (int64_t g = 0; g < ds.G; ++g) {
          const double* row = a->host_acc.data() + g * na;
          if (!(row[0] > 0)) continue;
          for (int c = 0; c < nk; ++c) {
            const int64_t v = ds.kmin[c] + (g / ds.stride[c]) % ds.range[c];
            double kv;
            if (a->plan.key_int[c]) std::memcpy(&kv, &v, 8);
            else kv = (double)v;
            a->host_keys.push_back(kv);
          }
          a->host_accs.insert(a->host_accs.end(), row, row + na);
        }
        a->fetched_host = true;
        a->fetched = na ? (int64_t)(a->host_accs.size() / na) : 0;
        *groups = a->fetched;
        return;
      }
    }
    a->fetched = agg_compact_into(a, a->fetch_buf, false);
    *groups = a->fetched;
  });
}
int ws_agg_fetch_copy(ws_agg* a, double* keys, double* accs) {
  return guard([&] {
    const int na = a->plan.naggs, nk = a->plan.nkeys;
    if (nk == 0) {
      a->hi.assign(na, 0);
      if (!a->nk_out) {  // no batches: identities
        for (int i = 0; i < na; ++i) accs[i] = host_identity(a->plan.ops[i]);
        return;
      }
      double h[2 * WS_MAX_AGGS];
      d2h_sync(h, a->nk_out, sizeof h);
      std::memcpy(accs, h, sizeof(double) * na);
      std::memcpy(a->hi.data(), h + WS_MAX_AGGS, 8 * na);
      return;
    }
    if (a->fetched < 0) throw Error("fetch_copy without fetch_begin", WS_ERR_INTERNAL);
    const int64_t g = a->fetched;
    if (a->fetched_host) {
      if (g) {
        std::memcpy(keys, a->host_keys.data(), sizeof(double) * g * nk);
        std::memcpy(accs, a->host_accs.data(), sizeof(double) * g * na);
      }
      return;
    }
    if (g) {
      WS_CK(hipMemcpyAsync(keys, a->fetch_buf.p, sizeof(double) * g * nk, hipMemcpyDeviceToHost, stream()));
      WS_CK(hipMemcpyAsync(accs, a->fetch_buf.p + g * nk, sizeof(double) * g * na, hipMemcpyDeviceToHost, stream()));
      WS_CK(hipStreamSynchronize(stream()));
    }
    a->fetch_buf.reset();
  });
}
int ws_agg_sum_bits_slots(const ws_agg* a, uint8_t* out) {
  const ws_agg_plan& p = a->plan;
  for (int i = 0; i < p.naggs; ++i)
    out[i] = p.nkeys ? p.agg_bits[i]
                     : (uint8_t)((p.ops[i] == WS_AGG_SUM || p.ops[i] == WS_AGG_MIN || p.ops[i] == WS_AGG_MAX) && p.agg_int[i]);
  return WS_OK;
}
int ws_agg_sum_high(const ws_agg* a, int64_t* out) {
  for (int i = 0; i < a->plan.naggs; ++i) out[i] = i < (int)a->hi.size() ? a->hi[i] : 0;
  return WS_OK;
}
int ws_agg_compact(ws_agg* a, double** keys, double** accs, int64_t* groups) {
  return guard([&] {
    DevBuf<double> k, ac;  // two separately owned buffers (the caller frees each)
    const int64_t g = agg_compact_into(a, k, true, &ac);
    *keys = k.release();
    *accs = ac.release();
    *groups = g;
  });
}
// Many dense aggregations of one plan shape in one launch (see batch_dense_body). Each aggregator
// keeps its own batches, environment (constants, columns) and accumulators.
int ws_agg_run_batch(int n, ws_agg* const* aggs, const ws_expr_env* const* envs, const ws_jit_program* jit) {
  return guard([&] {
    if (n <= 0) return;
    BatchBuild b = build_batch(n, aggs, envs, jit, false, [](size_t bytes) { return (uint8_t*)tmp_alloc(bytes); });
    if (b.ntiles == 0) {
      tmp_free(b.dev);
      return;
    }
    h2d_async_staged(b.dev, b.blob.host.data(), b.blob.host.size());
    launch_batch(b, n, true);
    for (int q = 0; q < n; ++q) aggs[q]->pending_init = false;
    tmp_free(b.dev);
  });
}

// Groups of many dense aggregations in one pass: counts[q] = groups of aggs[q]; keys (row-major,
// total x nkeys) and accs (total x naggs) hold them query after query, ascending group index.
// Fails (with counts set and *total the size needed) when the total exceeds cap.
int ws_agg_fetch_many(int n, ws_agg* const* aggs, int64_t* counts, int64_t* total, int64_t cap, double* keys, double* accs) {
  return guard([&] {
    *total = 0;
    if (n <= 0) return;
    const int nk = aggs[0]->plan.nkeys, na = aggs[0]->plan.naggs;
    std::vector<FetchQuery> fq(n);
    for (int q = 0; q < n; ++q) {
      ws_agg* A = aggs[q];
      prepare(A);  // never run: its identities (no groups)
      if (!A->is_dense || A->plan.nkeys != nk || A->plan.naggs != na)
        throw Error("batched fetch: every aggregation must be dense with the same plan shape", WS_ERR_INTERNAL);
      fq[q] = FetchQuery{A->acc, A->dense, A->dense.G};
    }
    std::vector<int64_t> cb;
    const std::vector<FetchChunk> ch = fetch_chunks(fq, cb);
    const int64_t nch = (int64_t)ch.size();
    if (nch == 0) {
      for (int q = 0; q < n; ++q) counts[q] = 0;
      return;
    }
    DevBuf<FetchQuery> dq(n);
    DevBuf<FetchChunk> dch(nch);
    h2d_async_staged(dq.p, fq.data(), sizeof(FetchQuery) * n);
    h2d_async_staged(dch.p, ch.data(), sizeof(FetchChunk) * nch);
    DevBuf<uint32_t> cnt(nch), offs(nch + 1);
    hipLaunchKernelGGL(dense_count_many_k, dim3((unsigned)nch), dim3(256), 0, stream(), (const FetchQuery*)dq.p,
                       (const FetchChunk*)dch.p, na, cnt.p);
    hipLaunchKernelGGL(part_start_k, dim3(1), dim3(1024), 0, stream(), (const uint32_t*)cnt.p, offs.p, (int)nch);
    check_launch("fetch_many_count");
    std::vector<uint32_t> h(nch + 1);
    d2h_sync(h.data(), offs.p, sizeof(uint32_t) * (nch + 1));
    for (int q = 0; q < n; ++q) counts[q] = (int64_t)(h[cb[q + 1]] - h[cb[q]]);
    *total = h[nch];
    if (*total > cap) throw Error("batched fetch: more groups than the buffers hold", WS_ERR_INTERNAL);
    if (!*total) return;
    DevBuf<double> pk(*total * nk), pa(*total * na);
    hipLaunchKernelGGL(dense_pack_many_k, dim3((unsigned)nch), dim3(256), 0, stream(), (const FetchQuery*)dq.p,
                       (const FetchChunk*)dch.p, aggs[0]->plan, nk, na, (const uint32_t*)offs.p, pk.p, pa.p);
    check_launch("fetch_many_pack");
    // straight into the caller's arrays (pinned ones, e.g. from ws_pinned_alloc, copy at link speed)
    if (nk) WS_CK(hipMemcpyAsync(keys, pk.p, sizeof(double) * (size_t)(*total * nk), hipMemcpyDeviceToHost, stream()));
    if (na) WS_CK(hipMemcpyAsync(accs, pa.p, sizeof(double) * (size_t)(*total * na), hipMemcpyDeviceToHost, stream()));
    WS_CK(hipStreamSynchronize(stream()));
  });
}

int ws_agg_batch_prepare(int n, ws_agg* const* aggs, const ws_expr_env* const* envs, const ws_jit_program* jit,
                         ws_agg_batch_plan** out) {
  return guard([&] {
    *out = nullptr;
    if (n <= 0 || n > 4096) throw Error("prepared batch: 1 to 4096 aggregations", WS_ERR_INTERNAL);
    auto P = std::make_unique<ws_agg_batch_plan>();
    P->n = n;
    P->nk = aggs[0]->plan.nkeys;
    P->na = aggs[0]->plan.naggs;
    P->plan = aggs[0]->plan;
    P->b = build_batch(n, aggs, envs, jit, true, [](size_t bytes) { return (uint8_t*)tmp_alloc(bytes); });
    if (P->b.ntiles == 0) throw Error("prepared batch: no rows", WS_ERR_INTERNAL);
    h2d_sync(P->b.dev, P->b.blob.host.data(), P->b.blob.host.size());
    std::vector<FetchQuery> fq(n);
    for (int q = 0; q < n; ++q) {
      fq[q] = FetchQuery{aggs[q]->acc, aggs[q]->dense, aggs[q]->dense.G};
      P->gcap += aggs[q]->dense.G;
      aggs[q]->pending_init = false;  // every replay gives the tables their initial values
    }
    const std::vector<FetchChunk> ch = fetch_chunks(fq, P->cb);
    P->nch = (int64_t)ch.size();
    if (P->nch == 0) throw Error("prepared batch: no groups", WS_ERR_INTERNAL);
    P->fq = DevBuf<FetchQuery>(n);
    h2d_sync(P->fq.p, fq.data(), sizeof(FetchQuery) * n);
    P->ch = DevBuf<FetchChunk>(P->nch);
    h2d_sync(P->ch.p, ch.data(), sizeof(FetchChunk) * P->nch);
    P->cnt = DevBuf<uint32_t>(P->nch);
    P->offs = DevBuf<uint32_t>(P->nch + 1);
    P->pk = DevBuf<double>(std::max<int64_t>(P->gcap * P->nk, 1));
    P->pa = DevBuf<double>(std::max<int64_t>(P->gcap * P->na, 1));
    WS_CK(hipStreamSynchronize(stream()));
    P->res = plan_res_acquire((size_t)P->nch + 1);
    P->cs = P->res.cs;
    P->ev_in = P->res.ev_in;
    P->ev_cnt = P->res.ev_cnt;
    P->hoffs = P->res.hoffs;
    WS_CK(hipStreamBeginCapture(P->cs, hipStreamCaptureModeThreadLocal));
    try {
      StreamOverride so(P->cs);
      launch_batch(P->b, n, false);
      hipLaunchKernelGGL(dense_count_many_k, dim3((unsigned)P->nch), dim3(256), 0, stream(), (const FetchQuery*)P->fq.p,
                         (const FetchChunk*)P->ch.p, P->na, P->cnt.p);
      hipLaunchKernelGGL(part_start_k, dim3(1), dim3(1024), 0, stream(), (const uint32_t*)P->cnt.p, P->offs.p, (int)P->nch);
      hipLaunchKernelGGL(dense_pack_many_k, dim3((unsigned)P->nch), dim3(256), 0, stream(), (const FetchQuery*)P->fq.p,
                         (const FetchChunk*)P->ch.p, P->plan, P->nk, P->na, (const uint32_t*)P->offs.p, P->pk.p, P->pa.p);
      check_launch("batch_plan_capture");
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(P->cs, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    WS_CK(hipStreamEndCapture(P->cs, &P->graph));
    WS_CK(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    *out = P.release();
  });
}
// The halves of a replay (ws_agg_batch_replay is launch + finish): the graph and the read-back of
// the group offsets on the plan's stream, ordered after the calling thread's stream; then, on the
// host, the wait for those offsets and the copies of the groups.
static void batch_plan_launch(ws_agg_batch_plan* P) {
  if (P->launched) throw Error("prepared batch: launched again before its finish", WS_ERR_INTERNAL);
  WS_CK(hipEventRecord(P->ev_in, stream()));
  WS_CK(hipStreamWaitEvent(P->cs, P->ev_in, 0));
  WS_CK(hipGraphLaunch(P->exec, P->cs));
  WS_CK(hipMemcpyAsync(P->hoffs, P->offs.p, sizeof(uint32_t) * (size_t)(P->nch + 1), hipMemcpyDeviceToHost, P->cs));
  WS_CK(hipEventRecord(P->ev_cnt, P->cs));
  P->launched = true;
}
static void batch_plan_finish(ws_agg_batch_plan* P, int64_t* counts, int64_t* total, int64_t cap, double* keys, double* accs) {
  *total = 0;
  if (!P->launched) throw Error("prepared batch: finished without a launch", WS_ERR_INTERNAL);
  P->launched = false;
  WS_CK(hipEventSynchronize(P->ev_cnt));
  const uint32_t* h = P->hoffs;
  for (int q = 0; q < P->n; ++q) counts[q] = (int64_t)(h[P->cb[q + 1]] - h[P->cb[q]]);
  *total = h[P->nch];
  if (*total > cap) throw Error("prepared batch: more groups than the buffers hold", WS_ERR_INTERNAL);
  if (!*total) return;
  if (P->nk) WS_CK(hipMemcpyAsync(keys, P->pk.p, sizeof(double) * (size_t)(*total * P->nk), hipMemcpyDeviceToHost, P->cs));
  if (P->na) WS_CK(hipMemcpyAsync(accs, P->pa.p, sizeof(double) * (size_t)(*total * P->na), hipMemcpyDeviceToHost, P->cs));
  WS_CK(hipStreamSynchronize(P->cs));
}
int ws_agg_batch_replay(ws_agg_batch_plan* P, int64_t* counts, int64_t* total, int64_t cap, double* keys, double* accs) {
  return guard([&] {
    batch_plan_launch(P);
    batch_plan_finish(P, counts, total, cap, keys, accs);
  });
}
int ws_agg_batch_launch(ws_agg_batch_plan* P) {
  return guard([&] { batch_plan_launch(P); });
}
int ws_agg_batch_finish(ws_agg_batch_plan* P, int64_t* counts, int64_t* total, int64_t cap, double* keys, double* accs) {
  return guard([&] { batch_plan_finish(P, counts, total, cap, keys, accs); });
}
void ws_agg_batch_release(ws_agg_batch_plan* P) { delete P; }

int64_t ws_agg_capacity(const ws_agg* a) { return a->cap; }
int ws_agg_group_count(ws_agg* a, int64_t* out) {
  return guard([&] {
    settle(a);
    const u64 c = a->keys || a->is_dense ? status(a).count : 0;
    *out = (int64_t)c + (a->part ? a->part_groups : 0);
  });
}
int ws_agg_set_strategy(ws_agg* a, int strategy, int64_t expected_groups) {
  return guard([&] {
    if (strategy < WS_STRAT_AUTO || strategy > WS_STRAT_PARTITIONED) throw Error("agg: bad strategy", WS_ERR_INTERNAL);
    a->strategy = strategy;
    a->expected_groups = expected_groups;
  });
}
int ws_agg_partitioned(const ws_agg* a) { return a->part ? 1 : 0; }
int ws_tune_set(const char* key, int64_t value) {
  return guard([&] {
    for (TuneKnob& k : g_tune)
      if (key && std::strcmp(key, k.name) == 0) {
        if (value < k.lo || value > k.hi) throw Error(std::string("tune: value out of range for ") + key, WS_ERR_INTERNAL);
        k.value.store(value, std::memory_order_relaxed);
        return;
      }
    throw Error(std::string("tune: unknown knob ") + (key ? key : "(null)"), WS_ERR_INTERNAL);
  });
}
int ws_tune_get(const char* key, int64_t* value) {
  return guard([&] {
    for (TuneKnob& k : g_tune)
      if (key && std::strcmp(key, k.name) == 0) {
        *value = k.value.load(std::memory_order_relaxed);
        return;
      }
    throw Error(std::string("tune: unknown knob ") + (key ? key : "(null)"), WS_ERR_INTERNAL);
  });
}

// ------------------------------------------------------------------ selection

static int64_t select_launch(const ws_expr_env& env, int has_where, ws_prog where, const ScanInput& in, bool cols_from_env,
                             const ws_jit_program* jit, int fn, int32_t* out_idx, std::vector<int64_t>* per_batch) {
  ScanDesc d{};
  d.tile_rows = kSelTileRows;
  ws_agg_plan fp{};
  fp.has_where = (uint8_t)(has_where != 0);
  fp.where = where;
  int64_t ntiles = 0;
  for (int64_t n : in.n) ntiles += (n + d.tile_rows - 1) / d.tile_rows;
  if (ntiles == 0) {
    if (per_batch) per_batch->assign(in.nb, 0);
    return 0;
  }
  // Long inputs take the chunked kernel: a chunk of consecutive tiles per block (select_chunks_per_cu
  // chunks per CU, at most 170 tiles so its ballots fit 32 KiB of LDS), one look-back per chunk.
  const int per_tile = kSelChunkRows * (kBlock / 64);  // ballots per chunked tile (2048 rows)
  const int64_t want = (int64_t)cu_count() * tune(kTuneSelectChunks);
  const bool chunked = ntiles > 2 * want;
  int64_t nunits = ntiles;
  size_t lds = 0;
  // Compiled plans over all but short inputs take the ballot pass: one block of kStagedBlock threads
  // per CU with the plan's set bitmaps (their leading words) staged in LDS, since a lookup in a
  // bitmap beyond L1 is otherwise one L2 request per row; then a scan of the per-tile counts and an
  // expansion of the ballots. No look-back: its ready frontier advances about one wave of tiles per
  // L2 round trip, which bounds the single-pass kernels on inputs of many tiles.
  int64_t set_words = 0, rows = 0;
  for (int k = 0; k < WS_MAX_SETS; ++k)
    if (env.sets[k] && env.set_size[k] > 0) set_words += ((int64_t)env.set_size[k] + 31) / 32;
  for (int64_t n : in.n) rows += n;
  constexpr int kStagedBlock = 1024;
  const bool mask_pass = jit && rows >= tune(kTuneMaskMinRows) && jit_function(jit, kJitSelectStaged) != nullptr;
  if (mask_pass) {
    d.tile_rows = kSelMaskRows * kStagedBlock;
    ntiles = 0;
    for (int64_t n : in.n) ntiles += (n + d.tile_rows - 1) / d.tile_rows;
    nunits = ntiles;
    const int64_t room = ((int64_t)kLdsPerCu - jit_staged_static_lds(jit)) / 4;
    d.set_words = (int32_t)std::max<int64_t>(0, std::min<int64_t>(set_words, room));
    lds = 4 * (size_t)d.set_words;
  } else if (chunked) {
    d.tile_rows = kSelChunkRows * kBlock;
    ntiles = 0;
    for (int64_t n : in.n) ntiles += (n + d.tile_rows - 1) / d.tile_rows;
    d.chunk_tiles = (int32_t)std::min<int64_t>(85, (ntiles + want - 1) / want);  // ballots within 32 KiB
    nunits = (ntiles + d.chunk_tiles - 1) / d.chunk_tiles;
    lds = (size_t)d.chunk_tiles * per_tile * 12;
  }
  DescUpload up;
  build_desc(up, d, in, fp, nullptr, env, cols_from_env, 0, nullptr);
  // [tile (or chunk) states][counter][total][err][per-batch counts]
  DevBuf<u64> st(ntiles + 3 + in.nb);
  fill_u64(st.p, st.n, 0);
  unsigned* ctr = (unsigned*)(st.p + ntiles);
  u64* total = st.p + ntiles + 1;
  int* err = (int*)(st.p + ntiles + 2);
  u64* pb = st.p + ntiles + 3;
  void* jf = jit ? jit_function(jit, mask_pass ? kJitSelectStaged : chunked ? kJitSelectChunk : fn) : nullptr;
  if (mask_pass) {
    const int words = d.tile_rows / 64;  // one expansion thread per ballot word
    constexpr int nw = 1;  // counts per tile
    DevBuf<u64> masks(ntiles * words);
    DevBuf<uint32_t> cnt(ntiles);  // per-tile counts, then their exclusive prefix
    DevBuf<int64_t> trow(ntiles);  // per tile: first global row * 2 + word layout (1: raw four-row ballots)
    if (words > kSelExpandThreads || words % 4) throw Error("selection tile of " + std::to_string(d.tile_rows) + " rows", WS_ERR_INTERNAL);
    fill_u32(cnt.p, ntiles, 0);
    void* args[] = {&up.ddesc, &masks.p, &cnt.p, &trow.p};
    jit_launch_raw(jf, (unsigned)std::min<int64_t>(ntiles, cu_count()), kStagedBlock, lds, args,
                   jit_prof_name(jit, "select_mask"));
    ProfScope ps("select_expand");
    hipLaunchKernelGGL(scan_counts_k, dim3(1), dim3(1024), 0, stream(), cnt.p, ntiles * nw, total);
    hipLaunchKernelGGL(batch_counts_k, dim3(1), dim3(256), 0, stream(), (const ScanDesc*)up.ddesc, (const uint32_t*)cnt.p, nw,
                       (const u64*)total, pb);
    hipLaunchKernelGGL(select_expand_k, dim3((unsigned)ntiles), dim3(words), 0, stream(), (const u64*)masks.p,
                       (const int64_t*)trow.p, (const uint32_t*)cnt.p, nw, out_idx);
    check_launch("select_expand");
  } else if (jf) {
    void* args[] = {&up.ddesc, &st.p, &ctr, &out_idx, &pb, &total, &err};
    jit_launch_raw(jf, (unsigned)nunits, kBlock, lds, args, jit_prof_name(jit, "select_jit"));
  } else {
    ProfScope ps("select");
    const int iv = !has_where || prog_is_leaf(env, where) ? kIvCols : prog_is_simple(env, where) ? kIvLean : kIvFull;
    if (chunked)
      hipLaunchKernelGGL(IV_PICK(iv, select_chunk), dim3((unsigned)nunits), dim3(kBlock), lds, stream(), up.ddesc, st.p, ctr,
                         out_idx, pb, total, err);
    else
      hipLaunchKernelGGL(IV_PICK(iv, select), dim3((unsigned)ntiles), dim3(kBlock), 0, stream(), up.ddesc, st.p, ctr,
                         out_idx, pb, total, err);
    check_launch("select");
  }
  std::vector<u64> h(3 + in.nb);
  d2h_sync(h.data(), st.p + ntiles, sizeof(u64) * h.size());
  if ((int)h[2]) throw Error("selection look-back did not complete", WS_ERR_INTERNAL);
  if (per_batch) {
    per_batch->resize(in.nb);
    for (int b = 0; b < in.nb; ++b) (*per_batch)[b] = (int64_t)h[3 + b];
  }
  return (int64_t)h[1];
}

int ws_select(const ws_expr_env* env, int has_where, ws_prog where, int64_t n, const uint32_t* del, int32_t* out_idx,
              const ws_jit_program* jit, int64_t* count) {
  return guard([&] {
    *count = 0;
    if (n <= 0) return;
    ScanInput in;
    in.nb = 1;
    in.ncols = WS_MAX_COLS;
    in.n = {n};
    in.del = {del};
    *count = select_launch(*env, has_where, where, in, true, jit, WS_JF_FLAG, out_idx, nullptr);
  });
}

int ws_select_multi(const ws_expr_env* env, int has_where, ws_prog where, int nb, const int64_t* bn,
                    const uint32_t* const* bdel, const ws_colview* bcols, int ncols, int32_t** out_idx,
                    int64_t* per_batch, const ws_jit_program* jit, int64_t* count) {
  return guard([&] {
    *count = 0;
    *out_idx = nullptr;
    ScanInput in;
    in.nb = nb;
    in.ncols = ncols;
    int64_t total = 0;
    for (int b = 0; b < nb; ++b) {
      in.n.push_back(bn[b]);
      in.del.push_back(bdel ? bdel[b] : nullptr);
      total += bn[b];
    }
    in.cols.assign(bcols, bcols + (size_t)nb * ncols);
    for (int b = 0; b < nb; ++b) per_batch[b] = 0;
    if (total <= 0) return;
    DevBuf<int32_t> out(total);
    std::vector<int64_t> pb;
    *count = select_launch(*env, has_where, where, in, false, jit, WS_JF_FLAGM, out.p, &pb);
    for (int b = 0; b < nb; ++b) per_batch[b] = pb[b];
    *out_idx = out.release();
  });
}

int ws_key_bitmap_where(const ws_expr_env* env, int has_where, ws_prog where, ws_prog key, int nb, const int64_t* bn,
                        const uint32_t* const* bdel, const ws_colview* bcols, int ncols, int64_t lo, int64_t nbits,
                        uint32_t* bits, uint64_t* count, const ws_jit_program* jit) {
  return guard([&] {
    if (nbits <= 0 || nbits > ((int64_t)1 << 31)) throw Error("ws_key_bitmap_where: bitmap of " + std::to_string(nbits) + " bits", WS_ERR_INTERNAL);
    if (!codegen::program_is_int(*env, key)) throw Error("ws_key_bitmap_where: the key must be an integer expression", WS_ERR_INTERNAL);
    fill_u32(bits, (nbits + 31) / 32, 0);
    ScanInput in;
    in.nb = nb;
    in.ncols = ncols;
    int64_t total = 0;
    for (int b = 0; b < nb; ++b) {
      in.n.push_back(bn[b]);
      in.del.push_back(bdel ? bdel[b] : nullptr);
      total += bn[b];
    }
    in.cols.assign(bcols, bcols + (size_t)nb * ncols);
    if (total <= 0) {
      if (count) fill_u64((u64*)count, 1, 0);
      return;
    }
    ws_agg_plan plan{};
    plan.has_where = has_where ? 1 : 0;
    plan.where = where;
    plan.nkeys = 1;
    plan.keys[0] = key;
    plan.key_int[0] = 1;
    // kKeyBitsRows rows per thread and tile (their loads issue together), a block per tile up to
    // 16 per CU.
    ScanDesc d{};
    d.tile_rows = kKeyBitsRows * kBlock;
    int64_t ntiles = 0;
    for (int64_t n : in.n) ntiles += (n + d.tile_rows - 1) / d.tile_rows;
    DescUpload up;
    build_desc(up, d, in, plan, nullptr, *env, false, 0, nullptr);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)cu_count() * 16));
    long long lo_ = lo, nbits_ = nbits;
    DevBuf<u64> parts(count ? grid : 0);  // per-block counts of passing rows, summed into *count
    unsigned long long* pp = (unsigned long long*)parts.p;
    if (void* jf = jit ? jit_function(jit, kJitKeyBits) : nullptr) {
      void* args[] = {&up.ddesc, &lo_, &nbits_, &bits, &pp};
      jit_launch_raw(jf, grid, kBlock, 0, args, jit_prof_name(jit, "key_bitmap_where"));
    } else {
      const int iv = interp_of(*env, plan);
      ProfScope ps("key_bitmap_where");
      hipLaunchKernelGGL(IV_PICK(iv, key_bits), dim3(grid), dim3(kBlock), 0, stream(), up.ddesc, lo_, nbits_, bits, pp);
      check_launch("key_bitmap_where");
    }
    if (count) {
      hipLaunchKernelGGL(sum_u64_k, dim3(1), dim3(1024), 0, stream(), (const u64*)parts.p, (int64_t)grid, (u64*)count);
      check_launch("key_bitmap_where_count");
    }
  });
}

}  // extern "C"

// ------------------------------------------------------------------ selection without allocation

// ws_select for callers that cannot allocate device memory (code under graph capture, serving
// loops that forbid it): the scan descriptor and the look-back state live in `scratch`, a
// file-scope device array that every GPU holds for the life of the process, instead of in pool
// blocks; the read-back goes through the caller's stack. The interpreted selection kernels are
// shared with ws_select. All calls on one GPU share that GPU's array: calls that overlap on one
// GPU (two host threads, or two streams) race and must be serialized by the caller.
namespace ws {
namespace {
constexpr size_t kScratchBytes = (size_t)2 << 20;     // 2 MiB per GPU
constexpr size_t kScratchDescBytes = 16 << 10;        // descriptor region at the front
constexpr int64_t kNoAllocMaxRows = (int64_t)1 << 28; // 268,435,456 rows
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// The current GPU's copy of `scratch` as a device address (looked up once per GPU, with that GPU
// current). Host code never uses the array's name as an address: only this pointer.
uint8_t* scratch_address() {
  int dev = 0;
  WS_CK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error("ws_select_noalloc: device ordinal out of range", WS_ERR_INTERNAL);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      throw Error(std::string("ws_select_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e), WS_ERR_INTERNAL);
    }
    g_scratch_ptr[dev] = (uint8_t*)p;
  }
  return g_scratch_ptr[dev];
}

// Host image of the descriptor region: ScanDesc, plan, env, then the one-batch arrays, each at a
// 16-byte boundary, with device addresses resolved against the region's base.
struct NoAllocDesc {
  alignas(16) uint8_t host[kScratchDescBytes];
  size_t used = 0;
  size_t put(const void* src, size_t bytes) {
    const size_t off = (used + 15) & ~(size_t)15;
    if (off + bytes > sizeof host) throw Error("ws_select_noalloc: descriptor region overflow", WS_ERR_INTERNAL);
    if (src) std::memcpy(host + off, src, bytes);
    else std::memset(host + off, 0, bytes);
    used = off + bytes;
    return off;
  }
};

int64_t select_noalloc(const ws_expr_env& env, int has_where, ws_prog where, int64_t n, const uint32_t* del,
                       int32_t* out_idx) {
  if (n > kNoAllocMaxRows)
    throw Error("ws_select_noalloc: " + std::to_string(n) + " rows exceed the maximum of " + std::to_string(kNoAllocMaxRows) +
                    " (the scratch array holds the look-back state of at most that many)",
                WS_ERR_INTERNAL);
  uint8_t* base = scratch_address();
  // tile geometry as ws_select chooses it for interpreted plans
  const int64_t want = (int64_t)cu_count() * tune(kTuneSelectChunks);
  int64_t ntiles = (n + kSelTileRows - 1) / kSelTileRows;
  const bool chunked = ntiles > 2 * want;
  ScanDesc d{};
  d.tile_rows = kSelTileRows;
  int64_t nunits = ntiles;
  size_t lds = 0;
  if (chunked) {
    d.tile_rows = kSelChunkRows * kBlock;
    ntiles = (n + d.tile_rows - 1) / d.tile_rows;
    d.chunk_tiles = (int32_t)std::min<int64_t>(85, (ntiles + want - 1) / want);
    nunits = (ntiles + d.chunk_tiles - 1) / d.chunk_tiles;
    lds = (size_t)d.chunk_tiles * kSelChunkRows * (kBlock / 64) * 12;
  }
  // [tile (or chunk) states][counter][total][err][per-batch count]
  const int64_t nstate = ntiles + 4;
  if (kScratchDescBytes + (size_t)nstate * 8 > kScratchBytes)
    throw Error("ws_select_noalloc: the look-back state of " + std::to_string(ntiles) + " tiles exceeds the scratch array",
                WS_ERR_INTERNAL);
  ws_agg_plan fp{};
  fp.has_where = (uint8_t)(has_where != 0);
  fp.where = where;
  NoAllocDesc h;
  const size_t o_desc = h.put(nullptr, sizeof(ScanDesc));
  const size_t o_plan = h.put(&fp, sizeof fp);
  const size_t o_env = h.put(&env, sizeof env);
  const size_t o_n = h.put(&n, 8);
  const size_t o_del = h.put(&del, 8);
  const int64_t tile0[2] = {0, ntiles}, row0 = 0;
  const size_t o_t0 = h.put(tile0, sizeof tile0);
  const size_t o_r0 = h.put(&row0, 8);
  d.nb = 1;
  d.ncols = WS_MAX_COLS;
  d.ntiles = ntiles;
  d.n = (const int64_t*)(base + o_n);
  d.del = (const uint32_t* const*)(base + o_del);
  d.tile0 = (const int64_t*)(base + o_t0);
  d.row0 = (const int64_t*)(base + o_r0);
  d.env = (const ws_expr_env*)(base + o_env);
  d.plan = (const ws_agg_plan*)(base + o_plan);
  d.cols = (const ws_colview*)(base + o_env + offsetof(ws_expr_env, cols));
  std::memcpy(h.host + o_desc, &d, sizeof d);
  WS_CK(hipMemcpyAsync(base, h.host, h.used, hipMemcpyHostToDevice, stream()));
  u64* st = (u64*)(base + kScratchDescBytes);
  fill_u64(st, nstate, 0);  // a fill kernel: the previous call's state is still in the array
  unsigned* ctr = (unsigned*)(st + ntiles);
  u64* total = st + ntiles + 1;
  int* err = (int*)(st + ntiles + 2);
  u64* pb = st + ntiles + 3;
  const ScanDesc* ddesc = (const ScanDesc*)(base + o_desc);
  const int iv = !has_where || prog_is_leaf(env, where) ? kIvCols : prog_is_simple(env, where) ? kIvLean : kIvFull;
  if (chunked)
    hipLaunchKernelGGL(IV_PICK(iv, select_chunk), dim3((unsigned)nunits), dim3(kBlock), lds, stream(), ddesc, st, ctr, out_idx,
                       pb, total, err);
  else
    hipLaunchKernelGGL(IV_PICK(iv, select), dim3((unsigned)ntiles), dim3(kBlock), 0, stream(), ddesc, st, ctr, out_idx, pb,
                       total, err);
  check_launch("select_noalloc");
  u64 back[3];  // total, err, per-batch count
  WS_CK(hipMemcpyAsync(back, total, sizeof back, hipMemcpyDeviceToHost, stream()));
  WS_CK(hipStreamSynchronize(stream()));
  if ((int)back[1]) throw Error("selection look-back did not complete", WS_ERR_INTERNAL);
  return (int64_t)back[0];
}
}  // namespace
}  // namespace ws

extern "C" {

int ws_select_noalloc(const ws_expr_env* env, int has_where, ws_prog where, int64_t n, const uint32_t* del, int32_t* out_idx,
                      int64_t* count) {
  return ws::guard([&] {
    *count = 0;
    if (n <= 0) return;
    *count = ws::select_noalloc(*env, has_where, where, n, del, out_idx);
  });
}

}  // extern "C"
