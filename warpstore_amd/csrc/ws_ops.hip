// Gathers with inline decode, dictionary code remaps, key bitmaps, per-row evaluation and GROUP BY
// derived-table columns: the operator surface of reference src/gpu.h:142-201 (semantics of
// gpu.cu:149-163, 757-1087).
//
// Design for gfx950 (every kernel here is one HBM-bound pass):
//   - index gathers move four elements per lane per step: one 16-byte load of four indices, then
//     the four random reads, then the stores, so each lane keeps four reads in flight;
//   - wide elements (VECTOR rows) are copied by a lane group per row, so a row's words are read as
//     one coalesced run instead of one lane per word;
//   - gathers over many batches stage the batch starts and views in LDS and find each row's batch
//     from the previous row's (ascending selections), with a binary search in LDS otherwise;
//   - bitmaps are built in LDS when the block's share of rows is large against the bitmap, then
//     OR-flushed word by word (only words that gained bits); otherwise with global atomics that
//     skip bits already set;
//   - integer evaluation is classified once on the host (one column / exact int64 program /
//     general), so the device runs the matching form without per-row validation;
//   - distinct values are compacted on the device (word popcounts, one-block scan, expansion)
//     when the bitmap is large, so only the values cross PCIe.
#include <algorithm>
#include <climits>
#include <cmath>

#include "ws_internal.h"

namespace ws {
namespace {

constexpr int kBlock = 256;

// Grid-stride loop over [0, n) with a device lambda: the small elementwise contracts below.
template <class F>
__global__ void __launch_bounds__(kBlock) each_k(int64_t n, F f) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) f(i);
}
template <class F>
void launch_each(const char* name, int64_t n, F f) {
  if (n <= 0) return;
  ProfScope ps(name);
  hipLaunchKernelGGL(each_k<F>, dim3(grid_for(n)), dim3(kBlock), 0, stream(), n, f);
  check_launch(name);
}

// ------------------------------------------------------------------ index gathers

// dst[j] = src[idx[j]] (idx[j] < 0: the NULL value when kNullFill), four elements per lane.
template <class T, bool kNullFill, bool kVecIdx>
__global__ void __launch_bounds__(kBlock) gather4_k(const T* __restrict__ src, const int32_t* __restrict__ idx, int64_t m,
                                                    T* __restrict__ dst, T nul) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x * 4;
  for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; j < m; j += step) {
    int32_t ix[4];
    if (kVecIdx && j + 4 <= m) {
      const int4 q = *(const int4*)(idx + j);
      ix[0] = q.x, ix[1] = q.y, ix[2] = q.z, ix[3] = q.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) ix[q] = j + q < m ? idx[j + q] : -1;
    }
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (kNullFill || j + q >= m) && ix[q] < 0 ? nul : src[ix[q]];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (j + q < m) dst[j + q] = v[q];
  }
}
template <class T, bool kNullFill>
void gather4(const char* name, const T* src, const int32_t* idx, int64_t m, T* dst, T nul) {
  if (m <= 0) return;
  ProfScope ps(name);
  const unsigned grid = grid_for((m + 3) / 4);
  if (((uintptr_t)idx & 15) == 0)
    hipLaunchKernelGGL((gather4_k<T, kNullFill, true>), dim3(grid), dim3(kBlock), 0, stream(), src, idx, m, dst, nul);
  else
    hipLaunchKernelGGL((gather4_k<T, kNullFill, false>), dim3(grid), dim3(kBlock), 0, stream(), src, idx, m, dst, nul);
  check_launch(name);
}

// Rows of `words` 32-bit words: a group of G lanes (a power of two, G <= 64) copies one row, the
// lanes reading consecutive words of it. A negative index writes a NULL row (float NaN words, the
// vector type's NULL marker) when `nulls`, else zeros.
__global__ void __launch_bounds__(kBlock) gather_rows_k(const uint32_t* __restrict__ src, int words, int group_log2,
                                                        const int32_t* __restrict__ idx, int64_t m, uint32_t* __restrict__ dst,
                                                        int nulls) {
  const int G = 1 << group_log2;
  const int lane = threadIdx.x & (G - 1);
  const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> group_log2);
  for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> group_log2) + (threadIdx.x >> group_log2); r < m; r += groups) {
    const int32_t i = idx[r];
    uint32_t* d = dst + r * words;
    if (i < 0) {
      const uint32_t fill = nulls ? 0x7fc00000u : 0u;
      for (int w = lane; w < words; w += G) d[w] = fill;
    } else {
      const uint32_t* s = src + (int64_t)i * words;
      for (int w = lane; w < words; w += G) d[w] = s[w];
    }
  }
}

// Gather from the concatenation of nb batches (batch b = rows [boff[b], boff[b+1])). Each block
// stages the batch starts and column views in LDS (kStaged; up to kStageBatches of them, else they
// are read from global memory), and every lane takes four consecutive indices: the batch of the
// first comes from a binary search over the starts, each next one is tried at the previous one's
// batch first (selections are ascending, so that almost always hits). Encoded batches are decoded
// into the plain type t (NULL sentinels of the nullable types are kept as stored).
constexpr int kStageBatches = 512;
struct MultiSrc {
  const ws_colview* views;  // [nb], types already set to the nullable form of t
  const int64_t* boff;      // [nb]
  int nb, t, esize;
};
template <bool kStaged>
__global__ void __launch_bounds__(kBlock) gather_batches_k(MultiSrc s, const int32_t* __restrict__ idx, int64_t m,
                                                           uint8_t* __restrict__ dst) {
  __shared__ int64_t s_off[kStaged ? kStageBatches : 1];
  __shared__ ws_colview s_view[kStaged ? kStageBatches : 1];
  if (kStaged) {
    for (int i = threadIdx.x; i < s.nb; i += blockDim.x) {
      s_off[i] = s.boff[i];
      s_view[i] = s.views[i];
    }
    __syncthreads();
  }
  const int nb = s.nb;
  auto off = [&](int b) { return kStaged ? s_off[b] : s.boff[b]; };
  auto batch = [&](int64_t g, int hint) {  // the last batch starting at or before g
    if (hint >= 0 && off(hint) <= g && (hint + 1 >= nb || g < off(hint + 1))) return hint;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off(mid) <= g) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  };
  const int64_t step = (int64_t)gridDim.x * blockDim.x * 4;
  const bool vec = ((uintptr_t)idx & 15) == 0;
  for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; j < m; j += step) {
    int64_t g[4];
    if (vec && j + 4 <= m) {
      const int4 q = *(const int4*)(idx + j);
      g[0] = q.x, g[1] = q.y, g[2] = q.z, g[3] = q.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) g[q] = j + q < m ? idx[j + q] : 0;
    }
    int b = -1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (j + q >= m) break;
      b = batch(g[q], b);
      const ws_colview v = kStaged ? s_view[b] : s.views[b];
      const int64_t r = g[q] - off(b), o = j + q;
      if (v.enc) {
        if (s.t == WS_P_F64) ((double*)dst)[o] = enc_dbl(v, r);
        else if (s.t == WS_P_I32) ((int32_t*)dst)[o] = (int32_t)enc_int(v, r);
        else ((int64_t*)dst)[o] = enc_int(v, r);
      } else if (s.esize == 4) {
        ((int32_t*)dst)[o] = ((const int32_t*)v.p)[r];
      } else if (s.esize == 8) {
        ((int64_t*)dst)[o] = ((const int64_t*)v.p)[r];
      } else {
        const uint8_t* p = (const uint8_t*)v.p + r * s.esize;
        for (int k = 0; k < s.esize; ++k) dst[o * s.esize + k] = p[k];
      }
    }
  }
}
__global__ void __launch_bounds__(kBlock) gather_encoded_k(ws_colview v, int t, const int32_t* __restrict__ idx, int64_t m,
                                                           void* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx[i];
    if (t == WS_P_F64) ((double*)dst)[i] = enc_dbl(v, r);
    else if (t == WS_P_I32) ((int32_t*)dst)[i] = (int32_t)enc_int(v, r);
    else ((int64_t*)dst)[i] = enc_int(v, r);
  }
}
uint8_t nullable_type_of(int ptype) { return ptype == WS_P_I32 ? WS_T_I32N : ptype == WS_P_I64 ? WS_T_I64N : WS_T_F64; }

// ------------------------------------------------------------------ value bitmaps

// Bit (v - base) of every non-NULL value v in [base, base + nbits) of an integer column view
// (doubles truncated); *saw_null (optional) set when a NULL is seen. kLds: the block builds its
// bitmap in LDS (nbits <= 32 * lds_words) and ORs the words that gained bits into `bits`.
struct BitSrc {
  const void* p;
  int kind;  // 0 int32 (NULL = INT32_MIN), 1 int64 (NULL = INT64_MIN), 2 double (NaN = NULL)
};
__device__ __forceinline__ bool bit_value(const BitSrc& s, int64_t i, int64_t& v) {
  if (s.kind == 0) {
    const int32_t x = ((const int32_t*)s.p)[i];
    v = x;
    return x != WS_NULL_I32;
  }
  if (s.kind == 1) {
    v = ((const int64_t*)s.p)[i];
    return v != WS_NULL_I64;
  }
  const double d = ((const double*)s.p)[i];
  v = (int64_t)d;
  return d == d;
}
// Set bits of a bitmap: per-thread popcounts, a wave sum, one atomic per wave.
__global__ void __launch_bounds__(kBlock) bitmap_count_k(const uint32_t* bits, int64_t words, unsigned long long* total) {
  unsigned long long c = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    c += (unsigned)__popc(bits[i]);
  for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_down(c, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(total, c);
}
template <bool kLds>
__global__ void __launch_bounds__(kBlock) value_bits_k(BitSrc s, int64_t n, int64_t base, int64_t nbits, uint32_t* bits,
                                                       int* saw_null) {
  extern __shared__ uint32_t lbits[];
  const int64_t words = (nbits + 31) >> 5;
  if (kLds) {
    for (int64_t w = threadIdx.x; w < words; w += blockDim.x) lbits[w] = 0;
    __syncthreads();
  }
  bool nul = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t v;
    if (!bit_value(s, i, v)) {
      nul = true;
      continue;
    }
    const uint64_t k = (uint64_t)v - (uint64_t)base;  // out of range below base wraps above nbits
    if (k >= (uint64_t)nbits) continue;
    const uint32_t bit = 1u << (k & 31);
    if (kLds) {
      atomicOr(&lbits[k >> 5], bit);
    } else if (!(bits[k >> 5] & bit)) {
      atomicOr(&bits[k >> 5], bit);
    }
  }
  if (saw_null && __any(nul) && (threadIdx.x & (kWave - 1)) == 0) *saw_null = 1;
  if (kLds) {
    __syncthreads();
    for (int64_t w = threadIdx.x; w < words; w += blockDim.x)
      if (lbits[w]) atomicOr(&bits[w], lbits[w]);
  }
}
void value_bits(const char* name, BitSrc s, int64_t n, int64_t base, int64_t nbits, uint32_t* bits, int* saw_null) {
  if (n <= 0 || nbits <= 0) return;
  ProfScope ps(name);
  const int64_t words = (nbits + 31) / 32;
  const unsigned grid = grid_for(n);
  // LDS form when it fits (64 KiB per block) and a block's rows outnumber the words it flushes
  if (words * 4 <= 64 * 1024 && (int64_t)grid * words * 4 <= n)
    hipLaunchKernelGGL(value_bits_k<true>, dim3(grid), dim3(kBlock), (size_t)words * 4, stream(), s, n, base, nbits, bits,
                       saw_null);
  else
    hipLaunchKernelGGL(value_bits_k<false>, dim3(grid), dim3(kBlock), 0, stream(), s, n, base, nbits, bits, saw_null);
  check_launch(name);
}
BitSrc bit_src_of(const ws_colview& c) {
  const int k = (c.t == WS_T_I32 || c.t == WS_T_I32N) ? 0 : c.t == WS_T_F64 ? 2 : 1;
  return BitSrc{c.p, k};
}

// ------------------------------------------------------------------ integer evaluation

// How an integer expression is evaluated, decided once per call on the host from the program and
// its column types (reference gpu.cu:826-883 evaluates the same three forms, choosing per row):
//   kOneCol: a single non-double column, its stored integer (a non-nullable INT's INT32_MIN reads
//            as the int64 NULL, as the reference's single-column path does);
//   kExact:  integer columns (not DECIMAL), integer literals, + - * and negation in exact int64,
//            NULL propagating, overflow raising env->ovf;
//   kDouble: anything else, through the VM in doubles, rounded (NaN -> NULL).
enum IntForm { kOneCol = 0, kExact = 1, kDouble = 2 };
constexpr int kExactDepth = 16;

// one_col = false: a lone column is classified like any other program (the distinct-value
// bitmap reads every expression in the exact form, reference gpu.cu:890-901).
IntForm int_form(const ws_expr_env& env, ws_prog p, bool one_col = true) {
  if (p.len <= 0 || p.start < 0 || p.start + p.len > WS_MAX_CODE) throw Error("eval: program out of range", WS_ERR_INTERNAL);
  const ws_instr first = env.code[p.start];
  if (one_col && p.len == 1 && first.op == WS_OP_COL && first.arg < WS_MAX_COLS && env.cols[first.arg].t != WS_T_F64)
    return kOneCol;
  int depth = 0;
  for (int i = 0; i < p.len; ++i) {
    const ws_instr in = env.code[p.start + i];
    switch (in.op) {
      case WS_OP_COL: {
        if (in.arg >= WS_MAX_COLS) return kDouble;
        const int t = env.cols[in.arg].t;
        if (t == WS_T_F64 || t == WS_T_VEC || t == WS_T_DECIMAL || depth >= kExactDepth) return kDouble;
        ++depth;
        break;
      }
      case WS_OP_CONST:
        if (!in.is_int || depth >= kExactDepth) return kDouble;
        ++depth;
        break;
      case WS_OP_ADD: case WS_OP_SUB: case WS_OP_MUL:
        if (depth < 2) return kDouble;
        --depth;
        break;
      case WS_OP_NEG:
        if (depth < 1) return kDouble;
        break;
      default: return kDouble;
    }
  }
  return depth == 1 ? kExact : kDouble;
}

// The exact form over a program the host has validated (no bounds or type checks per row).
__device__ i64 eval_exact(const ws_expr_env& env, ws_prog p, i64 r) {
  i64 val[kExactDepth];
  uint32_t nulls = 0;  // bit d: stack entry d is NULL
  int d = 0;
  for (int i = 0; i < p.len; ++i) {
    const ws_instr in = env.code[p.start + i];
    if (in.op == WS_OP_COL) {
      const ws_colview& c = env.cols[in.arg];
      const i64 v = col_int(c, r);
      const bool nl = (c.t == WS_T_I32N && v == (i64)WS_NULL_I32) || (c.t == WS_T_I64N && v == (i64)WS_NULL_I64);
      val[d] = v;
      nulls = (nulls & ~(1u << d)) | ((uint32_t)nl << d);
      ++d;
    } else if (in.op == WS_OP_CONST) {
      val[d] = dbits(env.consts[in.arg]);
      nulls &= ~(1u << d);
      ++d;
    } else if (in.op == WS_OP_NEG) {
      if (!((nulls >> (d - 1)) & 1u)) val[d - 1] = sub_chk(0, val[d - 1], env.ovf);
    } else {
      --d;
      const bool nl = ((nulls >> (d - 1)) | (nulls >> d)) & 1u;
      if (!nl) {
        const i64 a = val[d - 1], b = val[d];
        val[d - 1] = in.op == WS_OP_ADD ? add_chk(a, b, env.ovf) : in.op == WS_OP_SUB ? sub_chk(a, b, env.ovf) : mul_chk(a, b, env.ovf);
      }
      nulls = (nulls & ~(1u << (d - 1))) | ((uint32_t)nl << (d - 1));
    }
  }
  return (nulls & 1u) ? (i64)WS_NULL_I64 : val[0];
}
__device__ __forceinline__ i64 eval_int_form(const ws_expr_env& env, ws_prog p, int form, i64 r) {
  if (form == kOneCol) {
    const ws_colview& c = env.cols[env.code[p.start].arg];
    const i64 v = col_int(c, r);
    return (v == (i64)WS_NULL_I32 && (c.t == WS_T_I32 || c.t == WS_T_I32N)) ? (i64)WS_NULL_I64 : v;
  }
  if (form == kExact) return eval_exact(env, p, r);
  const double d = eval_prog(env, env.cols, p, r);
  return d != d ? (i64)WS_NULL_I64 : llround(d);
}
__global__ void __launch_bounds__(kBlock) eval_int_k(const ws_expr_env* envp, ws_prog p, int form, const int32_t* idx, int64_t m,
                                                     int64_t* out) {
  const ws_expr_env& env = *envp;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    out[j] = eval_int_form(env, p, form, idx ? idx[j] : j);
}
__global__ void __launch_bounds__(kBlock) eval_dbl_k(const ws_expr_env* envp, ws_prog prog, const int32_t* idx, int64_t m,
                                                     double* out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    out[j] = eval_prog(*envp, envp->cols, prog, idx ? idx[j] : j);
}
// A program that is one column (ORDER BY an aggregate): the VM's column load, without its stack.
__global__ void __launch_bounds__(kBlock) eval_col_k(ws_colview c, const int32_t* idx, int64_t m, double* out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    out[j] = load_col(c, idx ? idx[j] : j);
}

// Distinct values of an exact integer expression in [base, base + span): bit (v - base) in a
// 64-bit-word bitmap (NULLs and values outside the span skipped).
__global__ void __launch_bounds__(kBlock) distinct_mark_k(const ws_expr_env* envp, ws_prog p, int form, const int32_t* idx,
                                                          int64_t m, int64_t base, int64_t span, u64* bits) {
  const ws_expr_env& env = *envp;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const i64 v = eval_int_form(env, p, form, idx ? idx[j] : j);
    if (v == (i64)WS_NULL_I64) continue;
    const uint64_t k = (uint64_t)v - (uint64_t)base;
    if (k >= (uint64_t)span) continue;
    const u64 bit = 1ull << (k & 63);
    if (!(bits[k >> 6] & bit)) atomicOr((unsigned long long*)&bits[k >> 6], bit);
  }
}
// Device compaction of the bitmap: counts per run of kRunWords words, a one-block scan of the
// counts, then each run writes its values (ascending) at its offset.
constexpr int kRunWords = 16;
__global__ void __launch_bounds__(kBlock) run_popc_k(const u64* bits, int64_t words, uint32_t* cnt) {
  const int64_t runs = (words + kRunWords - 1) / kRunWords;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < runs; r += (int64_t)gridDim.x * blockDim.x) {
    uint32_t c = 0;
    const int64_t w1 = std::min<int64_t>(words, (r + 1) * kRunWords);
    for (int64_t w = r * kRunWords; w < w1; ++w) c += (uint32_t)__popcll(bits[w]);
    cnt[r] = c;
  }
}
__global__ void __launch_bounds__(1024) run_scan_k(uint32_t* cnt, int64_t n, u64* total) { scan_counts_body(cnt, n, total); }
__global__ void __launch_bounds__(kBlock) run_expand_k(const u64* bits, int64_t words, const uint32_t* off, int64_t base,
                                                       int64_t cap, int64_t* out) {
  const int64_t runs = (words + kRunWords - 1) / kRunWords;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < runs; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t o = off[r];
    const int64_t w1 = std::min<int64_t>(words, (r + 1) * kRunWords);
    for (int64_t w = r * kRunWords; w < w1 && o < cap; ++w)
      for (u64 b = bits[w]; b && o < cap; b &= b - 1) out[o++] = base + w * 64 + __builtin_ctzll(b);
  }
}

// ------------------------------------------------------------------ derived-table columns

// One GROUP BY result column of a temp table (reference gpu.h:155-166 TempColSpec): compacted
// results are column-major (slot j of group r at [j * g + r]). The spec is uniform, so its
// decisions are taken once on the host into `form`:
//   0 key (double), 1 key (int64 bits), 2 slot value / div, 3 slot int64 bits, 4 split sum
//   (hi * 2^32 + lo), 5 count, 6 average.
struct TempForm {
  int form;
  int idx, hi, cnt;      // slot indexes (cnt < 0: no count slot)
  int sum_bits, has_hi;  // average: the sum slot holds int64 bits, split in two halves
  int out;               // 0 int32, 1 int64, 2 double
  double div, mul;
};
__device__ __forceinline__ void temp_value(const TempForm& f, const double* keys, const double* accs, int64_t g, int64_t r,
                                           bool& nul, bool& is_int, i64& iv, double& dv) {
  auto slot = [&](int j) { return accs[(int64_t)j * g + r]; };
  is_int = false;
  iv = 0;
  dv = 0;
  nul = f.form >= 2 && f.cnt >= 0 && slot(f.cnt) == 0;
  switch (f.form) {
    case 0: dv = keys[(int64_t)f.idx * g + r]; nul = dv != dv; break;
    case 1:
      iv = dbits(keys[(int64_t)f.idx * g + r]);
      is_int = true;
      nul = iv == (i64)WS_NULL_I64 || iv == (i64)WS_NULL_I32;
      break;
    default:
      if (nul) break;
      if (f.form == 2) {
        dv = slot(f.idx) / f.div;
        nul = dv != dv;
      } else if (f.form == 3) {
        iv = dbits(slot(f.idx)), is_int = true;
      } else if (f.form == 4) {
        iv = (i64)(((u64)dbits(slot(f.hi)) << 32) + (u64)dbits(slot(f.idx))), is_int = true;
      } else if (f.form == 5) {
        iv = (i64)slot(f.idx), is_int = true;
      } else {
        double sum = slot(f.idx);
        if (f.sum_bits) sum = f.has_hi ? (double)dbits(slot(f.hi)) * 4294967296.0 + (double)dbits(slot(f.idx)) : (double)dbits(slot(f.idx));
        dv = sum / f.div / slot(f.cnt);
      }
  }
  if (f.mul != 1 && !nul) {
    dv = (is_int ? (double)iv : dv) * f.mul;
    is_int = false;
  }
}
__global__ void __launch_bounds__(kBlock) temp_col_k(TempForm f, const double* keys, const double* accs, int64_t g, void* out,
                                                     long long* mnmx) {
  i64 lo = kI64Max, hi = kI64Min;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < g; r += (int64_t)gridDim.x * blockDim.x) {
    bool nul, is_int;
    i64 iv;
    double dv;
    temp_value(f, keys, accs, g, r, nul, is_int, iv, dv);
    if (f.out == 2) {
      ((double*)out)[r] = nul ? nullv() : is_int ? (double)iv : dv;
      continue;
    }
    const i64 v = nul ? 0 : is_int ? iv : llround(dv);
    if (!nul) lo = v < lo ? v : lo, hi = v > hi ? v : hi;
    if (f.out == 1) ((int64_t*)out)[r] = nul ? (i64)WS_NULL_I64 : v;
    else ((int32_t*)out)[r] = nul ? WS_NULL_I32 : (int32_t)v;
  }
  if (f.out == 2) return;
  lo = wave_combine_i(WS_AGG_MIN, lo);
  hi = wave_combine_i(WS_AGG_MAX, hi);
  if ((threadIdx.x & (kWave - 1)) == 0 && lo <= hi) {
    atomicMin(&mnmx[0], (long long)lo);
    atomicMax(&mnmx[1], (long long)hi);
  }
}
TempForm temp_form_of(const ws_temp_col_spec& s) {
  TempForm f{};
  f.idx = s.idx, f.hi = s.hi, f.cnt = s.src == 1 ? s.cnt : -1, f.out = s.out, f.div = s.div, f.mul = s.mul;
  if (s.src == 0) f.form = s.key_int ? 1 : 0;
  else if (s.kind == 1) f.form = 3;
  else if (s.kind == 2) f.form = 4;
  else if (s.kind == 3) f.form = 5;
  else if (s.kind == 4) f.form = 6, f.sum_bits = s.sum_bits, f.has_hi = s.hi >= 0;
  else f.form = 2;
  return f;
}

// Device copy of an env (evaluation kernels read it through a pointer: 4.4 KB is too big for a
// kernel argument).
DevBuf<ws_expr_env> env_on_device(const ws_expr_env& env) {
  DevBuf<ws_expr_env> d(1);
  h2d_async_staged(d.p, &env, sizeof env);
  return d;
}

// Caller-freed device memory (the reference returns cudaMalloc'd buffers for devFree).
void* dev_malloc(size_t bytes) {
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    throw Error("GPU out of memory", WS_ERR_OOM);
  }
  WS_CK(e);
  return p;
}

}  // namespace
}  // namespace ws

using namespace ws;

extern "C" {

int ws_gather_multi(const ws_colview* src, const int64_t* boff, int nb, int ptype, int esize, const int32_t* idx, int64_t m,
                    void* dst) {
  return guard([&] {
    if (m <= 0) return;
    if (nb <= 0) throw Error("gather_multi: no batches", WS_ERR_INTERNAL);
    std::vector<ws_colview> views(src, src + nb);
    for (auto& v : views) v.t = nullable_type_of(ptype);  // encoded batches keep their NULL sentinels
    Blob blob;
    const size_t o_v = blob.put(views.data(), sizeof(ws_colview) * (size_t)nb);
    const size_t o_o = blob.put(boff, 8 * (size_t)nb);
    uint8_t* d = upload_blob(blob);
    MultiSrc s{(const ws_colview*)(d + o_v), (const int64_t*)(d + o_o), nb, ptype, esize};
    {
      ProfScope ps("gather_multi");
      const unsigned grid = grid_for((m + 3) / 4);
      if (nb <= kStageBatches)
        hipLaunchKernelGGL(gather_batches_k<true>, dim3(grid), dim3(kBlock), 0, stream(), s, idx, m, (uint8_t*)dst);
      else
        hipLaunchKernelGGL(gather_batches_k<false>, dim3(grid), dim3(kBlock), 0, stream(), s, idx, m, (uint8_t*)dst);
      check_launch("gather_multi");
    }
    tmp_free(d);
  });
}
int ws_gather_enc(const ws_colview* v, int ptype, const int32_t* idx, int64_t m, void* dst) {
  return guard([&] {
    if (m <= 0) return;
    ws_colview w = *v;
    w.t = nullable_type_of(ptype);
    ProfScope ps("gather_enc");
    hipLaunchKernelGGL(gather_encoded_k, dim3(grid_for(m)), dim3(kBlock), 0, stream(), w, ptype, idx, m, dst);
    check_launch("gather_enc");
  });
}
int ws_gather(const void* src, int ptype, const int32_t* idx, int64_t m, void* dst) {
  return guard([&] {
    if (ptype == WS_P_I32) gather4<int32_t, false>("gather", (const int32_t*)src, idx, m, (int32_t*)dst, 0);
    else gather4<int64_t, false>("gather", (const int64_t*)src, idx, m, (int64_t*)dst, 0);  // I64 / F64: 8-byte copy
  });
}
int ws_gather_null(const void* src, int ptype, const int32_t* idx, int64_t m, void* dst) {
  return guard([&] {
    if (ptype == WS_P_I32) gather4<int32_t, true>("gather_null", (const int32_t*)src, idx, m, (int32_t*)dst, WS_NULL_I32);
    else if (ptype == WS_P_I64)
      gather4<int64_t, true>("gather_null", (const int64_t*)src, idx, m, (int64_t*)dst, (int64_t)WS_NULL_I64);
    else gather4<double, true>("gather_null", (const double*)src, idx, m, (double*)dst, nullv());
  });
}
int ws_gather_bytes(const void* src, int esize, const int32_t* idx, int64_t m, void* dst, int nulls) {
  return guard([&] {
    if (m <= 0) return;
    if (esize <= 0 || esize % 4) throw Error("gather_bytes: element size must be a positive multiple of 4", WS_ERR_INTERNAL);
    const int words = esize / 4;
    int lg = 0;
    while ((1 << lg) < words && lg < 6) ++lg;
    ProfScope ps("gather_bytes");
    const int64_t rows_per_block = kBlock >> lg;
    const unsigned grid = (unsigned)std::min<int64_t>((m + rows_per_block - 1) / rows_per_block, (int64_t)cu_count() * 16);
    hipLaunchKernelGGL(gather_rows_k, dim3(std::max(grid, 1u)), dim3(kBlock), 0, stream(), (const uint32_t*)src, words, lg, idx, m,
                       (uint32_t*)dst, nulls);
    check_launch("gather_bytes");
  });
}
int ws_remap_codes(const int32_t* in, int64_t n, const int32_t* map, int64_t map_n, int32_t* out) {
  return guard([&] {
    launch_each("remap_codes", n, [=] __device__(int64_t j) {
      const int32_t c = in[j];
      out[j] = c == WS_NULL_I32 ? WS_NULL_I32 : ((uint64_t)(int64_t)c < (uint64_t)map_n ? map[c] : -2);
    });
  });
}
int ws_bits_to_double(const double* src, int64_t n, double* dst) {
  return guard([&] { launch_each("bits_to_double", n, [=] __device__(int64_t j) { dst[j] = (double)dbits(src[j]); }); });
}
int ws_key_bitmap(ws_colview key, int64_t n, int64_t lo, int64_t nbits, uint32_t* bits) {
  return guard([&] {
    fill_u32(bits, (nbits + 31) / 32, 0);
    value_bits("key_bitmap", bit_src_of(key), n, lo, nbits, bits, nullptr);
  });
}
int ws_bitmap_count(const uint32_t* bits, int64_t words, int64_t* count) {
  return guard([&] {
    if (!count) throw Error("ws_bitmap_count: null count", WS_ERR_INTERNAL);
    *count = 0;
    if (words <= 0) return;
    DevBuf<unsigned long long> total(1);
    fill_u64(total.p, 1, 0);
    unsigned long long* tp = total.p;
    ProfScope ps("bitmap_count");
    hipLaunchKernelGGL(bitmap_count_k, dim3((unsigned)std::min<int64_t>(grid_for(words), 1024)), dim3(kBlock), 0, stream(), bits,
                       words, tp);
    check_launch("bitmap_count");
    unsigned long long h = 0;
    d2h_sync(&h, tp, sizeof h);
    *count = (int64_t)h;
  });
}
int ws_set_bits(uint32_t* bits, const int32_t* idx, int64_t m) {
  return guard([&] {
    launch_each("set_bits", m, [=] __device__(int64_t j) {
      const uint32_t i = (uint32_t)idx[j];
      atomicOr(&bits[i >> 5], 1u << (i & 31));
    });
  });
}
int ws_map_rank(double* keys, const int32_t* rank, int64_t m) {
  return guard([&] {
    launch_each("map_rank", m, [=] __device__(int64_t j) {
      const double k = keys[j];
      keys[j] = k != k ? -1.0 : (double)rank[(int64_t)k];  // NULL sorts first
    });
  });
}
int ws_eval_at(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, double* out) {
  return guard([&] {
    if (m <= 0) return;
    if (prog.len == 1 && prog.start >= 0 && prog.start < WS_MAX_CODE && env->code[prog.start].op == WS_OP_COL &&
        env->code[prog.start].arg < WS_MAX_COLS) {
      ProfScope ps("eval_at");
      hipLaunchKernelGGL(eval_col_k, dim3(grid_for(m)), dim3(kBlock), 0, stream(), env->cols[env->code[prog.start].arg], idx, m,
                         out);
      check_launch("eval_at");
      return;
    }
    DevBuf<ws_expr_env> d = env_on_device(*env);
    ProfScope ps("eval_at");
    hipLaunchKernelGGL(eval_dbl_k, dim3(grid_for(m)), dim3(kBlock), 0, stream(), (const ws_expr_env*)d.p, prog, idx, m, out);
    check_launch("eval_at");
  });
}
int ws_eval_at_int(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, int64_t* out) {
  return guard([&] {
    if (m <= 0) return;
    const int form = int_form(*env, prog);
    DevBuf<ws_expr_env> d = env_on_device(*env);
    ProfScope ps("eval_at_int");
    hipLaunchKernelGGL(eval_int_k, dim3(grid_for(m)), dim3(kBlock), 0, stream(), (const ws_expr_env*)d.p, prog, form, idx, m, out);
    check_launch("eval_at_int");
  });
}
int ws_distinct_ints(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, int64_t base, int64_t span,
                     int64_t* out, int64_t out_cap, int64_t* n_out) {
  return guard([&] {
    *n_out = 0;
    if (m <= 0 || span <= 0) return;
    const int form = int_form(*env, prog, false);
    if (form != kExact) return;  // only exact integer expressions have distinct values here
    const int64_t words = (span + 63) / 64;
    DevBuf<u64> bits(words);
    fill_u64(bits.p, words, 0);
    DevBuf<ws_expr_env> d = env_on_device(*env);
    {
      ProfScope ps("distinct_mark");
      hipLaunchKernelGGL(distinct_mark_k, dim3(grid_for(m)), dim3(kBlock), 0, stream(), (const ws_expr_env*)d.p, prog, form, idx,
                         m, base, span, bits.p);
      check_launch("distinct_mark");
    }
    if (words <= 16384) {  // up to 128 KiB: read the bitmap back and list its bits here
      std::vector<u64> h((size_t)words);
      d2h_sync(h.data(), bits.p, 8 * (size_t)words);
      int64_t k = 0;
      for (int64_t w = 0; w < words; ++w)
        for (u64 b = h[(size_t)w]; b; b &= b - 1) {
          if (k < out_cap) out[k] = base + w * 64 + __builtin_ctzll(b);
          ++k;
        }
      *n_out = k;
      return;
    }
    const int64_t runs = (words + kRunWords - 1) / kRunWords;
    DevBuf<uint32_t> cnt(runs);
    DevBuf<u64> total(1);
    DevBuf<int64_t> vals(std::max<int64_t>(1, std::min<int64_t>(out_cap, span)));
    {
      ProfScope ps("distinct_compact");
      hipLaunchKernelGGL(run_popc_k, dim3(grid_for(runs)), dim3(kBlock), 0, stream(), bits.p, words, cnt.p);
      hipLaunchKernelGGL(run_scan_k, dim3(1), dim3(1024), 0, stream(), cnt.p, runs, total.p);
      hipLaunchKernelGGL(run_expand_k, dim3(grid_for(runs)), dim3(kBlock), 0, stream(), bits.p, words, cnt.p, base, vals.n,
                         vals.p);
      check_launch("distinct_compact");
    }
    u64 k = 0;
    d2h_sync(&k, total.p, 8);
    const int64_t c = std::min<int64_t>((int64_t)k, std::min<int64_t>(out_cap, vals.n));
    if (c > 0) d2h_sync(out, vals.p, 8 * (size_t)c);
    *n_out = (int64_t)k;
  });
}
int ws_bits_from_column(const void* col, int out, int64_t n, int64_t base, int64_t nbits, int* saw_null, uint32_t** bits) {
  return guard([&] {
    *bits = nullptr;
    const size_t bytes = std::max<size_t>(4, (size_t)((nbits + 31) / 32) * 4);
    uint32_t* b = (uint32_t*)dev_malloc(bytes);
    try {
      fill_u32(b, (int64_t)(bytes / 4), 0);
      DevBuf<int> nul(1);
      fill_u32((uint32_t*)nul.p, 1, 0);
      value_bits("bits_from_column", BitSrc{col, out == 1 ? 1 : 0}, n, base, nbits, b, nul.p);
      int h = 0;
      d2h_sync(&h, nul.p, 4);
      *saw_null = h != 0;
    } catch (...) {
      (void)hipFree(b);
      throw;
    }
    *bits = b;
  });
}
int ws_temp_column(const double* keys, int nk, const double* accs, int na, int64_t g, const ws_temp_col_spec* s, double* mn,
                   double* mx, void** out) {
  return guard([&] {
    (void)nk, (void)na;
    *out = nullptr;
    const size_t w = s->out == 0 ? 4 : 8;
    void* o = dev_malloc(std::max<size_t>(1, (size_t)g * w));
    try {
      DevBuf<long long> mm(2);
      const long long init[2] = {LLONG_MAX, LLONG_MIN};
      h2d_async_staged(mm.p, init, sizeof init);
      if (g > 0) {
        ProfScope ps("temp_column");
        hipLaunchKernelGGL(temp_col_k, dim3(grid_for(g)), dim3(kBlock), 0, stream(), temp_form_of(*s), keys, accs, g, o, mm.p);
        check_launch("temp_column");
      }
      long long h[2];
      d2h_sync(h, mm.p, sizeof h);
      const bool none = s->out == 2 || h[0] > h[1];
      *mn = none ? -INFINITY : (double)h[0];
      *mx = none ? INFINITY : (double)h[1];
    } catch (...) {
      (void)hipFree(o);
      throw;
    }
    *out = o;
  });
}

}  // extern "C"
