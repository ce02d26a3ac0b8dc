// Column statistics for sizing decisions: live and NULL row counts, min / max and HyperLogLog
// registers (p = 11, 2048 registers) of one column batch. A segment's statistics are computed once
// (at load) and merged per query; the sizing model reads them instead of the data.
//
// Distinct identity: integer types by exact int64 value (encoded columns decoded), doubles and
// DECIMAL's double view by the bit pattern of (v + 0.0), so -0.0 and 0.0 count once; NULL counts
// nothing. Register update (the published HyperLogLog algorithm, Flajolet et al. 2007): x =
// mix64(value bits), register x >> 53, rank = leading zeros of (x << 11 | 1 << 10) + 1.
#include <cmath>
#include <cstring>

#include "ws_internal.h"

namespace ws {

constexpr int kHllBits = 11, kHllRegs = 1 << kHllBits;

// Ordered key of a double for integer min / max (total order on non-NaN values).
__device__ __forceinline__ u64 dkey(double d) {
  const u64 b = (u64)dbits(d);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
__host__ __device__ __forceinline__ double dkey_inv(u64 k) {
  const u64 b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return bitsd((i64)b);
}

struct StatScalars {
  unsigned long long rows, nulls;
  long long imin, imax;
  unsigned long long kmin, kmax;  // dkey of the double min / max
};
__device__ __forceinline__ void stat_fold(StatScalars& a, const StatScalars& b) {
  a.rows += b.rows;
  a.nulls += b.nulls;
  a.imin = b.imin < a.imin ? b.imin : a.imin;
  a.imax = b.imax > a.imax ? b.imax : a.imax;
  a.kmin = b.kmin < a.kmin ? b.kmin : a.kmin;
  a.kmax = b.kmax > a.kmax ? b.kmax : a.kmax;
}
struct StatAcc {
  unsigned regs[kHllRegs];
  StatScalars s;
};

// One partial per block (registers as bytes, scalars as a StatScalars), all plain stores: the
// folding kernel combines them, so no two blocks touch the same address.
__global__ void __launch_bounds__(256) col_stats_k(ws_colview c, int64_t n, const uint32_t* del, StatScalars* part_s,
                                                   uint8_t* part_regs) {
  __shared__ unsigned regs[kHllRegs];
  __shared__ StatScalars s_w[4];
  for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x) regs[i] = 0;
  __syncthreads();
  const bool is_int = c.t != WS_T_F64;
  const bool nullable = c.t == WS_T_I32N || c.t == WS_T_I64N || c.t == WS_T_DECIMAL;
  const i64 nul = c.t == WS_T_I32N ? (i64)WS_NULL_I32 : (i64)WS_NULL_I64;
  u64 rows = 0, nulls = 0, kmin = ~0ull, kmax = 0;
  i64 imin = kI64Max, imax = kI64Min;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    if (del && ((del[r >> 5] >> (r & 31)) & 1u)) continue;
    ++rows;
    u64 bits;
    double dv;
    if (is_int) {
      const i64 v = col_int(c, r);
      if (nullable && v == nul) {
        ++nulls;
        continue;
      }
      bits = (u64)v;
      imin = v < imin ? v : imin;
      imax = v > imax ? v : imax;
      dv = c.t == WS_T_DECIMAL ? (double)v / pow10i(c.scale) : (double)v;
    } else {
      dv = col_dbl(c, r);
      if (dv != dv) {
        ++nulls;
        continue;
      }
      bits = (u64)dbits(dv + 0.0);
    }
    const u64 k = dkey(dv + 0.0);
    kmin = k < kmin ? k : kmin;
    kmax = k > kmax ? k : kmax;
    const u64 x = mix64(bits);
    const unsigned idx = (unsigned)(x >> (64 - kHllBits));
    const unsigned rank = (unsigned)__clzll((long long)((x << kHllBits) | (1ull << (kHllBits - 1)))) + 1;
    __hip_atomic_fetch_max(&regs[idx], rank, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // wave reductions, then one atomic per wave per field
  for (int o = 32; o > 0; o >>= 1) {
    rows += __shfl_xor(rows, o, 64);
    nulls += __shfl_xor(nulls, o, 64);
    const u64 a = __shfl_xor(kmin, o, 64), b = __shfl_xor(kmax, o, 64);
    kmin = a < kmin ? a : kmin;
    kmax = b > kmax ? b : kmax;
    const i64 e = __shfl_xor(imin, o, 64), f = __shfl_xor(imax, o, 64);
    imin = e < imin ? e : imin;
    imax = f > imax ? f : imax;
  }
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = StatScalars{rows, nulls, imin, imax, kmin, kmax};
  __syncthreads();
  if (threadIdx.x == 0) {
    StatScalars b = s_w[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) stat_fold(b, s_w[w]);
    part_s[blockIdx.x] = b;
  }
  // registers: one partial per block (plain stores), folded by stat_regs_k
  uint8_t* mine = part_regs + (int64_t)blockIdx.x * kHllRegs;
  for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x) mine[i] = (uint8_t)regs[i];
}
// Folds the block partials: block x takes registers [64 (x mod 32), +64) of the partials
// x / 32, x / 32 + kFoldGroups, ... (its 4 waves split them, 8 independent loads in flight per
// lane), combines the waves through LDS and max-combines into the zeroed acc->regs (kFoldGroups
// atomics per register). Block 0 also folds the scalar partials.
constexpr int kRegsPerFold = 64, kFoldWaves = 4, kFoldGroups = 8;
__global__ void __launch_bounds__(256) stat_regs_k(const uint8_t* part_regs, const StatScalars* part_s, int nblocks,
                                                   StatAcc* acc) {
  __shared__ unsigned s_m[kFoldWaves][kRegsPerFold];
  __shared__ StatScalars s_f[256];
  if (blockIdx.x == 0) {  // the scalars: a strided fold per thread, then a tree in LDS
    StatScalars f{0, 0, kI64Max, kI64Min, ~0ull, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) stat_fold(f, part_s[b]);
    s_f[threadIdx.x] = f;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) stat_fold(s_f[threadIdx.x], s_f[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x == 0) acc->s = s_f[0];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = blockIdx.x % (kHllRegs / kRegsPerFold), grp = blockIdx.x / (kHllRegs / kRegsPerFold);
  const int i = rb * kRegsPerFold + lane;
  const int step = kFoldWaves * kFoldGroups;
  unsigned m = 0;
  int b = grp * kFoldWaves + wave;
  for (; b + 7 * step < nblocks; b += 8 * step) {
    unsigned v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part_regs[(int64_t)(b + u * step) * kHllRegs + i];
#pragma unroll
    for (int u = 0; u < 8; ++u) m = v[u] > m ? v[u] : m;
  }
  for (; b < nblocks; b += step) {
    const unsigned v = part_regs[(int64_t)b * kHllRegs + i];
    m = v > m ? v : m;
  }
  s_m[wave][lane] = m;
  __syncthreads();
  if (wave == 0) {
    for (int w = 1; w < kFoldWaves; ++w) m = s_m[w][lane] > m ? s_m[w][lane] : m;
    if (m) __hip_atomic_fetch_max(&acc->regs[i], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace ws

using namespace ws;

extern "C" {

int ws_column_stats(const ws_colview* col, int64_t n, const uint32_t* del, ws_col_stats* out) {
  return guard([&] {
    ensure_init();
    std::memset(out, 0, sizeof *out);
    out->imin = kI64Max;
    out->imax = kI64Min;
    out->dmin = out->dmax = std::nan("");
    if (n <= 0) return;
    if (col->t == WS_T_VEC) throw Error("column statistics: vector columns have none", WS_ERR_INTERNAL);
    DevBuf<StatAcc> acc(1);
    // about four blocks per CU; each block leaves its partial for one folding kernel
    const unsigned grid = grid_for(n, 256, 4);
    DevBuf<uint8_t> parts((int64_t)grid * kHllRegs);
    DevBuf<StatScalars> parts_s(grid);
    fill_u32((uint32_t*)acc.p, kHllRegs, 0);  // the registers (the first member)
    hipLaunchKernelGGL(col_stats_k, dim3(grid), dim3(256), 0, stream(), *col, n, del, parts_s.p, parts.p);
    hipLaunchKernelGGL(stat_regs_k, dim3(kHllRegs / kRegsPerFold * kFoldGroups), dim3(64 * kFoldWaves), 0, stream(),
                       (const uint8_t*)parts.p, (const StatScalars*)parts_s.p, (int)grid, acc.p);
    check_launch("column_stats");
    StatAcc h;
    d2h_sync(&h, acc.p, sizeof h);
    out->rows = (int64_t)h.s.rows;
    out->nulls = (int64_t)h.s.nulls;
    if (col->t != WS_T_F64 && h.s.imin <= h.s.imax) {
      out->imin = h.s.imin;
      out->imax = h.s.imax;
    }
    if (h.s.kmin != ~0ull) {
      out->dmin = dkey_inv(h.s.kmin);
      out->dmax = dkey_inv(h.s.kmax);
    }
    for (int i = 0; i < kHllRegs; ++i) out->hll[i] = (uint8_t)h.regs[i];
  });
}

void ws_col_stats_merge(ws_col_stats* dst, const ws_col_stats* src) {
  dst->rows += src->rows;
  dst->nulls += src->nulls;
  dst->imin = src->imin < dst->imin ? src->imin : dst->imin;
  dst->imax = src->imax > dst->imax ? src->imax : dst->imax;
  if (src->dmin == src->dmin) dst->dmin = dst->dmin == dst->dmin ? std::fmin(dst->dmin, src->dmin) : src->dmin;
  if (src->dmax == src->dmax) dst->dmax = dst->dmax == dst->dmax ? std::fmax(dst->dmax, src->dmax) : src->dmax;
  for (int i = 0; i < kHllRegs; ++i) dst->hll[i] = src->hll[i] > dst->hll[i] ? src->hll[i] : dst->hll[i];
}

// Raw HyperLogLog estimate with the small-range (linear counting) correction; with 64-bit hashes
// no large-range correction is needed. Standard error about 1.04 / sqrt(2048) = 2.3 %.
double ws_hll_estimate(const uint8_t* regs) {
  const double m = kHllRegs;
  double sum = 0;
  int zeros = 0;
  for (int i = 0; i < kHllRegs; ++i) {
    sum += std::ldexp(1.0, -(int)regs[i]);
    zeros += regs[i] == 0;
  }
  const double alpha = 0.7213 / (1.0 + 1.079 / m);
  const double e = alpha * m * m / sum;
  if (e <= 2.5 * m && zeros > 0) return m * std::log(m / zeros);
  return e;
}

}  // extern "C"
