// Equi hash joins (reference src/gpu.cu:1089-1316): open-addressing multimap build, then a probe
// in two passes over tiles of the probe side. Pass 1 counts matches per tile and learns each
// tile's output offset by decoupled look-back (no per-row count array, no separate scan); pass 2
// ranks matches inside the tile with a block scan and writes (build, probe) pairs. Output is in
// probe order, and within one probe row in probe-sequence order, like the reference.
#include <algorithm>
#include <climits>

#include "ws_internal.h"

namespace ws {

constexpr i64 kJEmpty = (i64)0x8000000000000000ll;
constexpr int kJBlock = 256;
constexpr int kJRows = 4;  // probe rows per thread per tile
constexpr int kJTile = kJRows * kJBlock;

struct JEntry {  // 16-byte slot: one load per probe step
  i64 key;
  int32_t row;
  int32_t pad;
};

__device__ __forceinline__ i64 jkey_at(const ws_colview& c, i64 r) {
  switch (c.t) {
    case WS_T_I32: case WS_T_I32N: return ((const int32_t*)c.p)[r];
    case WS_T_I64: case WS_T_I64N: case WS_T_DECIMAL: return ((const int64_t*)c.p)[r];
    default: return (i64)((const double*)c.p)[r];
  }
}
__device__ __forceinline__ bool jkey_null(const ws_colview& c, i64 r) {
  switch (c.t) {
    case WS_T_I32: case WS_T_I32N: return ((const int32_t*)c.p)[r] == WS_NULL_I32;
    case WS_T_I64: case WS_T_I64N: case WS_T_DECIMAL: return ((const int64_t*)c.p)[r] == WS_NULL_I64;
    default: {
      const double v = ((const double*)c.p)[r];
      return v != v;
    }
  }
}

__global__ void join_fill_k(JEntry* t, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    t[i].key = kJEmpty;
    t[i].row = -1;
    t[i].pad = 0;
  }
}
// NULL keys never enter the table (SQL equality with NULL is unknown).
__global__ void join_build_k(ws_colview key, int64_t base, int64_t nb, JEntry* t, u64 mask) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
    if (jkey_null(key, base + i)) continue;
    const i64 k = jkey_at(key, base + i);
    u64 s = mix64((u64)k) & mask;
    for (u64 step = 0; step <= mask; ++step) {
      u64 prev = (u64)kJEmpty;
      if (__hip_atomic_compare_exchange_strong((u64*)&t[s].key, &prev, (u64)k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)) {
        t[s].row = (int32_t)(base + i);
        break;
      }
      s = (s + 1) & mask;
    }
  }
}

// Probe tiles: tiles never straddle probe partitions; partition p's table is tables[tpart[p]].
struct ProbeDesc {
  ws_colview key;
  int np;                  // probe partitions
  int64_t ntiles;
  const int64_t* p_off;    // [np + 1] probe row offsets
  const int64_t* tile0;    // [np + 1] first tile of each partition
  const JEntry* const* tab;  // [np] table of each probe partition
  const u64* mask;         // [np]
  int outer;               // 1: unmatched probe rows produce (-1, row)
};

__device__ __forceinline__ int part_of_tile(const ProbeDesc& d, i64 t) {
  int lo = 0, hi = d.np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.tile0[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Matches of probe row j (chain order), calling emit(build_row) for each.
template <class F>
__device__ __forceinline__ int probe_row(const ProbeDesc& d, const JEntry* tab, u64 mask, i64 j, F&& emit) {
  if (jkey_null(d.key, j)) return 0;
  const i64 k = jkey_at(d.key, j);
  u64 s = mix64((u64)k) & mask;
  int c = 0;
  for (u64 step = 0; step <= mask; ++step) {
    const JEntry e = tab[s];
    if (e.key == kJEmpty) break;
    if (e.key == k) {
      emit(e.row);
      ++c;
    }
    s = (s + 1) & mask;
  }
  return c;
}

// Match counts of a thread's kJRows rows of a tile (rows r0 + jj * kJBlock + tid): every row's key
// and first table slot are loaded before any chain is walked, so the rows' memory round trips
// overlap instead of running one after another. fr (when given) gets each row's first match.
__device__ __forceinline__ void probe_counts(const ProbeDesc& d, const JEntry* tab, u64 mask, i64 r0, i64 r1, int (&m)[kJRows],
                                             int32_t* fr = nullptr) {
  i64 k[kJRows];
  u64 s[kJRows];
  bool ok[kJRows];
#pragma unroll
  for (int jj = 0; jj < kJRows; ++jj) {
    const i64 j = r0 + (i64)jj * kJBlock + threadIdx.x;
    ok[jj] = tab && j < r1;
    const i64 jc = ok[jj] ? j : r0;  // (a valid row; its result is dropped)
    ok[jj] = ok[jj] && !jkey_null(d.key, jc);
    k[jj] = jkey_at(d.key, jc);
    s[jj] = mix64((u64)k[jj]) & mask;
  }
  JEntry e[kJRows];
#pragma unroll
  for (int jj = 0; jj < kJRows; ++jj) e[jj] = ok[jj] ? tab[s[jj]] : JEntry{kJEmpty, 0};
#pragma unroll
  for (int jj = 0; jj < kJRows; ++jj) {
    int c = 0;
    int32_t f = -1;
    for (u64 step = 0; step <= mask && e[jj].key != kJEmpty; ++step) {
      if (e[jj].key == k[jj]) {
        if (c == 0) f = e[jj].row;
        ++c;
      }
      s[jj] = (s[jj] + 1) & mask;
      e[jj] = tab[s[jj]];
    }
    m[jj] = c;
    if (fr) fr[jj] = f;
  }
}

// Pass 1: tile match counts -> exclusive output offsets by look-back.
__global__ void __launch_bounds__(kJBlock) join_count_k(ProbeDesc d, u64* state, unsigned* ctr, u64* tile_off, u64* total,
                                                        int* err) {
  __shared__ unsigned s_tile;
  __shared__ u64 s_wsum[kJBlock / 64];
  if (threadIdx.x == 0) s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  const i64 t = s_tile;
  const int p = part_of_tile(d, t);
  const JEntry* tab = d.tab[p];
  const u64 mask = d.mask[p];
  const i64 r0 = d.p_off[p] + (t - d.tile0[p]) * kJTile;
  const i64 r1 = min((long long)(r0 + kJTile), (long long)d.p_off[p + 1]);
  u64 c = 0;
  int mc[kJRows];
  probe_counts(d, tab, mask, r0, r1, mc);
#pragma unroll
  for (int jj = 0; jj < kJRows; ++jj) {
    const i64 j = r0 + (i64)jj * kJBlock + threadIdx.x;
    c += (u64)(d.outer && j < r1 && mc[jj] == 0 ? 1 : mc[jj]);
  }
  c = wave_sum_u64(c);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_wsum[wave] = c;
  __syncthreads();
  if (wave == 0) {
    u64 tc = 0;
    for (int w = 0; w < kJBlock / 64; ++w) tc += s_wsum[w];
    const u64 excl = lookback_exclusive(state, t, tc, err);
    if (lane == 0) {
      tile_off[t] = excl;
      if (t == d.ntiles - 1) *total = excl + tc;
    }
  }
}

// Pass 2: per-row counts, block exclusive scan in row order, then the pairs.
__global__ void __launch_bounds__(kJBlock) join_write_k(ProbeDesc d, const u64* tile_off, int32_t* ob, int32_t* op,
                                                        uint8_t* matched) {
  __shared__ u64 s_wtot[kJBlock / 64];
  const i64 t = blockIdx.x;
  const int p = part_of_tile(d, t);
  const JEntry* tab = d.tab[p];
  const u64 mask = d.mask[p];
  const i64 r0 = d.p_off[p] + (t - d.tile0[p]) * kJTile;
  const i64 r1 = min((long long)(r0 + kJTile), (long long)d.p_off[p + 1]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u64 carry = tile_off[t];
  int mc[kJRows];
  probe_counts(d, tab, mask, r0, r1, mc);
  for (int jj = 0; jj < kJRows; ++jj) {
    const i64 j = r0 + (i64)jj * kJBlock + threadIdx.x;
    int m = j < r1 ? mc[jj] : 0;
    const bool fill = d.outer && j < r1 && m == 0;
    const u64 cnt = (u64)(fill ? 1 : m);
    // inclusive wave scan, then wave totals through LDS
    u64 incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      const u64 y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (lane == 63) s_wtot[wave] = incl;
    __syncthreads();
    u64 before = 0, all = 0;
    for (int w = 0; w < kJBlock / 64; ++w) {
      if (w < wave) before += s_wtot[w];
      all += s_wtot[w];
    }
    u64 o = carry + before + incl - cnt;
    if (fill) {
      ob[o] = -1;
      op[o] = (int32_t)j;
    } else if (m) {
      probe_row(d, tab, mask, j, [&](int32_t br) {
        ob[o] = br;
        op[o] = (int32_t)j;
        if (matched) matched[br] = 1;
        ++o;
      });
    }
    carry += all;
    __syncthreads();
  }
}

// Single-probe form, for joins where no probe row has a second match (unique build keys, the usual
// case): one probe per row instead of three, and no look-back. Pass 1 (a block per tile) stores each
// row's first match (or -1) and the tile's output count (an unmatched row yields one pair in an
// outer join), and raises *multi when a row has a second match: the caller then runs the two-pass
// form instead. A one-block scan turns the counts into tile offsets; pass 2 writes the pairs.
__global__ void __launch_bounds__(kJBlock) join_probe1_k(ProbeDesc d, int32_t* first, uint32_t* tile_cnt, int* multi) {
  __shared__ uint32_t s_w[kJBlock / 64];
  const i64 t = blockIdx.x;
  const int p = part_of_tile(d, t);
  const JEntry* tab = d.tab[p];
  const u64 mask = d.mask[p];
  const i64 r0 = d.p_off[p] + (t - d.tile0[p]) * kJTile;
  const i64 r1 = min((long long)(r0 + kJTile), (long long)d.p_off[p + 1]);
  int mc[kJRows];
  int32_t fm[kJRows];
  probe_counts(d, tab, mask, r0, r1, mc, fm);
  uint32_t c = 0;
  bool many = false;
#pragma unroll
  for (int jj = 0; jj < kJRows; ++jj) {
    const i64 j = r0 + (i64)jj * kJBlock + threadIdx.x;
    if (j < r1) {
      first[j] = fm[jj];
      c += d.outer && mc[jj] == 0 ? 1u : (uint32_t)mc[jj];
      many |= mc[jj] > 1;
    }
  }
  if (many) *multi = 1;
  c = (uint32_t)wave_sum_u64(c);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_w[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tc = 0;
    for (int w = 0; w < kJBlock / 64; ++w) tc += s_w[w];
    tile_cnt[t] = tc;
  }
}
__global__ void __launch_bounds__(1024) join_scan_k(uint32_t* cnt, int64_t n, u64* total) { scan_counts_body(cnt, n, total); }
__global__ void __launch_bounds__(kJBlock) join_write1_k(ProbeDesc d, const int32_t* first, const uint32_t* tile_off, int32_t* ob,
                                                         int32_t* op, uint8_t* matched, const u64* total, u64 cap) {
  __shared__ uint32_t s_wtot[kJBlock / 64];
  if (*total > cap) return;  // buffers sized ahead of the count (learned size) and too small: written later
  const i64 t = blockIdx.x;
  const int p = part_of_tile(d, t);
  const i64 r0 = d.p_off[p] + (t - d.tile0[p]) * kJTile;
  const i64 r1 = min((long long)(r0 + kJTile), (long long)d.p_off[p + 1]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int32_t f[kJRows];
#pragma unroll
  for (int jj = 0; jj < kJRows; ++jj) {
    const i64 j = r0 + (i64)jj * kJBlock + threadIdx.x;
    f[jj] = j < r1 ? first[j] : -1;
  }
  u64 carry = tile_off[t];
#pragma unroll
  for (int jj = 0; jj < kJRows; ++jj) {
    const i64 j = r0 + (i64)jj * kJBlock + threadIdx.x;
    const bool emit = j < r1 && (f[jj] >= 0 || d.outer);
    const u64 m = __ballot(emit);
    if (lane == 0) s_wtot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    u64 before = 0, all = 0;
    for (int w = 0; w < kJBlock / 64; ++w) {
      if (w < wave) before += s_wtot[w];
      all += s_wtot[w];
    }
    if (emit) {
      const u64 o = carry + before + (u64)__popcll(m & ((1ull << lane) - 1));
      ob[o] = f[jj];
      op[o] = (int32_t)j;
      if (matched && f[jj] >= 0) matched[f[jj]] = 1;
    }
    carry += all;
    __syncthreads();
  }
}
// Output offset of each probe partition from the scanned tile offsets (total past the last).
__global__ void join_offsets1_k(const uint32_t* tile_off, const int64_t* tile0, int np, int64_t ntiles, const u64* total,
                                int64_t* out) {
  for (int p = threadIdx.x; p <= np; p += blockDim.x) out[p] = tile0[p] < ntiles ? (int64_t)tile_off[tile0[p]] : (int64_t)*total;
}

__global__ void gather_off_k(const u64* tile_off, const int64_t* tile0, int np, int64_t ntiles, u64 total, int64_t* out) {
  for (int p = threadIdx.x; p <= np; p += blockDim.x) out[p] = tile0[p] < ntiles ? (int64_t)tile_off[tile0[p]] : (int64_t)total;
}
__global__ void unmatched_k(const uint8_t* matched, int64_t n, uint8_t* flags) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = !matched[i];
}
__global__ void cross_k(int64_t na, int64_t nb, int32_t* oa, int32_t* ob) {
  const int64_t n = na * nb;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    oa[i] = (int32_t)(i / nb);
    ob[i] = (int32_t)(i % nb);
  }
}

struct JoinResult {
  DevBuf<int32_t> b, p;
  int64_t m = 0;
  std::vector<int64_t> out_off;
};

// Builds one table per build partition and probes every probe partition against its table
// (pb == 1: the one table is shared). outer: every probe row appears; matched: flags of build rows.
// expect > 0 (a predicted output size, e.g. from the learned join model): the single-probe form
// allocates its pair buffers at that size and writes them without first reading the count back;
// a prediction below the real size costs one more write pass, never a wrong result.
static JoinResult join_core(ws_colview bkey, const std::vector<int64_t>& b_off, ws_colview pkey, const std::vector<int64_t>& p_off,
                            bool outer, uint8_t* matched, int64_t expect = 0) {
  JoinResult res;
  const int pb = (int)b_off.size() - 1, pp = (int)p_off.size() - 1;
  if (pb != 1 && pb != pp) throw Error("internal: partition count mismatch in join", WS_ERR_INTERNAL);
  const int64_t np = p_off[pp];
  res.out_off.assign(pp + 1, 0);
  if (np == 0) return res;
  std::vector<int64_t> reg(pb + 1, 0), cap(pb);
  for (int q = 0; q < pb; ++q) {
    int64_t c = 2;
    while (c < (b_off[q + 1] - b_off[q]) * 2) c <<= 1;
    cap[q] = c;
    reg[q + 1] = reg[q] + c;
  }
  DevBuf<JEntry> tab(reg[pb]);
  {
    ProfScope ps("join_build");
    hipLaunchKernelGGL(join_fill_k, dim3(grid_for(reg[pb])), dim3(256), 0, stream(), tab.p, reg[pb]);
    for (int q = 0; q < pb; ++q) {
      const int64_t n = b_off[q + 1] - b_off[q];
      if (n)
        hipLaunchKernelGGL(join_build_k, dim3(grid_for(n)), dim3(256), 0, stream(), bkey, b_off[q], n, tab.p + reg[q],
                           (u64)(cap[q] - 1));
    }
    check_launch("join_build");
  }
  // probe descriptor tables
  std::vector<int64_t> tile0(pp + 1, 0);
  for (int p = 0; p < pp; ++p) tile0[p + 1] = tile0[p] + (p_off[p + 1] - p_off[p] + kJTile - 1) / kJTile;
  const int64_t ntiles = tile0[pp];
  std::vector<const JEntry*> tp(pp);
  std::vector<u64> masks(pp);
  const int64_t nbuild = b_off[pb];
  for (int p = 0; p < pp; ++p) {
    const int q = pb == 1 ? 0 : p;
    tp[p] = nbuild ? tab.p + reg[q] : nullptr;
    masks[p] = (u64)(cap[q] - 1);
  }
  Blob bl;
  const size_t o_off = bl.put(p_off.data(), 8 * p_off.size());
  const size_t o_t0 = bl.put(tile0.data(), 8 * tile0.size());
  const size_t o_tab = bl.put(tp.data(), 8 * tp.size());
  const size_t o_mask = bl.put(masks.data(), 8 * masks.size());
  uint8_t* dblob = upload_blob(bl);
  struct Free {
    uint8_t* p;
    ~Free() { tmp_free(p); }
  } free_blob{dblob};
  ProbeDesc d;
  d.key = pkey;
  d.np = pp;
  d.ntiles = ntiles;
  d.p_off = (const int64_t*)(dblob + o_off);
  d.tile0 = (const int64_t*)(dblob + o_t0);
  d.tab = (const JEntry* const*)(dblob + o_tab);
  d.mask = (const u64*)(dblob + o_mask);
  d.outer = outer ? 1 : 0;
  if (ntiles == 0) return res;
  {
    // single-probe form: [total][multi][out_off (pp + 1)] read back in one round trip
    DevBuf<int32_t> first(np);
    DevBuf<uint32_t> tcnt(ntiles);
    DevBuf<u64> s1(2 + pp + 1);
    fill_u64(s1.p, s1.n, 0);
    {
      ProfScope ps("join_probe");
      hipLaunchKernelGGL(join_probe1_k, dim3((unsigned)ntiles), dim3(kJBlock), 0, stream(), d, first.p, tcnt.p, (int*)(s1.p + 1));
      hipLaunchKernelGGL(join_scan_k, dim3(1), dim3(1024), 0, stream(), tcnt.p, ntiles, s1.p);
      hipLaunchKernelGGL(join_offsets1_k, dim3(1), dim3(64), 0, stream(), (const uint32_t*)tcnt.p, d.tile0, pp, ntiles,
                         (const u64*)s1.p, (int64_t*)(s1.p + 2));
      check_launch("join_probe");
    }
    const int64_t cap = expect > 0 ? std::min<int64_t>(expect, (int64_t)INT32_MAX) : 0;
    DevBuf<int32_t> eb, ep;
    if (cap > 0) {  // sized ahead: the write is queued behind the count, no round trip between
      eb = DevBuf<int32_t>(cap);
      ep = DevBuf<int32_t>(cap);
      ProfScope ps("join_write");
      hipLaunchKernelGGL(join_write1_k, dim3((unsigned)ntiles), dim3(kJBlock), 0, stream(), d, (const int32_t*)first.p,
                         (const uint32_t*)tcnt.p, eb.p, ep.p, matched, (const u64*)s1.p, (u64)cap);
      check_launch("join_write");
    }
    std::vector<u64> h(2 + pp + 1);
    d2h_sync(h.data(), s1.p, sizeof(u64) * h.size());
    if (!(int)h[1]) {
      const int64_t m = (int64_t)h[0];
      if (m > INT32_MAX) throw Error("join result too large (>2^31 rows)", WS_ERR_INTERNAL);
      for (int p = 0; p <= pp; ++p) res.out_off[p] = (int64_t)h[2 + p];
      res.m = m;
      if (cap > 0 && m <= cap) {
        res.b = std::move(eb);
        res.p = std::move(ep);
        return res;
      }
      res.b = DevBuf<int32_t>(std::max<int64_t>(m, 1));
      res.p = DevBuf<int32_t>(std::max<int64_t>(m, 1));
      if (m) {
        ProfScope ps("join_write");
        hipLaunchKernelGGL(join_write1_k, dim3((unsigned)ntiles), dim3(kJBlock), 0, stream(), d, (const int32_t*)first.p,
                           (const uint32_t*)tcnt.p, res.b.p, res.p.p, matched, (const u64*)s1.p, (u64)m);
        check_launch("join_write");
      }
      return res;
    }
  }
  // two-pass form (some probe row matches several build rows)
  DevBuf<u64> st(2 * ntiles + 4 + pp + 1);  // [states][tile_off][ctr][total][err][out_off]
  fill_u64(st.p, st.n, 0);
  u64* state = st.p;
  u64* tile_off = st.p + ntiles;
  unsigned* ctr = (unsigned*)(st.p + 2 * ntiles);
  u64* total = st.p + 2 * ntiles + 1;
  int* err = (int*)(st.p + 2 * ntiles + 2);
  int64_t* off_out = (int64_t*)(st.p + 2 * ntiles + 4);
  {
    ProfScope ps("join_count");
    hipLaunchKernelGGL(join_count_k, dim3((unsigned)ntiles), dim3(kJBlock), 0, stream(), d, state, ctr, tile_off, total, err);
    check_launch("join_count");
  }
  u64 h[3];
  d2h_sync(h, st.p + 2 * ntiles + 1, sizeof h);
  if (h[1]) throw Error("join look-back did not complete", WS_ERR_INTERNAL);
  const int64_t m = (int64_t)h[0];
  if (m > INT32_MAX) throw Error("join result too large (>2^31 rows)", WS_ERR_INTERNAL);
  hipLaunchKernelGGL(gather_off_k, dim3(1), dim3(64), 0, stream(), tile_off, d.tile0, pp, ntiles, (u64)m, off_out);
  check_launch("join_offsets");
  d2h_sync(res.out_off.data(), off_out, 8 * (size_t)(pp + 1));
  res.m = m;
  res.b = DevBuf<int32_t>(std::max<int64_t>(m, 1));
  res.p = DevBuf<int32_t>(std::max<int64_t>(m, 1));
  if (m) {
    ProfScope ps("join_write");
    hipLaunchKernelGGL(join_write_k, dim3((unsigned)ntiles), dim3(kJBlock), 0, stream(), d, (const u64*)tile_off, res.b.p, res.p.p,
                       matched);
    check_launch("join_write");
  }
  return res;
}

}  // namespace ws

using namespace ws;

extern "C" {

int ws_hash_join_parts(ws_colview build_key, const int64_t* b_off, int pb, ws_colview probe_key, const int64_t* p_off, int pp,
                       int32_t** out_build, int32_t** out_probe, int64_t* out_off, int64_t* m) {
  return guard([&] {
    *out_build = *out_probe = nullptr;
    *m = 0;
    std::vector<int64_t> bo(b_off, b_off + pb + 1), po(p_off, p_off + pp + 1);
    if (bo.back() == 0 || po.back() == 0) {
      if (out_off)
        for (int p = 0; p <= pp; ++p) out_off[p] = 0;
      return;
    }
    JoinResult r = join_core(build_key, bo, probe_key, po, false, nullptr);
    if (out_off)
      for (int p = 0; p <= pp; ++p) out_off[p] = r.out_off[p];
    *m = r.m;
    if (r.m) {
      *out_build = r.b.release();
      *out_probe = r.p.release();
    }
  });
}
int ws_hash_join_sized(ws_colview build_key, int64_t nb, ws_colview probe_key, int64_t np, int64_t expected_pairs,
                       int32_t** out_build, int32_t** out_probe, int64_t* m) {
  return guard([&] {
    *out_build = *out_probe = nullptr;
    *m = 0;
    if (nb <= 0 || np <= 0) return;  // as ws_hash_join: no pairs, no buffers
    JoinResult r = join_core(build_key, {0, nb}, probe_key, {0, np}, false, nullptr, expected_pairs);
    *m = r.m;
    if (r.m) {
      *out_build = r.b.release();
      *out_probe = r.p.release();
    }
  });
}
int ws_hash_join(ws_colview build_key, int64_t nb, ws_colview probe_key, int64_t np, int32_t** out_build, int32_t** out_probe,
                 int64_t* m) {
  const int64_t bo[2] = {0, nb}, po[2] = {0, np};
  return ws_hash_join_parts(build_key, bo, 1, probe_key, po, 1, out_build, out_probe, nullptr, m);
}
int ws_hash_join_outer(ws_colview build_key, int64_t nb, ws_colview probe_key, int64_t np, int32_t** out_build,
                       int32_t** out_probe, int full, int64_t* m) {
  return guard([&] {
    *out_build = *out_probe = nullptr;
    *m = 0;
    DevBuf<uint8_t> matched(full && nb ? nb : 0);
    if (matched.p) fill_u8(matched.p, nb, 0);
    JoinResult r;
    if (np) r = join_core(build_key, {0, nb}, probe_key, {0, np}, true, matched.p);
    int64_t extra = 0;
    DevBuf<int32_t> un;
    if (full && nb) {
      DevBuf<uint8_t> fl(nb);
      hipLaunchKernelGGL(unmatched_k, dim3(grid_for(nb)), dim3(256), 0, stream(), matched.p, nb, fl.p);
      check_launch("unmatched");
      un = DevBuf<int32_t>(nb);
      extra = compact_flags(fl.p, nb, un.p);
    }
    const int64_t total = r.m + extra;
    DevBuf<int32_t> ob(std::max<int64_t>(total, 1)), op(std::max<int64_t>(total, 1));
    if (r.m) {
      WS_CK(hipMemcpyAsync(ob.p, r.b.p, 4 * (size_t)r.m, hipMemcpyDeviceToDevice, stream()));
      WS_CK(hipMemcpyAsync(op.p, r.p.p, 4 * (size_t)r.m, hipMemcpyDeviceToDevice, stream()));
    }
    if (extra) {
      WS_CK(hipMemcpyAsync(ob.p + r.m, un.p, 4 * (size_t)extra, hipMemcpyDeviceToDevice, stream()));
      fill_i32(op.p + r.m, extra, -1);
    }
    *m = total;
    *out_build = ob.release();
    *out_probe = op.release();
  });
}
int ws_cross_join(int64_t na, int64_t nb, int32_t** out_a, int32_t** out_b, int64_t* m) {
  return guard([&] {
    const int64_t n = na * nb;
    if (n > 200000000) throw Error("cross join too large (" + std::to_string(n) + " rows); add a join condition", WS_ERR_INTERNAL);
    DevBuf<int32_t> a(std::max<int64_t>(n, 1)), b(std::max<int64_t>(n, 1));
    if (n) {
      ProfScope ps("cross_join");
      hipLaunchKernelGGL(cross_k, dim3(grid_for(n)), dim3(256), 0, stream(), na, nb, a.p, b.p);
      check_launch("cross_join");
    }
    *m = n;
    *out_a = a.release();
    *out_b = b.release();
  });
}

}  // extern "C"
