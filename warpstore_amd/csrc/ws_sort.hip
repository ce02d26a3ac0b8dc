// Stable multi-key sort permutation for ORDER BY (reference src/gpu.cu:1318-1386, which uses CUB's
// radix sort). Each ORDER BY key, last to first, becomes an order-preserving u64 (NULL -> -inf,
// -0 -> 0, descending keys complemented) and the (key, row) pairs go through a stable LSD radix
// sort with 8-bit digits. Digit positions where every key agrees are skipped (one reduction up
// front), so small-integer or narrow-range keys cost only the passes over their varying bytes.
// A pass: per-tile digit histograms (LDS atomics) -> one exclusive scan of the digit-major
// histogram table -> scatter, ranking equal digits inside a tile with wave ballots so the order of
// equal keys is kept.
#include <algorithm>

#include "ws_internal.h"

namespace ws {

constexpr int kSBlock = 256;
constexpr int kSItems = 16;                       // items per thread per tile
constexpr int kSTile = kSBlock * kSItems;         // 4096 keys per tile
constexpr int kScanTile = kSelRows * kSBlock;     // 1024 values per scan tile

__device__ __forceinline__ u64 order_key(double v, bool desc) {
  v = v != v ? -__builtin_huge_val() : (v == 0 ? 0.0 : v);
  const u64 b = (u64)dbits(v);
  const u64 u = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
  return desc ? ~u : u;
}

__global__ void seq_k(int32_t* p, int64_t m) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) p[j] = (int32_t)j;
}
__global__ void sort_key_k(const int32_t* perm, const double* key, int64_t m, int desc, u64* out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    out[j] = order_key(key[perm[j]], desc != 0);
}
// OR over (key ^ key[0]): the bits that vary between keys. One atomic per block (a grid of about
// one block per CU), since atomics on one address serialize.
__global__ void __launch_bounds__(256) varying_bits_k(const u64* k, int64_t m, unsigned long long* out) {
  __shared__ u64 s_or[4];
  const u64 k0 = k[0];
  u64 acc = 0;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) acc |= k[j] ^ k0;
  for (int o = 32; o > 0; o >>= 1) acc |= __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) s_or[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u64 v = s_or[0] | s_or[1] | s_or[2] | s_or[3];
    if (v) atomicOr(out, (unsigned long long)v);
  }
}

__global__ void __launch_bounds__(kSBlock) digit_hist_k(const u64* keys, int64_t m, int shift, int64_t ntiles, uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSTile;
  for (int j = 0; j < kSItems; ++j) {
    const int64_t i = base + (int64_t)j * kSBlock + threadIdx.x;
    if (i < m) __hip_atomic_fetch_add(&h[(keys[i] >> shift) & 255], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of n u32 values (single pass, decoupled look-back).
__global__ void __launch_bounds__(kSBlock) scan_u32_k(const uint32_t* in, uint32_t* out, int64_t n, u64* state, unsigned* ctr,
                                                      int* err) {
  __shared__ unsigned s_tile;
  __shared__ u64 s_w[kSBlock / 64];
  __shared__ u64 s_excl;
  if (threadIdx.x == 0) s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  const i64 t = s_tile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Each thread owns kSelRows consecutive values.
  const int64_t i0 = t * kScanTile + (int64_t)threadIdx.x * kSelRows;
  u64 v[kSelRows];
  u64 sum = 0;
  for (int j = 0; j < kSelRows; ++j) {
    v[j] = i0 + j < n ? in[i0 + j] : 0;
    sum += v[j];
  }
  u64 incl = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const u64 y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  u64 before = 0, tile_total = 0;
  for (int w = 0; w < kSBlock / 64; ++w) {
    if (w < wave) before += s_w[w];
    tile_total += s_w[w];
  }
  if (wave == 0) {
    const u64 excl = lookback_exclusive(state, t, tile_total, err);
    if (lane == 0) s_excl = excl;
  }
  __syncthreads();
  u64 run = s_excl + before + incl - sum;
  for (int j = 0; j < kSelRows; ++j) {
    if (i0 + j < n) out[i0 + j] = (uint32_t)run;
    run += v[j];
  }
}

// Stable scatter of one digit pass.
__global__ void __launch_bounds__(kSBlock) digit_scatter_k(const u64* kin, const int32_t* vin, int64_t m, int shift, int64_t ntiles,
                                                           const uint32_t* offs, u64* kout, int32_t* vout) {
  __shared__ uint32_t s_base[256];      // global position of this tile's first key of each digit
  __shared__ uint32_t s_run[256];       // keys of each digit already placed by earlier rounds
  __shared__ uint32_t s_wc[kSBlock / 64][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  s_base[threadIdx.x] = offs[(int64_t)threadIdx.x * ntiles + blockIdx.x];
  s_run[threadIdx.x] = 0;
  const int64_t base = (int64_t)blockIdx.x * kSTile;
  for (int j = 0; j < kSItems; ++j) {
    for (int w = 0; w < kSBlock / 64; ++w) s_wc[w][threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = base + (int64_t)j * kSBlock + threadIdx.x;
    const bool valid = i < m;
    const u64 k = valid ? kin[i] : 0;
    const int32_t v = valid ? vin[i] : 0;
    const unsigned dg = (unsigned)((k >> shift) & 255);
    // Lanes of this wave holding the same digit (and valid).
    u64 peers = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const u64 bal = __ballot((dg >> b) & 1);
      peers &= ((dg >> b) & 1) ? bal : ~bal;
    }
    const unsigned rank = (unsigned)__popcll(peers & ((1ull << lane) - 1));
    if (valid && rank == 0) s_wc[wave][dg] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t before = 0;
      for (int w = 0; w < wave; ++w) before += s_wc[w][dg];
      const uint32_t pos = s_base[dg] + s_run[dg] + before + rank;
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    uint32_t tot = 0;
    for (int w = 0; w < kSBlock / 64; ++w) tot += s_wc[w][threadIdx.x];
    s_run[threadIdx.x] += tot;
    __syncthreads();
  }
}

static void scan_u32(const uint32_t* in, uint32_t* out, int64_t n) {
  const int64_t nt = (n + kScanTile - 1) / kScanTile;
  DevBuf<u64> st(nt + 2);
  fill_u64(st.p, st.n, 0);
  hipLaunchKernelGGL(scan_u32_k, dim3((unsigned)nt), dim3(kSBlock), 0, stream(), in, out, n, st.p, (unsigned*)(st.p + nt),
                     (int*)(st.p + nt + 1));
  check_launch("scan_u32");
}

// Stable ascending sort of (keys, vals) by keys; results end in (keys, vals) (buffers swapped
// with the scratch pair as passes run).
static void radix_sort_pairs(DevBuf<u64>& keys, DevBuf<int32_t>& vals, DevBuf<u64>& k2, DevBuf<int32_t>& v2, int64_t m) {
  DevBuf<unsigned long long> vary(1);
  fill_u64((u64*)vary.p, 1, 0);
  hipLaunchKernelGGL(varying_bits_k, dim3(grid_for(m, 256, 1)), dim3(256), 0, stream(), keys.p, m, vary.p);
  check_launch("varying_bits");
  unsigned long long bits = 0;
  d2h_sync(&bits, vary.p, 8);
  const int64_t ntiles = (m + kSTile - 1) / kSTile;
  DevBuf<uint32_t> hist(256 * ntiles), offs(256 * ntiles);
  for (int pass = 0; pass < 8; ++pass) {
    const int shift = 8 * pass;
    if (((bits >> shift) & 255) == 0) continue;  // every key has the same digit here
    hipLaunchKernelGGL(digit_hist_k, dim3((unsigned)ntiles), dim3(kSBlock), 0, stream(), keys.p, m, shift, ntiles, hist.p);
    check_launch("digit_hist");
    scan_u32(hist.p, offs.p, 256 * ntiles);
    hipLaunchKernelGGL(digit_scatter_k, dim3((unsigned)ntiles), dim3(kSBlock), 0, stream(), keys.p, vals.p, m, shift, ntiles,
                       offs.p, k2.p, v2.p);
    check_launch("digit_scatter");
    std::swap(keys, k2);
    std::swap(vals, v2);
  }
}

// keys[i] sorted stably (ascending unless desc[i]); keys[0] most significant.
static DevBuf<int32_t> sort_perm(const std::vector<std::pair<const double*, bool>>& keys, int64_t m) {
  DevBuf<int32_t> perm(std::max<int64_t>(m, 1));
  if (m <= 0) return perm;
  ProfScope ps("sort_perm");
  hipLaunchKernelGGL(seq_k, dim3(grid_for(m)), dim3(256), 0, stream(), perm.p, m);
  check_launch("seq");
  DevBuf<int32_t> perm2(m);
  DevBuf<u64> k1(m), k2(m);
  for (int i = (int)keys.size() - 1; i >= 0; --i) {
    hipLaunchKernelGGL(sort_key_k, dim3(grid_for(m)), dim3(256), 0, stream(), perm.p, keys[i].first, m, keys[i].second ? 1 : 0,
                       k1.p);
    check_launch("sort_key");
    radix_sort_pairs(k1, perm, k2, perm2, m);
  }
  return perm;
}

// ------------------------------------------------------------------ ORDER BY ... LIMIT k
// The first k entries of the stable sort permutation without sorting all m rows: rows rank by
// (order keys, row index), a total order whose first k are exactly the stable sort's. A block
// keeps its best k so far in LDS, with the k-th as a threshold. A step offers 256 rows (one per
// thread); those that rank before the threshold are appended after the list, and every entry of
// the union computes its new rank in parallel (a list entry: its position plus the appended
// entries before it; an appended entry: a binary search in the list plus the appended entries
// before it) and moves to that slot of the other buffer if the rank is below k. After the first
// steps few rows pass the threshold, so a step costs little more than reading its rows. A second
// launch of one block selects the final k from the blocks' lists.
constexpr int kTopMaxK = 256, kTopMaxKeys = 4, kTopThreads = 256, kTopSlots = kTopMaxK + kTopThreads;

struct TopIn {
  const double* key[kTopMaxKeys];  // first pass: the key columns
  int desc[kTopMaxKeys];
  const u64* cok;     // second pass: candidates' order keys [n][nk]
  const int32_t* ci;  // and row indexes [n] (INT32_MAX: none)
};

template <int NK>
struct TopRow {
  u64 k[NK];
  int32_t i;  // INT32_MAX: no row
};
template <int NK>
__device__ __forceinline__ bool top_before(const TopRow<NK>& a, const TopRow<NK>& b) {
#pragma unroll
  for (int j = 0; j < NK; ++j)
    if (a.k[j] != b.k[j]) return a.k[j] < b.k[j];
  return a.i < b.i;
}

template <int NK, bool kFirst>
__device__ __forceinline__ TopRow<NK> top_load(const TopIn& in, int64_t r, int64_t r1) {
  TopRow<NK> e;
  e.i = INT32_MAX;
#pragma unroll
  for (int j = 0; j < NK; ++j) e.k[j] = ~0ull;  // no row: defined, sorts last
  if (r < r1) {
    if (kFirst) {
#pragma unroll
      for (int j = 0; j < NK; ++j) e.k[j] = order_key(in.key[j][r], in.desc[j] != 0);
      e.i = (int32_t)r;
    } else {
      e.i = in.ci[r];
#pragma unroll
      for (int j = 0; j < NK; ++j) e.k[j] = in.cok[r * NK + j];
    }
  }
  return e;
}

// LDS candidate buffer: key j of slot x at k[j][x] (a broadcast read per key), index at i[x].
template <int NK>
struct TopSlots {
  u64 k[NK][kTopSlots];
  int32_t i[kTopSlots];
  __device__ TopRow<NK> get(int x) const {
    TopRow<NK> e;
#pragma unroll
    for (int j = 0; j < NK; ++j) e.k[j] = k[j][x];
    e.i = i[x];
    return e;
  }
  __device__ void put(int x, const TopRow<NK>& e) {
#pragma unroll
    for (int j = 0; j < NK; ++j) k[j][x] = e.k[j];
    i[x] = e.i;
  }
};

// New rank of slot x of the union of the sorted list [0, cnt) and the new entries [cnt, n):
// an old entry moves up by the new entries before it; a new entry's rank is its lower bound in
// the old list plus the new entries before it.
template <int NK>
__device__ __forceinline__ int top_rank(const TopSlots<NK>& b, int x, int cnt, int n) {
  const TopRow<NK> me = b.get(x);
  int rank;
  if (x < cnt) {
    rank = x;
  } else {
    int lo = 0, hi = cnt;  // first old entry not before me
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (top_before(b.get(mid), me)) lo = mid + 1;
      else hi = mid;
    }
    rank = lo;
  }
#pragma unroll 8
  for (int y = cnt; y < n; ++y) rank += top_before(b.get(y), me);
  return rank;
}

// A step's new entries (slots [cnt, cnt + c) of b, any order, c <= kTopThreads) sorted in place by a
// bitonic sort over the block, one entry per thread (strides below a wave by shuffles, the two
// wider ones through `tmp`); slots past the new entries sort last (no row: the largest key and
// index). Every thread of the block calls it.
template <int NK>
__device__ __forceinline__ void top_sort_new(TopSlots<NK>& b, TopSlots<NK>& tmp, int cnt, int c) {
  const int t = threadIdx.x;
  TopRow<NK> me;
  if (t < c) {
    me = b.get(cnt + t);
  } else {
#pragma unroll
    for (int j = 0; j < NK; ++j) me.k[j] = ~0ull;
    me.i = INT32_MAX;
  }
  for (int size = 2; size <= kTopThreads; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      TopRow<NK> other;
      if (stride < kWave) {
#pragma unroll
        for (int j = 0; j < NK; ++j) other.k[j] = (u64)__shfl_xor((long long)me.k[j], stride, kWave);
        other.i = __shfl_xor(me.i, stride, kWave);
      } else {
        __syncthreads();  // the previous exchange's reads are done
        tmp.put(t, me);
        __syncthreads();
        other = tmp.get(t ^ stride);
      }
      // pair (t, t ^ stride): the lower thread keeps the smaller entry in an ascending run of
      // `size`, the larger in a descending one (the whole block is one ascending run at the end)
      const bool up = (t & size) == 0, lower = (t & stride) == 0;
      const bool take = lower == up ? top_before(other, me) : top_before(me, other);
      if (take) me = other;
    }
  }
  __syncthreads();  // every unsorted entry has been read (and the exchanges are done)
  if (t < c) b.put(cnt + t, me);
  __syncthreads();
}
// Rank of slot x in the union of two ascending runs of b, [0, cnt) (the list) and [cnt, n) (the
// step's sorted new entries): its index in its run plus the entries of the other run before it.
template <int NK>
__device__ __forceinline__ int top_rank_sorted(const TopSlots<NK>& b, int x, int cnt, int n) {
  const TopRow<NK> me = b.get(x);
  const bool old = x < cnt;
  int lo = old ? cnt : 0, hi = old ? n : cnt;
  const int base = old ? x : x - cnt;
  const int first = lo;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (top_before(b.get(mid), me)) lo = mid + 1;
    else hi = mid;
  }
  return base + (lo - first);
}
// Steps bringing more new entries than this sort them (top_sort_new) instead of ranking each
// against all the others.
constexpr int kTopSortFrom = 64;

// Whether entry e (pass: it is a candidate) is among the k best candidates of its wave: a
// candidate with k better ones in its own wave cannot be among the block's k best. The wave's
// candidates are compared through lane reads (no LDS, no barrier); a wave with at most k
// candidates keeps them all.
template <int NK>
__device__ __forceinline__ bool top_wave_keep(const TopRow<NK>& e, bool pass, int k) {
  const u64 pm = __ballot(pass);
  if (__popcll(pm) <= k) return pass;  // (wave-uniform)
  int better = 0;
#pragma unroll 16
  for (int j = 0; j < kWave; ++j) {
    TopRow<NK> o;
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)e.k[q], j);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(e.k[q] >> 32), j);
      o.k[q] = ((u64)hi << 32) | lo;
    }
    o.i = __builtin_amdgcn_readlane(e.i, j);
    better += ((pm >> j) & 1) && top_before(o, e) ? 1 : 0;
  }
  return pass && better < k;
}

template <int NK, bool kFirst>
__global__ void __launch_bounds__(kTopThreads) topk_k(TopIn in, int64_t m, int64_t per_block, int k, u64* out_ok,
                                                      int32_t* out_idx) {
  __shared__ TopSlots<NK> s_b[2];
  __shared__ int s_c;
  const int t = threadIdx.x;
  if (t == 0) s_c = 0;
  int cur = 0, cnt = 0;  // the list: entries [0, cnt) of buffer cur, ascending
  TopRow<NK> thr;        // its k-th entry, once cnt == k
  const int64_t r0 = (int64_t)blockIdx.x * per_block, r1 = r0 + per_block < m ? r0 + per_block : m;
  TopRow<NK> nxt = top_load<NK, kFirst>(in, r0 + t, r1);
  __syncthreads();
  for (int64_t base = r0; base < r1; base += kTopThreads) {
    const TopRow<NK> e = nxt;
    nxt = top_load<NK, kFirst>(in, base + kTopThreads + t, r1);  // next step's row, in flight meanwhile
    const bool pass = e.i != INT32_MAX && (cnt < k || top_before(e, thr));
    if (top_wave_keep(e, pass, k))
      s_b[cur].put(cnt + __hip_atomic_fetch_add(&s_c, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP), e);
    __syncthreads();
    const int c = s_c;
    __syncthreads();  // every thread has read the count before it is reset or appended to again
    if (c == 0) continue;  // (uniform)
    if (t == 0) s_c = 0;
    const int n = cnt + c, o = cur ^ 1;
    if (c > kTopSortFrom) {  // (uniform)
      top_sort_new(s_b[cur], s_b[o], cnt, c);
      for (int x = t; x < n; x += kTopThreads) {
        const int rank = top_rank_sorted(s_b[cur], x, cnt, n);
        if (rank < k) s_b[o].put(rank, s_b[cur].get(x));
      }
    } else {
      for (int x = t; x < n; x += kTopThreads) {
        const int rank = top_rank(s_b[cur], x, cnt, n);
        if (rank < k) s_b[o].put(rank, s_b[cur].get(x));
      }
    }
    __syncthreads();
    cur = o;
    cnt = n < k ? n : k;
    if (cnt == k) thr = s_b[cur].get(k - 1);
    // (the next step appends from slot cnt >= k on, past the threshold entry read here)
  }
  for (int x = t; x < k; x += kTopThreads) {
    const int64_t ob = (int64_t)blockIdx.x * k + x;
    const TopRow<NK> e = s_b[cur].get(x < cnt ? x : 0);
#pragma unroll
    for (int j = 0; j < NK; ++j) out_ok[ob * NK + j] = x < cnt ? e.k[j] : ~0ull;
    out_idx[ob] = x < cnt ? e.i : INT32_MAX;
  }
}

template <int NK>
static void topk_launch(TopIn in, int64_t m, int64_t kk, int32_t* out) {
  // first pass: about topk_steps steps of kTopThreads rows per block, at most topk_blocks_per_cu
  // blocks per CU (each step waits for its rows and three barriers, so fewer steps per block
  // shorten the pass while the candidates it leaves stay few); then passes of one step per block
  const int64_t steps = tune_knob("topk_steps"), bpc = tune_knob("topk_blocks_per_cu");
  const int64_t want = std::max<int64_t>(
      1, std::min<int64_t>((m + steps * kTopThreads - 1) / (steps * kTopThreads), cu_count() * bpc));
  const int64_t per = (m + want - 1) / want;
  const int64_t nb = (m + per - 1) / per;
  DevBuf<u64> cok(nb * kk * NK), fok(kk * NK);
  DevBuf<int32_t> ci(nb * kk);
  hipLaunchKernelGGL((topk_k<NK, true>), dim3((unsigned)nb), dim3(kTopThreads), 0, stream(), in, m, per, (int)kk, cok.p, ci.p);
  in.cok = cok.p;
  in.ci = ci.p;
  // While more than one step's worth of candidates remains and k is small enough for a step to
  // shrink them (k <= 64: 256 candidates -> k), blocks of one step each keep their best k; the
  // last pass is one block (a single block over all candidates would take a step per 256).
  int64_t n = nb * kk;
  DevBuf<u64> cok2[2];
  DevBuf<int32_t> ci2[2];
  for (int cur = 0; n > kTopThreads && kk * 4 <= kTopThreads; cur ^= 1) {
    const int64_t nb2 = (n + kTopThreads - 1) / kTopThreads;
    cok2[cur] = DevBuf<u64>(nb2 * kk * NK);
    ci2[cur] = DevBuf<int32_t>(nb2 * kk);
    hipLaunchKernelGGL((topk_k<NK, false>), dim3((unsigned)nb2), dim3(kTopThreads), 0, stream(), in, n, (int64_t)kTopThreads,
                       (int)kk, cok2[cur].p, ci2[cur].p);
    in.cok = cok2[cur].p;
    in.ci = ci2[cur].p;
    n = nb2 * kk;
  }
  hipLaunchKernelGGL((topk_k<NK, false>), dim3(1), dim3(kTopThreads), 0, stream(), in, n, n, (int)kk, fok.p, out);
  check_launch("topk");
}

// First min(k, m) entries of sort_perm(keys, m), by selection when k and the key count allow.
static DevBuf<int32_t> topk_perm(const std::vector<std::pair<const double*, bool>>& keys, int64_t m, int64_t k) {
  const int64_t kk = std::min(k, m);
  const int nk = (int)keys.size();
  if (kk <= 0 || nk == 0) return DevBuf<int32_t>(1);
  if (kk > kTopMaxK || nk > kTopMaxKeys || kk * 16 >= m) {  // a full sort is as cheap
    DevBuf<int32_t> perm = sort_perm(keys, m);
    if (kk == m) return perm;
    DevBuf<int32_t> out(kk);
    WS_CK(hipMemcpyAsync(out.p, perm.p, 4 * (size_t)kk, hipMemcpyDeviceToDevice, stream()));
    return out;
  }
  ProfScope ps("topk_perm");
  TopIn in{};
  for (int j = 0; j < nk; ++j) in.key[j] = keys[j].first, in.desc[j] = keys[j].second ? 1 : 0;
  DevBuf<int32_t> out(kk);
  switch (nk) {
    case 1: topk_launch<1>(in, m, kk, out.p); break;
    case 2: topk_launch<2>(in, m, kk, out.p); break;
    case 3: topk_launch<3>(in, m, kk, out.p); break;
    default: topk_launch<4>(in, m, kk, out.p); break;
  }
  return out;
}

template <class T>
__global__ void to_double_k(const T* src, int64_t n, double* dst) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) dst[j] = (double)src[j];
}

}  // namespace ws

using namespace ws;

extern "C" {

int ws_sort_perm(const double* const* keys, const uint8_t* desc, int nkeys, int64_t m, int32_t** perm) {
  return guard([&] {
    *perm = nullptr;
    std::vector<std::pair<const double*, bool>> k;
    for (int i = 0; i < nkeys; ++i) k.push_back({keys[i], desc && desc[i]});
    if (m > INT32_MAX) throw Error("sort: too many rows", WS_ERR_INTERNAL);
    *perm = sort_perm(k, m).release();
  });
}
int ws_topk_perm(const double* const* keys, const uint8_t* desc, int nkeys, int64_t m, int64_t k, int32_t** perm) {
  return guard([&] {
    *perm = nullptr;
    std::vector<std::pair<const double*, bool>> kv;
    for (int i = 0; i < nkeys; ++i) kv.push_back({keys[i], desc && desc[i]});
    if (m > INT32_MAX - 1) throw Error("sort: too many rows", WS_ERR_INTERNAL);
    if (k < 0) throw Error("top-k: negative k", WS_ERR_INTERNAL);
    *perm = topk_perm(kv, m, k).release();
  });
}
int ws_sort_perm_host(const void* host_col, int ptype, int64_t n, int32_t* out) {
  return guard([&] {
    if (n <= 0) return;
    const size_t es = ptype == WS_P_I32 ? 4 : 8;
    DevBuf<uint8_t> raw(n * (int64_t)es);
    h2d_sync(raw.p, host_col, n * es);
    DevBuf<double> k(n);
    if (ptype == WS_P_I32)
      hipLaunchKernelGGL(to_double_k<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream(), (const int32_t*)raw.p, n, k.p);
    else if (ptype == WS_P_I64)
      hipLaunchKernelGGL(to_double_k<int64_t>, dim3(grid_for(n)), dim3(256), 0, stream(), (const int64_t*)raw.p, n, k.p);
    else
      WS_CK(hipMemcpyAsync(k.p, raw.p, 8 * (size_t)n, hipMemcpyDeviceToDevice, stream()));
    check_launch("to_double");
    DevBuf<int32_t> perm = sort_perm({{k.p, false}}, n);
    d2h_sync(out, perm.p, 4 * (size_t)n);
  });
}

}  // extern "C"
