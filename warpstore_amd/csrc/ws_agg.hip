// GpuAggregator and row selection over batches (reference src/gpu.cu:393-755), planned for gfx950.
#include <algorithm>
#include <climits>
#include <cmath>
#include <memory>

#include "ws_internal.h"

// Ahead-of-time kernels: the shared bodies with the bytecode interpreter.
WS_KERNEL_NOKEYS(ws::InterpEval, ws_i_nokeys)
WS_KERNEL_DENSE(ws::InterpEval, ws_i_dense)
WS_KERNEL_GROUPED(ws::InterpEval, ws_i_grouped)
WS_KERNEL_SELECT(ws::InterpEval, ws_i_select)
WS_KERNEL_NOKEYS(ws::InterpEvalLean, ws_l_nokeys)
WS_KERNEL_DENSE(ws::InterpEvalLean, ws_l_dense)
WS_KERNEL_GROUPED(ws::InterpEvalLean, ws_l_grouped)
WS_KERNEL_SELECT(ws::InterpEvalLean, ws_l_select)

namespace ws {

constexpr int kBlock = 256;
constexpr int kAggTileRows = 4096;
constexpr int kSelTileRows = kSelRows * kBlock;
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr size_t kStaticLds = 2 * 1024 + 512;  // staged column views + small shared state

// ------------------------------------------------------------------ small kernels

__global__ void init_table_k(GTable t, int nk, int na, ws_agg_plan plan) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= (int64_t)t.mask; i += (int64_t)gridDim.x * blockDim.x) {
    t.keys[i] = kEmpty;
    for (int a = 0; a < na; ++a) t.acc[i * na + a] = slot_identity(plan.agg_bits[a] != 0, plan.ops[a]);
  }
}
__global__ void init_dense_k(double* acc, int64_t G, int na, ws_agg_plan plan) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < G * na; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(i % na);
    acc[i] = slot_identity(plan.agg_bits[a] != 0, plan.ops[a]);
  }
}
__global__ void dense_reduce_k(ws_agg_plan plan, ws_dense_spec ds, const i64* parts, int nparts, double* gacc) {
  dense_reduce_body(plan, ds, parts, nparts, gacc);
}
// Compacted groups, row-major (fetch) or column-major (compact), keys from the table.
__global__ void gather_groups_k(const double* kv, int nk, const double* acc, int na, const int32_t* idx, int64_t g,
                                double* keys_out, double* accs_out, int colmajor) {
  const int64_t total = g * (nk + na);
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = j / (nk + na);
    const int c = (int)(j - i * (nk + na));
    const int64_t s = idx[i];
    if (c < nk) {
      const double v = kv[s * nk + c];
      if (colmajor) keys_out[(int64_t)c * g + i] = v;
      else keys_out[i * nk + c] = v;
    } else {
      const int a = c - nk;
      const double v = acc[s * na + a];
      if (colmajor) accs_out[(int64_t)a * g + i] = v;
      else accs_out[i * na + a] = v;
    }
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
  int64_t cap = 1;
  u64* keys = nullptr;
  double* kv = nullptr;
  double* acc = nullptr;
  int* flags = nullptr;
  u64* count = nullptr;
  double* nk_out = nullptr;  // no GROUP BY: [2 * WS_MAX_AGGS] running result
  std::vector<int64_t> hi;
  int ncols = 0;
  std::vector<int64_t> bn;
  std::vector<const uint32_t*> bdel;
  std::vector<ws_colview> bcols;
  int64_t rows_done = 0;
  int batches_done = 0, windows = 0;
  // fetch state
  int64_t fetched = -1;
  DevBuf<double> fetch_buf;
  ~ws_agg() {
    for (void* p : {(void*)keys, (void*)kv, (void*)acc, (void*)flags, (void*)count, (void*)nk_out}) tmp_free(p);
  }
};

// The lean interpreter serves plans whose programs use no scalar-function opcodes.
static bool plan_is_simple(const ws_expr_env& env, const ws_agg_plan& plan) {
  if (plan.has_where && !prog_is_simple(env, plan.where)) return false;
  for (int k = 0; k < plan.nkeys; ++k)
    if (!prog_is_simple(env, plan.keys[k])) return false;
  for (int a = 0; a < plan.naggs; ++a)
    if (plan.ops[a] != WS_AGG_COUNT && !prog_is_simple(env, plan.aggs[a])) return false;
  return true;
}

static double host_identity(uint8_t op) { return op == WS_AGG_MIN ? INFINITY : op == WS_AGG_MAX ? -INFINITY : 0.0; }

static void agg_run(ws_agg* A, const ws_expr_env& env, const ws_jit_program* jit) {
  const ws_agg_plan& plan = A->plan;
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
  const bool lean = !jf && plan_is_simple(env, plan);
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
      hipLaunchKernelGGL(lean ? ws_l_nokeys : ws_i_nokeys, dim3(grid), dim3(kBlock), 0, stream(), up.ddesc, out, parts.p, done);
      check_launch("agg_nokeys");
    }
    return;
  }
  if (A->is_dense) {
    const ws_dense_spec& ds = A->dense;
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
      hipLaunchKernelGGL(lean ? ws_l_dense : ws_i_dense, dim3((unsigned)grid), dim3(kBlock), dp.lds, stream(), up.ddesc, gacc, pp,
                         flags);
      check_launch("agg_dense");
    }
    if (dp.mode != 0) {
      ProfScope ps("agg_dense_reduce");
      hipLaunchKernelGGL(dense_reduce_k, dim3(grid_for(S)), dim3(kBlock), 0, stream(), plan, ds, (const i64*)pp, (int)grid,
                         gacc);
      check_launch("agg_dense_reduce");
    }
    return;
  }
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
  GTable t{A->keys, A->kv, A->acc, (u64)(A->cap - 1), A->flags, A->count};
  if (jf) {
    void* args[] = {&up.ddesc, &t};
    jit_launch_raw(jf, grid, kBlock, lds, args, "agg_hash_jit");
  } else {
    ProfScope ps("agg_hash");
    hipLaunchKernelGGL(lean ? ws_l_grouped : ws_i_grouped, dim3(grid), dim3(kBlock), lds, stream(), up.ddesc, t);
    check_launch("agg_hash");
  }
}

// Compacts the groups into a tmp buffer: keys then accs, row-major (fetch) or column-major.
static int64_t agg_compact_into(ws_agg* A, DevBuf<double>& buf, bool colmajor) {
  const int nk = A->plan.nkeys, na = A->plan.naggs;
  DevBuf<int32_t> idx(A->cap);
  const int64_t g = A->is_dense ? compact_dense(A->acc, na, A->dense.G, idx.p) : compact_occupied_hash(A->keys, A->cap, idx.p);
  buf = DevBuf<double>(std::max<int64_t>(g * (nk + na), 1));
  if (g) {
    double* ko = buf.p;
    double* ao = buf.p + g * nk;
    ProfScope ps("gather_groups");
    if (A->is_dense)
      hipLaunchKernelGGL(gather_dense_k, dim3(grid_for(g * (nk + na))), dim3(256), 0, stream(), A->dense, A->plan, A->acc, nk,
                         na, idx.p, g, ko, ao, colmajor ? 1 : 0);
    else
      hipLaunchKernelGGL(gather_groups_k, dim3(grid_for(g * (nk + na))), dim3(256), 0, stream(), A->kv, nk, A->acc, na, idx.p,
                         g, ko, ao, colmajor ? 1 : 0);
    check_launch("gather_groups");
  }
  return g;
}

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
      const int nk = plan->nkeys, na = std::max(plan->naggs, 1);
      A->flags = (int*)tmp_alloc(sizeof(int) * 4);
      A->count = (u64*)tmp_alloc(sizeof(u64));
      fill_u32((uint32_t*)A->flags, 4, 0);
      fill_u64(A->count, 1, 0);
      A->acc = (double*)tmp_alloc(sizeof(double) * (size_t)A->cap * na);
      if (A->is_dense) {
        hipLaunchKernelGGL(init_dense_k, dim3(grid_for(A->cap * na)), dim3(256), 0, stream(), A->acc, A->cap, plan->naggs, A->plan);
      } else {
        A->keys = (u64*)tmp_alloc(sizeof(u64) * (size_t)A->cap);
        A->kv = (double*)tmp_alloc(sizeof(double) * (size_t)A->cap * nk);
        GTable t{A->keys, A->kv, A->acc, (u64)(A->cap - 1), A->flags, A->count};
        hipLaunchKernelGGL(init_table_k, dim3(grid_for(A->cap)), dim3(256), 0, stream(), t, nk, plan->naggs, A->plan);
      }
      check_launch("init_table");
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
    agg_run(a, *env, jit);
    a->fetched = -1;
  });
}
int ws_agg_run_window(ws_agg* a, const ws_expr_env* env, const ws_jit_program* jit) {
  return guard([&] {
    if (a->bn.empty()) return;
    agg_run(a, *env, jit);
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
    int f[4];
    d2h_sync(f, a->flags, sizeof f);
    if (f[1]) throw Error("dense GROUP BY key outside the planned key range", WS_ERR_INTERNAL);
    if (!a->is_dense) *out = f[0] != 0;
  });
}
int ws_agg_fetch_begin(ws_agg* a, int64_t* groups) {
  return guard([&] {
    if (a->plan.nkeys == 0) {
      *groups = 1;
      a->fetched = 1;
      return;
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
    DevBuf<double> buf;
    const int64_t g = agg_compact_into(a, buf, true);
    const int nk = a->plan.nkeys, na = a->plan.naggs;
    // Two separately owned buffers (the caller frees each): split the column-major result.
    DevBuf<double> k(std::max<int64_t>(g * nk, 1)), ac(std::max<int64_t>(g * na, 1));
    if (g) {
      WS_CK(hipMemcpyAsync(k.p, buf.p, sizeof(double) * g * nk, hipMemcpyDeviceToDevice, stream()));
      WS_CK(hipMemcpyAsync(ac.p, buf.p + g * nk, sizeof(double) * g * na, hipMemcpyDeviceToDevice, stream()));
    }
    *keys = k.release();
    *accs = ac.release();
    *groups = g;
  });
}
int64_t ws_agg_capacity(const ws_agg* a) { return a->cap; }
int ws_agg_group_count(ws_agg* a, int64_t* out) {
  return guard([&] {
    u64 c = 0;
    if (a->count) d2h_sync(&c, a->count, sizeof c);
    *out = (int64_t)c;
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
  DescUpload up;
  build_desc(up, d, in, fp, nullptr, env, cols_from_env, 0, nullptr);
  // [tile states][counter][total][err][per-batch counts]
  DevBuf<u64> st(ntiles + 3 + in.nb);
  fill_u64(st.p, st.n, 0);
  unsigned* ctr = (unsigned*)(st.p + ntiles);
  u64* total = st.p + ntiles + 1;
  int* err = (int*)(st.p + ntiles + 2);
  u64* pb = st.p + ntiles + 3;
  void* jf = jit ? jit_function(jit, fn) : nullptr;
  if (jf) {
    void* args[] = {&up.ddesc, &st.p, &ctr, &out_idx, &pb, &total, &err};
    jit_launch_raw(jf, (unsigned)ntiles, kBlock, 0, args, "select_jit");
  } else {
    ProfScope ps("select");
    const bool lean = !has_where || prog_is_simple(env, where);
    hipLaunchKernelGGL(lean ? ws_l_select : ws_i_select, dim3((unsigned)ntiles), dim3(kBlock), 0, stream(), up.ddesc, st.p, ctr,
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

}  // extern "C"
