// Learned sizing: a small MLP that predicts the group count of a GROUP BY from segment statistics,
// trained on the GPU by hand-written forward / backward / Adam kernels on exact counts that the
// aggregation itself logs. The prediction only chooses physical parameters (table capacity,
// partition count, shared-table versus partitioned plan); results never depend on it, and an
// under-estimate is caught by the plain path's overflow check and retried exactly.
//
// Network: 16 inputs -> 32 ReLU -> 32 ReLU -> 1, fp32; target log2(groups + 1) / 32.
// Loss: mean squared error over the batch. All FLOPs are negligible next to a scan.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "ws_internal.h"

namespace ws {

constexpr int kIn = WS_SIZER_INPUTS, kH = 32;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kIn, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH, kParams = kOffB3 + 1;
static_assert(kParams == WS_SIZER_PARAMS, "parameter count");

// Forward pass, one thread per sample; the parameters are staged in LDS. Stores the hidden
// activations when h1 / h2 are given (training).
__global__ void __launch_bounds__(256) mlp_forward_k(const float* __restrict__ prm, const float* __restrict__ x, int64_t B,
                                                     float* __restrict__ y, float* __restrict__ h1o, float* __restrict__ h2o) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) w[i] = prm[i];
  __syncthreads();
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  float xi[kIn], h1[kH], h2[kH];
#pragma unroll
  for (int j = 0; j < kIn; ++j) xi[j] = x[b * kIn + j];
#pragma unroll
  for (int i = 0; i < kH; ++i) {
    float a = w[kOffB1 + i];
#pragma unroll
    for (int j = 0; j < kIn; ++j) a = fmaf(w[kOffW1 + i * kIn + j], xi[j], a);
    h1[i] = a > 0.f ? a : 0.f;
  }
#pragma unroll
  for (int i = 0; i < kH; ++i) {
    float a = w[kOffB2 + i];
#pragma unroll
    for (int j = 0; j < kH; ++j) a = fmaf(w[kOffW2 + i * kH + j], h1[j], a);
    h2[i] = a > 0.f ? a : 0.f;
  }
  float o = w[kOffB3];
#pragma unroll
  for (int j = 0; j < kH; ++j) o = fmaf(w[kOffW3 + j], h2[j], o);
  y[b] = o;
  if (h1o) {
#pragma unroll
    for (int i = 0; i < kH; ++i) h1o[b * kH + i] = h1[i], h2o[b * kH + i] = h2[i];
  }
}

// Backward pass, part 1, one thread per sample: output and hidden-layer deltas of the MSE loss
// L = (1/B) sum (y - t)^2.
__global__ void __launch_bounds__(256) mlp_delta_k(const float* __restrict__ prm, const float* __restrict__ t, int64_t B,
                                                   const float* __restrict__ y, const float* __restrict__ h1,
                                                   const float* __restrict__ h2, float* __restrict__ d1, float* __restrict__ d2,
                                                   float* __restrict__ d3) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) w[i] = prm[i];
  __syncthreads();
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float g3 = 2.f * (y[b] - t[b]) / (float)B;
  d3[b] = g3;
  float g2[kH];
#pragma unroll
  for (int i = 0; i < kH; ++i) {
    g2[i] = h2[b * kH + i] > 0.f ? w[kOffW3 + i] * g3 : 0.f;
    d2[b * kH + i] = g2[i];
  }
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    float a = 0.f;
#pragma unroll
    for (int i = 0; i < kH; ++i) a = fmaf(w[kOffW2 + i * kH + j], g2[i], a);
    d1[b * kH + j] = h1[b * kH + j] > 0.f ? a : 0.f;
  }
}

// Backward pass, part 2, one thread per parameter: the gradient as a sum over the batch in sample
// order (deterministic, no atomics).
__global__ void __launch_bounds__(256) mlp_grad_k(const float* __restrict__ x, int64_t B, const float* __restrict__ h1,
                                                  const float* __restrict__ h2, const float* __restrict__ d1,
                                                  const float* __restrict__ d2, const float* __restrict__ d3,
                                                  float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  float g = 0.f;
  if (p < kOffB1) {
    const int i = p / kIn, j = p % kIn;
    for (int64_t b = 0; b < B; ++b) g = fmaf(d1[b * kH + i], x[b * kIn + j], g);
  } else if (p < kOffW2) {
    const int i = p - kOffB1;
    for (int64_t b = 0; b < B; ++b) g += d1[b * kH + i];
  } else if (p < kOffB2) {
    const int i = (p - kOffW2) / kH, j = (p - kOffW2) % kH;
    for (int64_t b = 0; b < B; ++b) g = fmaf(d2[b * kH + i], h1[b * kH + j], g);
  } else if (p < kOffW3) {
    const int i = p - kOffB2;
    for (int64_t b = 0; b < B; ++b) g += d2[b * kH + i];
  } else if (p < kOffB3) {
    const int j = p - kOffW3;
    for (int64_t b = 0; b < B; ++b) g = fmaf(d3[b], h2[b * kH + j], g);
  } else {
    for (int64_t b = 0; b < B; ++b) g += d3[b];
  }
  grad[p] = g;
}

// Adam step (Kingma & Ba 2015) with bias correction.
__global__ void adam_k(float* __restrict__ prm, const float* __restrict__ grad, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  prm[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void mse_k(const float* __restrict__ y, const float* __restrict__ t, int64_t B, float* __restrict__ out) {
  __shared__ float s[256];
  float a = 0.f;
  for (int64_t b = threadIdx.x; b < B; b += blockDim.x) {
    const float e = y[b] - t[b];
    a += e * e;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0] / (float)B;
}

// Counter-based uniform in [0, 1) (splitmix64 of seed and index).
static double unif(u64 seed, u64 i) {
  u64 x = seed * 0x9e3779b97f4a7c15ull + i + 0x632be59bd9b4e019ull;
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// Exact counts logged by aggregations that carried features (ws_agg_set_features).
std::mutex g_log_mu;
std::vector<float> g_log_x, g_log_t;

void sizing_log(const float* features, int64_t groups) {
  std::lock_guard<std::mutex> g(g_log_mu);
  g_log_x.insert(g_log_x.end(), features, features + kIn);
  g_log_t.push_back((float)(std::log2((double)groups + 1.0) / 32.0));
}
// Exact join output sizes logged by the callers (ws_join_log_add), the join model's training set.
std::mutex g_jlog_mu;
std::vector<float> g_jlog_x, g_jlog_t;

}  // namespace ws

using namespace ws;

struct ws_sizer {
  DevBuf<float> prm, grad, m, v;
  int64_t steps = 0;
};

namespace {
// Activations and deltas of one batch.
struct Work {
  DevBuf<float> x, t, y, h1, h2, d1, d2, d3, loss;
  explicit Work(int64_t B)
      : x(B * kIn), t(B), y(B), h1(B * kH), h2(B * kH), d1(B * kH), d2(B * kH), d3(B), loss(1) {}
};
void forward_backward(ws_sizer* s, Work& w, int64_t B, bool grads) {
  const unsigned g = (unsigned)((B + 255) / 256);
  hipLaunchKernelGGL(mlp_forward_k, dim3(g), dim3(256), 0, stream(), (const float*)s->prm.p, (const float*)w.x.p, B, w.y.p,
                     w.h1.p, w.h2.p);
  hipLaunchKernelGGL(mse_k, dim3(1), dim3(256), 0, stream(), (const float*)w.y.p, (const float*)w.t.p, B, w.loss.p);
  if (grads) {
    hipLaunchKernelGGL(mlp_delta_k, dim3(g), dim3(256), 0, stream(), (const float*)s->prm.p, (const float*)w.t.p, B,
                       (const float*)w.y.p, (const float*)w.h1.p, (const float*)w.h2.p, w.d1.p, w.d2.p, w.d3.p);
    hipLaunchKernelGGL(mlp_grad_k, dim3((kParams + 255) / 256), dim3(256), 0, stream(), (const float*)w.x.p, B,
                       (const float*)w.h1.p, (const float*)w.h2.p, (const float*)w.d1.p, (const float*)w.d2.p,
                       (const float*)w.d3.p, s->grad.p);
  }
  check_launch("sizer");
}
}  // namespace

extern "C" {

int ws_sizer_create(uint64_t seed, ws_sizer** out) {
  return guard([&] {
    *out = nullptr;
    ensure_init();
    auto s = std::make_unique<ws_sizer>();
    s->prm = DevBuf<float>(kParams);
    s->grad = DevBuf<float>(kParams);
    s->m = DevBuf<float>(kParams);
    s->v = DevBuf<float>(kParams);
    // He-uniform weights, zero biases (output bias at the middle of the target range)
    std::vector<float> h(kParams, 0.f);
    for (int i = 0; i < kH * kIn; ++i) h[kOffW1 + i] = (float)((2 * unif(seed, i) - 1) * std::sqrt(6.0 / kIn));
    for (int i = 0; i < kH * kH; ++i) h[kOffW2 + i] = (float)((2 * unif(seed, 10000 + i) - 1) * std::sqrt(6.0 / kH));
    for (int i = 0; i < kH; ++i) h[kOffW3 + i] = (float)((2 * unif(seed, 20000 + i) - 1) * std::sqrt(6.0 / kH));
    h[kOffB3] = 0.4f;
    h2d_sync(s->prm.p, h.data(), sizeof(float) * kParams);
    std::vector<float> z(kParams, 0.f);
    h2d_sync(s->m.p, z.data(), sizeof(float) * kParams);
    h2d_sync(s->v.p, z.data(), sizeof(float) * kParams);
    *out = s.release();
  });
}
void ws_sizer_destroy(ws_sizer* s) { delete s; }

int ws_sizer_get_params(const ws_sizer* s, float* out) {
  return guard([&] { d2h_sync(out, s->prm.p, sizeof(float) * kParams); });
}
int ws_sizer_set_params(ws_sizer* s, const float* in) {
  return guard([&] { h2d_sync(s->prm.p, in, sizeof(float) * kParams); });
}

int ws_sizer_predict(ws_sizer* s, const float* x, int64_t B, float* y) {
  return guard([&] {
    if (B <= 0) return;
    DevBuf<float> dx(B * kIn), dy(B);
    h2d_sync(dx.p, x, sizeof(float) * B * kIn);
    hipLaunchKernelGGL(mlp_forward_k, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, stream(), (const float*)s->prm.p,
                       (const float*)dx.p, B, dy.p, (float*)nullptr, (float*)nullptr);
    check_launch("sizer_predict");
    d2h_sync(y, dy.p, sizeof(float) * B);
  });
}

int ws_sizer_loss_grad(ws_sizer* s, const float* x, const float* t, int64_t B, float* loss, float* grad) {
  return guard([&] {
    if (B <= 0) throw Error("sizer: empty batch", WS_ERR_INTERNAL);
    Work w(B);
    h2d_sync(w.x.p, x, sizeof(float) * B * kIn);
    h2d_sync(w.t.p, t, sizeof(float) * B);
    forward_backward(s, w, B, grad != nullptr);
    d2h_sync(loss, w.loss.p, sizeof(float));
    if (grad) d2h_sync(grad, s->grad.p, sizeof(float) * kParams);
  });
}

// Minibatch Adam over n samples for `epochs` passes (sample order shuffled per epoch from seed).
int ws_sizer_train(ws_sizer* s, const float* x, const float* t, int64_t n, int epochs, int batch, float lr, uint64_t seed,
                   float* final_loss) {
  return guard([&] {
    if (n <= 0 || batch <= 0 || epochs <= 0) throw Error("sizer: nothing to train on", WS_ERR_INTERNAL);
    const int64_t B = std::min<int64_t>(batch, n);
    Work w(B);
    std::vector<int64_t> order(n);
    std::vector<float> bx(B * kIn), bt(B);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
      for (int64_t i = 0; i < n; ++i) order[i] = i;
      for (int64_t i = n - 1; i > 0; --i) std::swap(order[i], order[(int64_t)(unif(seed + e, i) * (i + 1))]);
      for (int64_t o = 0; o + B <= n; o += B) {
        for (int64_t k = 0; k < B; ++k) {
          std::memcpy(&bx[k * kIn], &x[order[o + k] * kIn], sizeof(float) * kIn);
          bt[k] = t[order[o + k]];
        }
        h2d_async_staged(w.x.p, bx.data(), sizeof(float) * B * kIn);
        h2d_async_staged(w.t.p, bt.data(), sizeof(float) * B);
        forward_backward(s, w, B, true);
        ++s->steps;
        const float c1 = 1.f - std::pow(b1, (float)s->steps), c2 = 1.f - std::pow(b2, (float)s->steps);
        hipLaunchKernelGGL(adam_k, dim3((kParams + 255) / 256), dim3(256), 0, stream(), s->prm.p, (const float*)s->grad.p,
                           s->m.p, s->v.p, lr, b1, b2, eps, c1, c2);
        check_launch("sizer_adam");
      }
    }
    if (final_loss) {
      // loss over all samples with the trained parameters
      Work all(n);
      h2d_sync(all.x.p, x, sizeof(float) * n * kIn);
      h2d_sync(all.t.p, t, sizeof(float) * n);
      forward_backward(s, all, n, false);
      d2h_sync(final_loss, all.loss.p, sizeof(float));
    }
  });
}

// Features of a GROUP BY for the sizing model, from merged statistics of the key columns.
//   0 log2(rows + 1) / 32            1 selectivity estimate (0..1)   2 log2(rows * sel + 1) / 32
//   3 + 2k: log2(distinct_k + 1) / 32 (HyperLogLog)   4 + 2k: log2(span_k + 1) / 32 (integer span or
//   dictionary size; 0 for doubles), keys k = 0..3     11 nkeys / 4
//   12 log2(prod distinct + 1) / 64   13 log2(min(prod distinct, rows * sel) + 1) / 32
//   14 log2(max distinct + 1) / 32    15 NULL fraction of the keys (mean)
int ws_sizing_features(const ws_col_stats* keys, int nkeys, const int64_t* dict_sizes, int64_t rows, double selectivity,
                       float* out) {
  return guard([&] {
    if (nkeys < 1 || nkeys > 4) throw Error("sizing features: 1 to 4 keys", WS_ERR_INTERNAL);
    std::memset(out, 0, sizeof(float) * kIn);
    const double sel = std::min(1.0, std::max(0.0, selectivity));
    const double live = (double)std::max<int64_t>(rows, 0);
    out[0] = (float)(std::log2(live + 1) / 32);
    out[1] = (float)sel;
    out[2] = (float)(std::log2(live * sel + 1) / 32);
    double lprod = 0, dmax = 0, nullf = 0;
    for (int k = 0; k < nkeys; ++k) {
      const ws_col_stats& c = keys[k];
      const double d = ws_hll_estimate(c.hll) + (c.nulls ? 1 : 0);  // NULL is one more group
      double span = 0;
      if (dict_sizes && dict_sizes[k] > 0) span = (double)dict_sizes[k];
      else if (c.imin <= c.imax) span = (double)c.imax - (double)c.imin + 1;
      out[3 + 2 * k] = (float)(std::log2(d + 1) / 32);
      out[4 + 2 * k] = (float)(std::log2(span + 1) / 32);
      lprod += std::log2(d + 1);
      dmax = std::max(dmax, d);
      nullf += c.rows ? (double)c.nulls / (double)c.rows : 0;
    }
    out[11] = (float)nkeys / 4;
    out[12] = (float)(lprod / 64);
    out[13] = (float)(std::min(lprod, std::log2(live * sel + 1)) / 32);
    out[14] = (float)(std::log2(dmax + 1) / 32);
    out[15] = (float)(nullf / nkeys);
  });
}

// Predicted group count of one feature vector (2^(32 y) - 1, at least 1).
int ws_sizer_groups(ws_sizer* s, const float* features, int64_t* groups) {
  return guard([&] {
    float y = 0;
    int rc = ws_sizer_predict(s, features, 1, &y);
    if (rc) throw Error(ws_last_error(), rc);
    const double g = std::exp2(std::min(62.0, std::max(0.0, 32.0 * (double)y))) - 1.0;
    *groups = std::max<int64_t>(1, (int64_t)std::llround(g));
  });
}

// Features of an equi-join's output size from the key columns' statistics (table-level: the
// build side may have been filtered to build_rows of build_table_rows rows, so its distinct keys
// are capped there). Slots: log2 probe rows, log2 build rows, build selectivity, log2 distinct
// build / probe keys, log2 key spans, key-range overlap over the probe span, log2 build rows per
// distinct key, log2 of the containment estimate |B| |P| / max(d_B, d_P), NULL fractions.
int ws_join_features(const ws_col_stats* build_key, int64_t build_rows, int64_t build_table_rows, const ws_col_stats* probe_key,
                     int64_t probe_rows, float* out) {
  return guard([&] {
    std::memset(out, 0, sizeof(float) * kIn);
    const double nb = (double)std::max<int64_t>(build_rows, 0), np = (double)std::max<int64_t>(probe_rows, 0);
    const double sel = build_table_rows > 0 ? std::min(1.0, nb / (double)build_table_rows) : 1.0;
    auto span = [](const ws_col_stats& c) { return c.imin <= c.imax ? (double)c.imax - (double)c.imin + 1 : 0.0; };
    const double db = std::min(nb, ws_hll_estimate(build_key->hll) * sel + 1), dp = ws_hll_estimate(probe_key->hll);
    const double sb = span(*build_key), sp = span(*probe_key);
    double overlap = 0;
    if (sb > 0 && sp > 0) {
      const double lo = std::max((double)build_key->imin, (double)probe_key->imin);
      const double hi = std::min((double)build_key->imax, (double)probe_key->imax);
      overlap = hi >= lo ? (hi - lo + 1) / sp : 0.0;
    }
    out[0] = (float)(std::log2(np + 1) / 32);
    out[1] = (float)(std::log2(nb + 1) / 32);
    out[2] = (float)sel;
    out[3] = (float)(std::log2(db + 1) / 32);
    out[4] = (float)(std::log2(dp + 1) / 32);
    out[5] = (float)(std::log2(sb + 1) / 32);
    out[6] = (float)(std::log2(sp + 1) / 32);
    out[7] = (float)overlap;
    out[8] = (float)(std::log2(nb / std::max(db, 1.0) + 1) / 32);
    out[9] = (float)(std::log2(nb * np * overlap / std::max(std::max(db, dp), 1.0) + 1) / 32);
    out[10] = (float)(probe_key->rows ? (double)probe_key->nulls / (double)probe_key->rows : 0.0);
    out[11] = (float)(build_key->rows ? (double)build_key->nulls / (double)build_key->rows : 0.0);
  });
}
void ws_join_log_add(const float* features, int64_t pairs) {
  std::lock_guard<std::mutex> g(g_jlog_mu);
  g_jlog_x.insert(g_jlog_x.end(), features, features + kIn);
  g_jlog_t.push_back((float)(std::log2((double)std::max<int64_t>(pairs, 0) + 1.0) / 32.0));
}
int64_t ws_join_log_size(void) {
  std::lock_guard<std::mutex> g(g_jlog_mu);
  return (int64_t)g_jlog_t.size();
}
int64_t ws_join_log_copy(float* x, float* t, int64_t cap) {
  std::lock_guard<std::mutex> g(g_jlog_mu);
  const int64_t n = std::min<int64_t>(cap, (int64_t)g_jlog_t.size());
  if (x) std::memcpy(x, g_jlog_x.data(), sizeof(float) * kIn * n);
  if (t) std::memcpy(t, g_jlog_t.data(), sizeof(float) * n);
  return n;
}
void ws_join_log_clear(void) {
  std::lock_guard<std::mutex> g(g_jlog_mu);
  g_jlog_x.clear();
  g_jlog_t.clear();
}

int64_t ws_sizing_log_size(void) {
  std::lock_guard<std::mutex> g(g_log_mu);
  return (int64_t)g_log_t.size();
}
int64_t ws_sizing_log_copy(float* x, float* t, int64_t cap) {
  std::lock_guard<std::mutex> g(g_log_mu);
  const int64_t n = std::min<int64_t>(cap, (int64_t)g_log_t.size());
  if (x) std::memcpy(x, g_log_x.data(), sizeof(float) * kIn * n);
  if (t) std::memcpy(t, g_log_t.data(), sizeof(float) * n);
  return n;
}
void ws_sizing_log_clear(void) {
  std::lock_guard<std::mutex> g(g_log_mu);
  g_log_x.clear();
  g_log_t.clear();
}

}  // extern "C"
