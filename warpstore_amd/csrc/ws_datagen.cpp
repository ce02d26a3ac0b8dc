// Seeded synthetic inputs for the measured workloads, generated on the host and uploaded.
//
// Every value is a pure function of (seed, column, row): a counter-based generator built from the
// public splitmix64 finalizer. So the CPU checker, the CPU baseline and the GPU see identical bytes,
// and any slice can be regenerated independently. Distributions:
//   uniform ints   : floor(u * k) with a 128-bit multiply of the 64-bit draw
//   Zipf(s) ranks  : inverse CDF over a precomputed cumulative table (binary search)
//   lognormal      : exp(mu + sigma * z), z from Box-Muller on two uniforms
// Also the segment column encoder (frame of reference, the engine's storage encoding, reference
// storage.cu:1138-1170), used to store the web table's region column as uint16 codes.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "warpstore_amd/ws_abi.h"

namespace {

typedef unsigned long long u64;

inline u64 mix(u64 x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}
inline u64 draw(u64 seed, u64 col, u64 row) {
  return mix(mix(seed * 0x9e3779b97f4a7c15ull + col + 1) ^ (row * 0xd1b54a32d192ed03ull + 0x632be59bd9b4e019ull));
}
inline double unif(u64 seed, u64 col, u64 row) { return (double)(draw(seed, col, row) >> 11) * 0x1p-53; }
inline u64 below(u64 seed, u64 col, u64 row, u64 k) {
  return (u64)(((unsigned __int128)draw(seed, col, row) * k) >> 64);
}
inline double normal(u64 seed, u64 col, u64 row) {
  const double u1 = 1.0 - unif(seed, col, row);  // (0, 1]
  const double u2 = unif(seed, col + 1000, row);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Zipf {
  std::vector<double> cdf;
  Zipf(int64_t n, double s) : cdf(n) {
    double acc = 0;
    for (int64_t k = 0; k < n; ++k) {
      acc += std::pow((double)(k + 1), -s);
      cdf[k] = acc;
    }
    for (auto& c : cdf) c /= acc;
  }
  int64_t rank(double u) const {  // 0-based rank
    const int64_t i = std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
    return std::min<int64_t>(i, (int64_t)cdf.size() - 1);
  }
};

// Fisher-Yates permutation of [0, n) from the generator.
std::vector<int64_t> permutation(u64 seed, u64 col, int64_t n) {
  std::vector<int64_t> p(n);
  for (int64_t i = 0; i < n; ++i) p[i] = i;
  for (int64_t i = n - 1; i > 0; --i) std::swap(p[i], p[(int64_t)below(seed, col, (u64)i, (u64)i + 1)]);
  return p;
}

}  // namespace

extern "C" {

// Plain filter + GROUP BY table: k int32 U[0,1000), v int64 U[0,1e6), p float64 U[0,1000).
void wsg_c1(uint64_t seed, int64_t n, int32_t* k, int64_t* v, double* p) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    k[i] = (int32_t)below(seed, 1, i, 1000);
    v[i] = (int64_t)below(seed, 2, i, 1000000);
    p[i] = unif(seed, 3, i) * 1000.0;
  }
}

// Web-analytics table: user_id int64 = perm[Zipf(1.1) rank over n_users]; region U[0,256);
// event_date = 19723 + U[0,365); url code = Zipf(1.1) rank over n_urls; dur = round(lognormal(ln 30000, 1)).
void wsg_c2(uint64_t seed, int64_t n, int64_t n_users, int64_t n_urls, int64_t* user_id, int32_t* region, int32_t* event_date,
            int32_t* url, int32_t* dur) {
  const Zipf zu(n_users, 1.1), zl(n_urls, 1.1);
  const std::vector<int64_t> perm = permutation(seed, 99, n_users);
  const double mu = std::log(30000.0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    user_id[i] = perm[zu.rank(unif(seed, 1, i))];
    region[i] = (int32_t)below(seed, 2, i, 256);
    event_date[i] = 19723 + (int32_t)below(seed, 3, i, 365);
    url[i] = (int32_t)zl.rank(unif(seed, 4, i));
    dur[i] = (int32_t)std::llround(std::exp(mu + normal(seed, 5, i)));
  }
}

// Star schema: dim key = shuffled [0, n_dim), attr U[0,1000); fact fk U over dim keys, qty U[1,50],
// price (DECIMAL(15,2) as scaled int64) U[100, 10000000], flag U[0,25) (stored as int32).
void wsg_c3_dim(uint64_t seed, int64_t n_dim, int32_t* key, int32_t* attr) {
  const std::vector<int64_t> perm = permutation(seed, 98, n_dim);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_dim; ++i) {
    key[i] = (int32_t)perm[i];
    attr[i] = (int32_t)below(seed, 6, i, 1000);
  }
}
void wsg_c3_fact(uint64_t seed, int64_t n_fact, int64_t n_dim, int32_t* fk, int64_t* qty, int64_t* price, int32_t* flag) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_fact; ++i) {
    fk[i] = (int32_t)below(seed, 7, i, (u64)n_dim);
    qty[i] = 1 + (int64_t)below(seed, 8, i, 50);
    price[i] = 100 + (int64_t)below(seed, 9, i, 10000000 - 100 + 1);
    flag[i] = (int32_t)below(seed, 10, i, 25);
  }
}

// Batched queries: rows_q = clamp(round(lognormal(ln 131072, 1.25)), 4096, 4194304);
// distinct keys d_q = round(10^U[1,5]); threshold theta_q = U[1e5, 9e5].
void wsg_c4_shapes(uint64_t seed, int64_t nq, int64_t* rows, int64_t* distinct, int64_t* theta) {
  for (int64_t q = 0; q < nq; ++q) {
    const double r = std::exp(std::log(131072.0) + 1.25 * normal(seed, 20, q));
    rows[q] = std::min<int64_t>(4194304, std::max<int64_t>(4096, std::llround(r)));
    distinct[q] = std::llround(std::pow(10.0, 1.0 + 4.0 * unif(seed, 22, q)));
    theta[q] = 100000 + (int64_t)(unif(seed, 23, q) * 800000.0);
  }
}
// One batched query's segment: key int32 U[0, d), val int64 U[0, 1e6).
void wsg_c4_data(uint64_t seed, int64_t q, int64_t rows, int64_t distinct, int32_t* key, int64_t* val) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < rows; ++i) {
    key[i] = (int32_t)below(seed + 1000003ull * (u64)(q + 1), 24, i, (u64)distinct);
    val[i] = (int64_t)below(seed + 1000003ull * (u64)(q + 1), 25, i, 1000000);
  }
}

}  // extern "C"
