// CPU baseline of the measured query shapes. TEST INFRASTRUCTURE ONLY: timed beside the GPU path
// by bench.py (kind "port") and checked against the single-thread oracle restatements in
// ws_oracle.cpp (or_c1 / or_c2 / or_c3 shapes) by tests/test_oracle.py; the product path never
// calls it.
//
// Each query runs on `threads` host threads the way a CPU column engine would: every thread
// filters and pre-aggregates a contiguous share of the rows into private state, then the shares
// are merged. Semantics are those of the single-thread restatements (reference query shapes:
// engine_select.cpp:1403-1492 for COUNT DISTINCT, engine_core.h:4514-4611 for the inner join with
// build-side multiplicity, kernels.cuh:188-633 for the slot semantics); only the work split differs.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

typedef long long i64;

template <class F>
void parallel_ranges(int threads, int64_t n, F&& f) {
  threads = std::max(1, threads);
  if (threads == 1) {
    f(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    const int64_t a = n * t / threads, b = n * (t + 1) / threads;
    pool.emplace_back([&f, t, a, b] { f(t, a, b); });
  }
  for (auto& th : pool) th.join();
}

uint64_t mix(uint64_t x) {
  x ^= x >> 31;
  x *= 0x7fb5d329728ea185ull;
  x ^= x >> 27;
  x *= 0x81dadef4bc2dd44dull;
  return x ^ (x >> 33);
}

}  // namespace

extern "C" {

int cb_hardware_threads(void) { return (int)std::max(1u, std::thread::hardware_concurrency()); }

// SELECT k, COUNT(*), SUM(v), MIN(v), MAX(v), SUM(p) FROM t WHERE k < k_lt AND v > v_gt GROUP BY k.
// out rows (ascending k): k, count, sum v, min v, max v, then sums of p in psum[]. Returns groups.
int64_t cb_c1(const int32_t* k, const int64_t* v, const double* p, int64_t n, int32_t k_lt, int64_t v_gt, int threads,
              int64_t* out, double* psum, int64_t cap) {
  struct A {
    int64_t cnt = 0, sum = 0, mn = INT64_MAX, mx = INT64_MIN;
    double ps = 0;
  };
  std::vector<std::unordered_map<int32_t, A>> part(std::max(1, threads));
  parallel_ranges(threads, n, [&](int t, int64_t a, int64_t b) {
    auto& m = part[t];
    for (int64_t i = a; i < b; ++i) {
      if (!(k[i] < k_lt && v[i] > v_gt)) continue;
      A& g = m[k[i]];
      ++g.cnt;
      g.sum += v[i];
      g.mn = std::min(g.mn, v[i]);
      g.mx = std::max(g.mx, v[i]);
      g.ps += p[i];
    }
  });
  std::unordered_map<int32_t, A> all;
  for (auto& m : part)
    for (auto& [key, g] : m) {
      A& d = all[key];
      d.cnt += g.cnt, d.sum += g.sum, d.mn = std::min(d.mn, g.mn), d.mx = std::max(d.mx, g.mx), d.ps += g.ps;
    }
  std::vector<int32_t> keys;
  for (auto& kv : all) keys.push_back(kv.first);
  std::sort(keys.begin(), keys.end());
  if ((int64_t)keys.size() > cap) return -(int64_t)keys.size();
  for (size_t i = 0; i < keys.size(); ++i) {
    const A& g = all[keys[i]];
    out[i * 5 + 0] = keys[i], out[i * 5 + 1] = g.cnt, out[i * 5 + 2] = g.sum, out[i * 5 + 3] = g.mn, out[i * 5 + 4] = g.mx;
    psum[i] = g.ps;
  }
  return (int64_t)keys.size();
}

// SELECT user_id, COUNT(*), SUM(dur), COUNT(DISTINCT url) FROM h WHERE event_date < date_lt AND
// region < region_lt GROUP BY user_id, every group ordered by count desc then user_id asc.
// Threads filter their share and route (user, url, dur) by a hash of user_id to `threads`
// partitions; each thread then sorts and groups one partition.
int64_t cb_c2(const int64_t* user, const int32_t* region, const int32_t* date, const int32_t* url, const int32_t* dur,
              int64_t n, int32_t date_lt, int32_t region_lt, int threads, int64_t* out, int64_t cap) {
  struct R {
    int64_t user;
    int32_t url, dur;
  };
  const int T = std::max(1, threads);
  std::vector<std::vector<std::vector<R>>> bins(T, std::vector<std::vector<R>>(T));
  parallel_ranges(T, n, [&](int t, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i)
      if (date[i] < date_lt && region[i] < region_lt)
        bins[t][mix((uint64_t)user[i]) % (uint64_t)T].push_back({user[i], url[i], dur[i]});
  });
  std::vector<std::vector<std::array<int64_t, 4>>> groups(T);
  parallel_ranges(T, T, [&](int, int64_t a, int64_t b) {
    for (int64_t p = a; p < b; ++p) {
      std::vector<R> rows;
      for (int t = 0; t < T; ++t) rows.insert(rows.end(), bins[t][p].begin(), bins[t][p].end());
      std::sort(rows.begin(), rows.end(),
                [](const R& x, const R& y) { return x.user != y.user ? x.user < y.user : x.url < y.url; });
      for (size_t i = 0; i < rows.size();) {
        size_t j = i;
        int64_t cnt = 0, sum = 0, distinct = 0;
        for (; j < rows.size() && rows[j].user == rows[i].user; ++j) {
          distinct += j == i || rows[j].url != rows[j - 1].url;
          ++cnt;
          sum += rows[j].dur;
        }
        groups[p].push_back({rows[i].user, cnt, sum, distinct});
        i = j;
      }
    }
  });
  std::vector<std::array<int64_t, 4>> g;
  for (auto& v : groups) g.insert(g.end(), v.begin(), v.end());
  std::sort(g.begin(), g.end(), [](const std::array<int64_t, 4>& a, const std::array<int64_t, 4>& b) {
    return a[1] != b[1] ? a[1] > b[1] : a[0] < b[0];
  });
  if ((int64_t)g.size() > cap) return -(int64_t)g.size();
  for (size_t i = 0; i < g.size(); ++i)
    for (int c = 0; c < 4; ++c) out[i * 4 + c] = g[i][c];
  return (int64_t)g.size();
}

// SELECT f.flag, COUNT(*), SUM(f.price), SUM(f.qty) FROM fact f JOIN dim d ON f.fk = d.key WHERE
// d.attr < attr_lt GROUP BY f.flag: rows (flag, count, raw price sum, qty sum) ordered by flag.
// Build: the filtered dimension keys with their multiplicity, direct-addressed over the key span
// when it is below 2^28 (else a hash map); probe: each thread scans a share of the fact rows.
int64_t cb_c3(const int32_t* fk, const int64_t* qty, const int64_t* price, const int32_t* flag, int64_t nf,
              const int32_t* dkey, const int32_t* dattr, int64_t nd, int32_t attr_lt, int threads, int64_t* out,
              int64_t cap) {
  int64_t kmin = INT64_MAX, kmax = INT64_MIN;
  for (int64_t i = 0; i < nd; ++i)
    if (dattr[i] < attr_lt) kmin = std::min<int64_t>(kmin, dkey[i]), kmax = std::max<int64_t>(kmax, dkey[i]);
  const bool direct = kmin <= kmax && kmax - kmin < (1ll << 28);
  std::vector<uint32_t> mult(direct ? (size_t)(kmax - kmin + 1) : 0, 0);
  std::unordered_map<int32_t, uint32_t> hmult;
  for (int64_t i = 0; i < nd; ++i)
    if (dattr[i] < attr_lt) {
      if (direct) ++mult[(size_t)(dkey[i] - kmin)];
      else ++hmult[dkey[i]];
    }
  struct A {
    __int128 cnt = 0, price = 0, qty = 0;
  };
  const int T = std::max(1, threads);
  std::vector<std::unordered_map<int32_t, A>> part(T);
  parallel_ranges(T, nf, [&](int t, int64_t a, int64_t b) {
    auto& m = part[t];
    for (int64_t i = a; i < b; ++i) {
      uint32_t c = 0;
      if (direct) {
        const int64_t o = (int64_t)fk[i] - kmin;
        if (o >= 0 && o <= kmax - kmin) c = mult[(size_t)o];
      } else if (!hmult.empty()) {
        auto it = hmult.find(fk[i]);
        if (it != hmult.end()) c = it->second;
      }
      if (!c) continue;
      A& g = m[flag[i]];
      g.cnt += c;
      g.price += (__int128)price[i] * c;
      g.qty += (__int128)qty[i] * c;
    }
  });
  std::unordered_map<int32_t, A> all;
  for (auto& m : part)
    for (auto& [f, g] : m) {
      A& d = all[f];
      d.cnt += g.cnt, d.price += g.price, d.qty += g.qty;
    }
  std::vector<int32_t> flags;
  for (auto& kv : all) flags.push_back(kv.first);
  std::sort(flags.begin(), flags.end());
  if ((int64_t)flags.size() > cap) return -(int64_t)flags.size();
  for (size_t i = 0; i < flags.size(); ++i) {
    const A& g = all[flags[i]];
    out[i * 4 + 0] = flags[i], out[i * 4 + 1] = (int64_t)g.cnt, out[i * 4 + 2] = (int64_t)g.price,
    out[i * 4 + 3] = (int64_t)g.qty;
  }
  return (int64_t)flags.size();
}

}  // extern "C"
