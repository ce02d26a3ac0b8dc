// Segment column encodings as the storage layer writes them (reference src/storage.cu:1097-1194,
// Table::encodeColumn and encodeDecimalDoubles); the kernels decode them inline (ws_vm.h). Host
// code: segments are encoded when they are sealed, before upload.
//
// Formats (a 16-byte header, then the codes, then 16 bytes of padding so vector loads of the last
// codes stay inside the buffer):
//   1 / 2 / 3  integer frame of reference: header int64 base (the minimum); u8 / u16 / u32 codes
//              value - base, the top code meaning NULL (3 only for int64 columns);
//   4          double dictionary of at most 255 distinct values: 256 doubles after the header, then
//              u8 codes;
//   5 / 6      decimal doubles: header int64 base, double 10^s; u16 / u32 codes k - base with
//              k = round(x * 10^s), the top code meaning NULL; chosen with the smallest s (0..4)
//              for which every value satisfies k / 10^s == x exactly (and |x * 10^s| < 4e15).
//
// One pass gathers everything the choice needs: integer min / max, or, for doubles, the distinct
// values (a 512-slot open-addressing table that gives up at the 256th) together with the exact
// decimal candidates for all five scales at once.
#include <cmath>
#include <cstring>

#include "ws_internal.h"

namespace ws {
namespace {

constexpr int64_t kHeader = WS_ENC_HEADER;
constexpr int64_t kPad = 16;

template <class Code>
void pack_codes(uint8_t* dst, int64_t n, const int64_t* vals, int64_t base, int64_t null_value) {
  Code* c = (Code*)dst;
  const Code top = (Code)~(Code)0;
  for (int64_t i = 0; i < n; ++i) c[i] = vals[i] == null_value ? top : (Code)(uint64_t)(vals[i] - base);
}

int encode_ints(const void* col, bool wide, int64_t n, std::vector<uint8_t>& out) {
  const int64_t nul = wide ? (int64_t)WS_NULL_I64 : (int64_t)WS_NULL_I32;
  std::vector<int64_t> v((size_t)n);
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  for (int64_t i = 0; i < n; ++i) {
    v[(size_t)i] = wide ? ((const int64_t*)col)[i] : (int64_t)((const int32_t*)col)[i];
    if (v[(size_t)i] == nul) continue;
    lo = std::min(lo, v[(size_t)i]);
    hi = std::max(hi, v[(size_t)i]);
  }
  if (lo > hi) lo = hi = 0;  // every row NULL
  const uint64_t span = (uint64_t)hi - (uint64_t)lo;
  const int enc = span < 0xFFull ? 1 : span < 0xFFFFull ? 2 : (wide && span < 0xFFFFFFFFull) ? 3 : 0;
  if (!enc) return 0;
  const int w = enc_width((uint8_t)enc);
  out.assign((size_t)(kHeader + n * w + kPad), 0);
  std::memcpy(out.data(), &lo, 8);
  uint8_t* d = out.data() + kHeader;
  if (w == 1) pack_codes<uint8_t>(d, n, v.data(), lo, nul);
  else if (w == 2) pack_codes<uint16_t>(d, n, v.data(), lo, nul);
  else pack_codes<uint32_t>(d, n, v.data(), lo, nul);
  return enc;
}

// Distinct double bit patterns, up to 255 (the dictionary's capacity).
struct SmallDict {
  uint64_t key[512];
  int16_t code[512];
  int size = 0;
  bool full = false;
  SmallDict() { std::memset(code, -1, sizeof code); }
  int find_or_add(uint64_t bits) {
    uint64_t h = bits * 0x9e3779b97f4a7c15ull;
    for (unsigned s = (unsigned)(h >> 55);; s = (s + 1) & 511) {
      if (code[s] < 0) {
        if (size == 255) {
          full = true;
          return -1;
        }
        key[s] = bits;
        code[s] = (int16_t)size;
        return size++;
      }
      if (key[s] == bits) return code[s];
    }
  }
};

int encode_doubles(const double* x, int64_t n, std::vector<uint8_t>& out) {
  SmallDict dict;
  std::vector<uint8_t> dcode((size_t)n);
  double table[256] = {};
  // decimal candidates: scale s is feasible while every value is an exact s-digit decimal
  bool dec_ok[5] = {true, true, true, true, true};
  int64_t dlo[5], dhi[5];
  for (int s = 0; s < 5; ++s) dlo[s] = INT64_MAX, dhi[s] = INT64_MIN;
  for (int64_t i = 0; i < n; ++i) {
    const double v = x[i];
    if (!dict.full) {
      uint64_t bits;
      std::memcpy(&bits, &v, 8);
      const int c = dict.find_or_add(bits);
      if (c >= 0) {
        dcode[(size_t)i] = (uint8_t)c;
        table[c] = v;
      }
    }
    if (v != v) continue;  // NULL
    for (int s = 0; s < 5; ++s) {
      if (!dec_ok[s]) continue;
      const double p = pow10i(s), y = v * p;
      if (!(std::fabs(y) < 4e15)) {
        dec_ok[s] = false;
        continue;
      }
      const int64_t k = std::llround(y);
      if ((double)k / p != v) {
        dec_ok[s] = false;
        continue;
      }
      dlo[s] = std::min(dlo[s], k);
      dhi[s] = std::max(dhi[s], k);
    }
  }
  if (!dict.full) {
    out.assign((size_t)(kHeader + 256 * 8 + n + kPad), 0);
    std::memcpy(out.data() + kHeader, table, sizeof table);
    std::memcpy(out.data() + kHeader + 256 * 8, dcode.data(), (size_t)n);
    return 4;
  }
  for (int s = 0; s < 5; ++s) {
    if (!dec_ok[s]) continue;
    int64_t lo = dlo[s], hi = dhi[s];
    if (lo > hi) lo = hi = 0;
    const uint64_t span = (uint64_t)(hi - lo);
    const int enc = span < 0xFFFFull ? 5 : span < 0xFFFFFFFFull ? 6 : 0;
    if (!enc) return 0;  // a larger scale only widens the span
    const int w = enc_width((uint8_t)enc);
    const double p = pow10i(s);
    out.assign((size_t)(kHeader + n * w + kPad), 0);
    std::memcpy(out.data(), &lo, 8);
    std::memcpy(out.data() + 8, &p, 8);
    std::vector<int64_t> k((size_t)n);
    const int64_t nul = INT64_MIN;  // marks NULL rows for the packer
    for (int64_t i = 0; i < n; ++i) k[(size_t)i] = x[i] != x[i] ? nul : std::llround(x[i] * p);
    if (w == 2) pack_codes<uint16_t>(out.data() + kHeader, n, k.data(), lo, nul);
    else pack_codes<uint32_t>(out.data() + kHeader, n, k.data(), lo, nul);
    return enc;
  }
  return 0;
}

}  // namespace
}  // namespace ws

using namespace ws;

extern "C" {

size_t ws_encoded_bytes_max(int64_t n) { return (size_t)(kHeader + 256 * 8 + (n > 0 ? n : 0) * 4 + kPad); }

int ws_encode_column(const void* col, int ptype, int64_t n, uint8_t* out, size_t cap, int* enc, size_t* bytes) {
  return guard([&] {
    *enc = 0;
    *bytes = 0;
    if (n < 0) throw Error("encode: negative row count", WS_ERR_INTERNAL);
    std::vector<uint8_t> buf;
    int e = 0;
    if (ptype == WS_P_I32 || ptype == WS_P_I64) e = encode_ints(col, ptype == WS_P_I64, n, buf);
    else if (ptype == WS_P_F64) e = encode_doubles((const double*)col, n, buf);
    else throw Error("encode: unsupported column type", WS_ERR_INTERNAL);
    if (!e) return;
    if (buf.size() > cap) throw Error("encode: output buffer too small", WS_ERR_INTERNAL);
    std::memcpy(out, buf.data(), buf.size());
    *enc = e;
    *bytes = buf.size();
  });
}

}  // extern "C"
