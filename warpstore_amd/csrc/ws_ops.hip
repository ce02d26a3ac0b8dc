// Gathers with inline decode, dictionary code remap, join-filter bitmaps, per-row evaluation and
// GROUP BY temp columns (reference src/gpu.cu:149-163, 757-1087). Grid-stride kernels: each is one
// streaming pass, bound by HBM (or by the random side of a gather).
#include <algorithm>
#include <climits>
#include <cmath>

#include "ws_internal.h"

namespace ws {

// ------------------------------------------------------------------ gathers

// Gather from a batch-concatenated row space: binary search of the batch, decode if encoded.
__global__ void gather_multi_k(const ws_colview* src, const int64_t* boff, int nb, int t, int esize, const int32_t* idx,
                               int64_t m, uint8_t* dst) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = idx[j];
    int lo = 0, hi = nb - 1;  // last batch with boff <= g
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (boff[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    const ws_colview& v = src[lo];
    const int64_t r = g - boff[lo];
    if (v.enc) {
      if (t == WS_P_F64) ((double*)dst)[j] = enc_dbl(v, r);
      else if (t == WS_P_I32) ((int32_t*)dst)[j] = (int32_t)enc_int(v, r);
      else ((int64_t*)dst)[j] = enc_int(v, r);
    } else if (esize == 4) {
      ((int32_t*)dst)[j] = ((const int32_t*)v.p)[r];
    } else if (esize == 8) {
      ((int64_t*)dst)[j] = ((const int64_t*)v.p)[r];
    } else {
      const uint8_t* s = (const uint8_t*)v.p + r * esize;
      for (int k = 0; k < esize; ++k) dst[j * esize + k] = s[k];
    }
  }
}
__global__ void gather_enc_k(ws_colview v, int t, const int32_t* idx, int64_t m, void* dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx[i];
    if (t == WS_P_F64) ((double*)dst)[i] = enc_dbl(v, r);
    else if (t == WS_P_I32) ((int32_t*)dst)[i] = (int32_t)enc_int(v, r);
    else ((int64_t*)dst)[i] = enc_int(v, r);
  }
}
template <class T>
__global__ void gather_k(const T* src, const int32_t* idx, int64_t m, T* dst) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) dst[j] = src[idx[j]];
}
template <class T>
__global__ void gather_null_k(const T* src, const int32_t* idx, int64_t m, T* dst, T nul) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    dst[j] = idx[j] < 0 ? nul : src[idx[j]];
}
__global__ void gather_words_k(const uint32_t* src, int words, const int32_t* idx, int64_t m, uint32_t* dst, int nulls) {
  const int64_t n = m * words;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = j / words, w = j - r * words;
    const int32_t i = idx[r];
    dst[j] = i < 0 ? (nulls ? 0x7fc00000u : 0u) : src[(int64_t)i * words + w];  // float NaN marks a NULL vector
  }
}

// ------------------------------------------------------------------ codes, bits, bitmaps

__global__ void remap_codes_k(const int32_t* in, int64_t n, const int32_t* map, int64_t map_n, int32_t* out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = in[j];
    out[j] = c == WS_NULL_I32 ? WS_NULL_I32 : (c >= 0 && c < map_n) ? map[c] : -2;
  }
}
__global__ void bits_to_double_k(const double* src, int64_t n, double* dst) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    dst[j] = (double)dbits(src[j]);
}
__device__ __forceinline__ i64 key_at(const ws_colview& c, i64 r) {
  switch (c.t) {
    case WS_T_I32: case WS_T_I32N: return ((const int32_t*)c.p)[r];
    case WS_T_I64: case WS_T_I64N: case WS_T_DECIMAL: return ((const int64_t*)c.p)[r];
    default: return (i64)((const double*)c.p)[r];
  }
}
__device__ __forceinline__ bool key_null(const ws_colview& c, i64 r) {
  switch (c.t) {
    case WS_T_I32: case WS_T_I32N: return ((const int32_t*)c.p)[r] == WS_NULL_I32;
    case WS_T_I64: case WS_T_I64N: case WS_T_DECIMAL: return ((const int64_t*)c.p)[r] == WS_NULL_I64;
    default: {
      const double v = ((const double*)c.p)[r];
      return v != v;
    }
  }
}
__global__ void key_bitmap_k(ws_colview key, int64_t n, int64_t lo, int64_t nbits, uint32_t* bits) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (key_null(key, i)) continue;
    const int64_t v = key_at(key, i) - lo;
    if (v >= 0 && v < nbits) atomicOr(&bits[v >> 5], 1u << (v & 31));
  }
}
__global__ void set_bits_k(uint32_t* bits, const int32_t* idx, int64_t m) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    atomicOr(&bits[idx[j] >> 5], 1u << (idx[j] & 31));
}
__global__ void map_rank_k(double* keys, const int32_t* rank, int64_t m) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    keys[j] = keys[j] != keys[j] ? -1.0 : (double)rank[(int64_t)keys[j]];  // NULL first
}

// ------------------------------------------------------------------ evaluation

__global__ void eval_at_k(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, double* out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    out[j] = eval_prog(*env, env->cols, prog, idx ? idx[j] : j);
}
// A program that is one column (ORDER BY an aggregate): the VM's column load, without the VM's
// stack (which spills to scratch memory).
__global__ void eval_col_at_k(ws_colview c, const int32_t* idx, int64_t m, double* out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    out[j] = load_col(c, idx ? idx[j] : j);
}
// Exact int64 evaluation of integer columns, integer literals, + - * and negation (NULLs propagate,
// overflow raises env->ovf); false for anything else (then the double VM). Reference gpu.cu:826-866.
__device__ bool eval_int_exact(const ws_expr_env& env, ws_prog p, i64 r, i64& out) {
  i64 st[16];
  bool nul[16];
  int sp = 0;
  for (int i = 0; i < p.len; ++i) {
    const ws_instr in = env.code[p.start + i];
    switch (in.op) {
      case WS_OP_COL: {
        const ws_colview& c = env.cols[in.arg];
        if (c.t == WS_T_F64 || c.t == WS_T_VEC || c.t == WS_T_DECIMAL || sp >= 16) return false;
        const i64 v = col_int(c, r);
        st[sp] = v;
        nul[sp++] = (c.t == WS_T_I32N && v == (i64)WS_NULL_I32) || (c.t == WS_T_I64N && v == (i64)WS_NULL_I64);
        break;
      }
      case WS_OP_CONST:
        if (!in.is_int || sp >= 16) return false;
        st[sp] = dbits(env.consts[in.arg]);
        nul[sp++] = false;
        break;
      case WS_OP_ADD: case WS_OP_SUB: case WS_OP_MUL: {
        if (sp < 2) return false;
        --sp;
        const bool n = nul[sp - 1] || nul[sp];
        if (!n)
          st[sp - 1] = in.op == WS_OP_ADD ? add_chk(st[sp - 1], st[sp], env.ovf)
                       : in.op == WS_OP_SUB ? sub_chk(st[sp - 1], st[sp], env.ovf)
                                            : mul_chk(st[sp - 1], st[sp], env.ovf);
        nul[sp - 1] = n;
        break;
      }
      case WS_OP_NEG:
        if (sp < 1) return false;
        if (!nul[sp - 1]) st[sp - 1] = sub_chk(0, st[sp - 1], env.ovf);
        break;
      default: return false;
    }
  }
  if (sp != 1) return false;
  out = nul[0] ? (i64)WS_NULL_I64 : st[0];
  return true;
}
__global__ void eval_at_int_k(const ws_expr_env* envp, ws_prog prog, const int32_t* idx, int64_t m, int64_t* out) {
  const ws_expr_env& env = *envp;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx ? idx[j] : j;
    const ws_instr in = env.code[prog.start];
    i64 iv;
    if (prog.len == 1 && in.op == WS_OP_COL && env.cols[in.arg].t != WS_T_F64) {
      const ws_colview& c = env.cols[in.arg];
      const i64 v = col_int(c, r);
      out[j] = (v == (i64)WS_NULL_I32 && (c.t == WS_T_I32 || c.t == WS_T_I32N)) ? (i64)WS_NULL_I64 : v;
    } else if (eval_int_exact(env, prog, r, iv)) {
      out[j] = iv;
    } else {
      const double d = eval_prog(env, env.cols, prog, r);
      out[j] = d != d ? (i64)WS_NULL_I64 : llround(d);
    }
  }
}
__global__ void distinct_bits_k(const ws_expr_env* envp, ws_prog prog, const int32_t* idx, int64_t m, int64_t base, int64_t span,
                                unsigned long long* bits) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx ? idx[j] : j;
    i64 v;
    if (!eval_int_exact(*envp, prog, r, v) || v == (i64)WS_NULL_I64) continue;
    const int64_t k = v - base;
    if (k < 0 || k >= span) continue;
    const unsigned long long bit = 1ull << (k & 63);
    if (!(bits[k >> 6] & bit)) atomicOr(&bits[k >> 6], bit);
  }
}
__global__ void bits_from_col_k(const void* col, int out, int64_t n, long long base, long long nbits, uint32_t* bits, int* nul) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    long long v;
    if (out == 1) {
      v = ((const long long*)col)[i];
      if (v == (long long)WS_NULL_I64) {
        *nul = 1;
        continue;
      }
    } else {
      const int32_t x = ((const int32_t*)col)[i];
      if (x == WS_NULL_I32) {
        *nul = 1;
        continue;
      }
      v = x;
    }
    const long long k = v - base;
    if (k >= 0 && k < nbits) atomicOr(&bits[k >> 5], 1u << (k & 31));
  }
}
// A derived-table column from compacted GROUP BY results (column-major keys / accs).
__global__ void temp_col_k(const double* keys, int nk, const double* accs, int na, int64_t g, ws_temp_col_spec s, void* out,
                           long long* mnmx) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < g; r += (int64_t)gridDim.x * blockDim.x) {
    auto A = [&](int j) { return accs[(int64_t)j * g + r]; };
    bool nul = s.src == 1 && s.cnt >= 0 && A(s.cnt) == 0;
    bool is_int = false;
    i64 iv = 0;
    double dv = 0;
    if (s.src == 0) {
      const double k = keys[(int64_t)s.idx * g + r];
      if (s.key_int) {
        iv = dbits(k);
        is_int = true;
        nul = iv == (i64)WS_NULL_I64 || iv == (i64)WS_NULL_I32;
      } else {
        dv = k;
        nul = k != k;
      }
    } else if (!nul) {
      switch (s.kind) {
        case 1: iv = dbits(A(s.idx)); is_int = true; break;
        case 2: iv = dbits(A(s.hi)) * 4294967296ll + dbits(A(s.idx)); is_int = true; break;
        case 3: iv = (i64)A(s.idx); is_int = true; break;
        case 4: {
          const double sum = s.sum_bits ? (s.hi >= 0 ? (double)dbits(A(s.hi)) * 4294967296.0 + (double)dbits(A(s.idx))
                                                     : (double)dbits(A(s.idx)))
                                        : A(s.idx);
          dv = sum / s.div / A(s.cnt);
          break;
        }
        default: dv = A(s.idx) / s.div; nul = dv != dv;
      }
    }
    if (s.mul != 1 && !nul) {
      dv = (is_int ? (double)iv : dv) * s.mul;
      is_int = false;
    }
    if (s.out == 2) {
      ((double*)out)[r] = nul ? nullv() : is_int ? (double)iv : dv;
      continue;
    }
    const long long v = nul ? 0 : is_int ? iv : llround(dv);
    if (!nul) {
      atomicMin(&mnmx[0], v);
      atomicMax(&mnmx[1], v);
    }
    if (s.out == 1) ((long long*)out)[r] = nul ? (long long)WS_NULL_I64 : v;
    else ((int32_t*)out)[r] = nul ? WS_NULL_I32 : (int32_t)v;
  }
}

// Device copy of an env (the evaluation kernels read it through a pointer: 4.4 KB is too big for a
// kernel argument).
static DevBuf<ws_expr_env> env_on_device(const ws_expr_env& env) {
  DevBuf<ws_expr_env> d(1);
  h2d_async_staged(d.p, &env, sizeof env);
  return d;
}

}  // namespace ws

using namespace ws;

extern "C" {

int ws_gather_multi(const ws_colview* src, const int64_t* boff, int nb, int ptype, int esize, const int32_t* idx, int64_t m,
                    void* dst) {
  return guard([&] {
    if (m <= 0) return;
    std::vector<ws_colview> vs(src, src + nb);
    for (auto& v : vs) v.t = ptype == WS_P_I32 ? WS_T_I32N : ptype == WS_P_I64 ? WS_T_I64N : WS_T_F64;  // keep NULL sentinels
    Blob b;
    const size_t o_v = b.put(vs.data(), sizeof(ws_colview) * nb);
    const size_t o_o = b.put(boff, 8 * (size_t)nb);
    uint8_t* d = upload_blob(b);
    ProfScope ps("gather_multi");
    hipLaunchKernelGGL(gather_multi_k, dim3((unsigned)std::min<int64_t>((m + 255) / 256, 8192)), dim3(256), 0, stream(),
                       (const ws_colview*)(d + o_v), (const int64_t*)(d + o_o), nb, ptype, esize, idx, m, (uint8_t*)dst);
    tmp_free(d);
    check_launch("gather_multi");
  });
}
int ws_gather_enc(const ws_colview* v, int ptype, const int32_t* idx, int64_t m, void* dst) {
  return guard([&] {
    if (m <= 0) return;
    ws_colview w = *v;
    w.t = ptype == WS_P_I32 ? WS_T_I32N : ptype == WS_P_I64 ? WS_T_I64N : WS_T_F64;
    ProfScope ps("gather_enc");
    hipLaunchKernelGGL(gather_enc_k, dim3((unsigned)std::min<int64_t>((m + 255) / 256, 4096)), dim3(256), 0, stream(), w, ptype,
                       idx, m, dst);
    check_launch("gather_enc");
  });
}
int ws_gather(const void* src, int ptype, const int32_t* idx, int64_t m, void* dst) {
  return guard([&] {
    if (m <= 0) return;
    ProfScope ps("gather");
    if (ptype == WS_P_I32)
      hipLaunchKernelGGL(gather_k<int32_t>, dim3(grid_for(m)), dim3(256), 0, stream(), (const int32_t*)src, idx, m, (int32_t*)dst);
    else
      hipLaunchKernelGGL(gather_k<int64_t>, dim3(grid_for(m)), dim3(256), 0, stream(), (const int64_t*)src, idx, m, (int64_t*)dst);
    check_launch("gather");
  });
}
int ws_gather_null(const void* src, int ptype, const int32_t* idx, int64_t m, void* dst) {
  return guard([&] {
    if (m <= 0) return;
    ProfScope ps("gather_null");
    if (ptype == WS_P_I32)
      hipLaunchKernelGGL(gather_null_k<int32_t>, dim3(grid_for(m)), dim3(256), 0, stream(), (const int32_t*)src, idx, m,
                         (int32_t*)dst, (int32_t)WS_NULL_I32);
    else if (ptype == WS_P_I64)
      hipLaunchKernelGGL(gather_null_k<int64_t>, dim3(grid_for(m)), dim3(256), 0, stream(), (const int64_t*)src, idx, m,
                         (int64_t*)dst, (int64_t)WS_NULL_I64);
    else
      hipLaunchKernelGGL(gather_null_k<double>, dim3(grid_for(m)), dim3(256), 0, stream(), (const double*)src, idx, m,
                         (double*)dst, nullv());
    check_launch("gather_null");
  });
}
int ws_gather_bytes(const void* src, int esize, const int32_t* idx, int64_t m, void* dst, int nulls) {
  return guard([&] {
    if (m <= 0) return;
    if (esize % 4) throw Error("gather_bytes: element size must be a multiple of 4", WS_ERR_INTERNAL);
    ProfScope ps("gather_bytes");
    hipLaunchKernelGGL(gather_words_k, dim3(grid_for(m * (esize / 4))), dim3(256), 0, stream(), (const uint32_t*)src, esize / 4,
                       idx, m, (uint32_t*)dst, nulls);
    check_launch("gather_bytes");
  });
}
int ws_remap_codes(const int32_t* in, int64_t n, const int32_t* map, int64_t map_n, int32_t* out) {
  return guard([&] {
    if (n <= 0) return;
    ProfScope ps("remap_codes");
    hipLaunchKernelGGL(remap_codes_k, dim3(grid_for(n)), dim3(256), 0, stream(), in, n, map, map_n, out);
    check_launch("remap_codes");
  });
}
int ws_bits_to_double(const double* src, int64_t n, double* dst) {
  return guard([&] {
    if (n <= 0) return;
    ProfScope ps("bits_to_double");
    hipLaunchKernelGGL(bits_to_double_k, dim3(grid_for(n)), dim3(256), 0, stream(), src, n, dst);
    check_launch("bits_to_double");
  });
}
int ws_key_bitmap(ws_colview key, int64_t n, int64_t lo, int64_t nbits, uint32_t* bits) {
  return guard([&] {
    fill_u32(bits, (nbits + 31) / 32, 0);
    if (n <= 0) return;
    ProfScope ps("key_bitmap");
    hipLaunchKernelGGL(key_bitmap_k, dim3(grid_for(n)), dim3(256), 0, stream(), key, n, lo, nbits, bits);
    check_launch("key_bitmap");
  });
}
int ws_set_bits(uint32_t* bits, const int32_t* idx, int64_t m) {
  return guard([&] {
    if (m <= 0) return;
    ProfScope ps("set_bits");
    hipLaunchKernelGGL(set_bits_k, dim3(grid_for(m)), dim3(256), 0, stream(), bits, idx, m);
    check_launch("set_bits");
  });
}
int ws_map_rank(double* keys, const int32_t* rank, int64_t m) {
  return guard([&] {
    if (m <= 0) return;
    ProfScope ps("map_rank");
    hipLaunchKernelGGL(map_rank_k, dim3(grid_for(m)), dim3(256), 0, stream(), keys, rank, m);
    check_launch("map_rank");
  });
}
int ws_eval_at(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, double* out) {
  return guard([&] {
    if (m <= 0) return;
    if (prog.len == 1 && prog.start >= 0 && prog.start < WS_MAX_CODE && env->code[prog.start].op == WS_OP_COL &&
        env->code[prog.start].arg < WS_MAX_COLS) {
      ProfScope ps("eval_at");
      hipLaunchKernelGGL(eval_col_at_k, dim3(grid_for(m)), dim3(256), 0, stream(), env->cols[env->code[prog.start].arg], idx, m,
                         out);
      check_launch("eval_at");
      return;
    }
    DevBuf<ws_expr_env> d = env_on_device(*env);
    ProfScope ps("eval_at");
    hipLaunchKernelGGL(eval_at_k, dim3(grid_for(m)), dim3(256), 0, stream(), (const ws_expr_env*)d.p, prog, idx, m, out);
    check_launch("eval_at");
  });
}
int ws_eval_at_int(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, int64_t* out) {
  return guard([&] {
    if (m <= 0) return;
    DevBuf<ws_expr_env> d = env_on_device(*env);
    ProfScope ps("eval_at_int");
    hipLaunchKernelGGL(eval_at_int_k, dim3(grid_for(m)), dim3(256), 0, stream(), (const ws_expr_env*)d.p, prog, idx, m, out);
    check_launch("eval_at_int");
  });
}
int ws_distinct_ints(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, int64_t base, int64_t span,
                     int64_t* out, int64_t out_cap, int64_t* n_out) {
  return guard([&] {
    *n_out = 0;
    if (m <= 0 || span <= 0) return;
    const int64_t words = (span + 63) / 64;
    DevBuf<unsigned long long> bits(words);
    fill_u64((u64*)bits.p, words, 0);
    DevBuf<ws_expr_env> d = env_on_device(*env);
    {
      ProfScope ps("distinct_bits");
      hipLaunchKernelGGL(distinct_bits_k, dim3(grid_for(m)), dim3(256), 0, stream(), (const ws_expr_env*)d.p, prog, idx, m, base,
                         span, bits.p);
      check_launch("distinct_bits");
    }
    std::vector<unsigned long long> h(words);
    d2h_sync(h.data(), bits.p, 8 * (size_t)words);
    int64_t k = 0;
    for (int64_t w = 0; w < words; ++w)
      for (unsigned long long b = h[w]; b; b &= b - 1) {
        if (k < out_cap) out[k] = base + w * 64 + __builtin_ctzll(b);
        ++k;
      }
    *n_out = k;
  });
}
int ws_bits_from_column(const void* col, int out, int64_t n, int64_t base, int64_t nbits, int* saw_null, uint32_t** bits) {
  return guard([&] {
    *bits = nullptr;
    const size_t bytes = std::max<size_t>(4, (size_t)((nbits + 31) / 32) * 4);
    uint32_t* b = nullptr;
    WS_CK(hipMalloc((void**)&b, bytes));
    try {
      fill_u32(b, (int64_t)(bytes / 4), 0);
      DevBuf<int> nul(1);
      fill_u32((uint32_t*)nul.p, 1, 0);
      if (n > 0) {
        ProfScope ps("bits_from_column");
        hipLaunchKernelGGL(bits_from_col_k, dim3(grid_for(n)), dim3(256), 0, stream(), col, out, n, (long long)base,
                           (long long)nbits, b, nul.p);
        check_launch("bits_from_column");
      }
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
    *out = nullptr;
    const size_t w = s->out == 0 ? 4 : 8;
    void* o = nullptr;
    WS_CK(hipMalloc(&o, std::max<size_t>(1, (size_t)g * w)));
    try {
      DevBuf<long long> mm(2);
      const long long init[2] = {LLONG_MAX, LLONG_MIN};
      h2d_async_staged(mm.p, init, sizeof init);
      if (g > 0) {
        ProfScope ps("temp_column");
        hipLaunchKernelGGL(temp_col_k, dim3(grid_for(g)), dim3(256), 0, stream(), keys, nk, accs, na, g, *s, o, mm.p);
        check_launch("temp_column");
      }
      long long h[2];
      d2h_sync(h, mm.p, sizeof h);
      *mn = s->out == 2 || h[0] > h[1] ? -INFINITY : (double)h[0];
      *mx = s->out == 2 || h[0] > h[1] ? INFINITY : (double)h[1];
    } catch (...) {
      (void)hipFree(o);
      throw;
    }
    *out = o;
  });
}

}  // extern "C"
