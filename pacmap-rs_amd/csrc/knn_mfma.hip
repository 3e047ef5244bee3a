// knn_mfma.hip -- certified exact kNN with MFMA candidate generation (gfx950).
//
// The reference's exact kNN (knn.rs:137-194) ranks every pair by the FP32
// distance of distance.rs:50-75.  Recomputing all n^2 distances in that order
// is VALU work; here the n^2 part runs on the matrix cores and only a short
// candidate list per row is recomputed in the exact order:
//
//  1. prep: rows are centred (c = x - mean, distances are unchanged) and split
//     into bf16 hi + lo parts stored in MFMA fragment order; per row
//     a = |c|^2 (1 - kappa), where kappa bounds every rounding error of step 2
//     relative to |c_q|^2 + |c_j|^2.
//  2. candidates: lb(q, j) = a_q + a_j - 2 <c_q, c_j> via 3 bf16 MFMAs
//     (hi*hi + hi*lo + lo*hi, FP32 accumulation) is a lower bound of the true
//     squared distance.  A threshold tau_q is taken from a 1/ss sample of the
//     candidate tiles (the 16th smallest lb), then every j with lb <= tau_q is
//     appended to the row's list.
//  3. rerank: the listed candidates get the exact FP32 distance of the
//     reference, are sorted by (distance, index), and the row is certified
//     when every unlisted candidate must be strictly farther than the k-th
//     (sqrt(tau) with rounding margins > D_k).  Uncertified rows are redone by
//     the exact direct kernel, so the result always equals the exact path.
#include "internal.h"

#include <cstdint>
#include <hipcub/hipcub.hpp>
#include <type_traits>
#include <vector>
#include <algorithm>

namespace pm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t tkey_m(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_tkey_m(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(b);
}
__device__ __forceinline__ uint64_t shfl_xor_u64m(uint64_t v, int m) {
  uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
  lo = __shfl_xor(lo, m, 64);
  hi = __shfl_xor(hi, m, 64);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_u64m(uint64_t v, int src) {
  uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
  lo = __shfl(lo, src, 64);
  hi = __shfl(hi, src, 64);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

template <int R>
__device__ __forceinline__ void wave_bitonic_m(uint64_t (&v)[R], int lane) {
  constexpr int N = 64 * R;
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 64) {
        const int rs = stride >> 6;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((r & rs) == 0) {
            const bool asc = (((r * 64 + lane) & size) == 0);
            const uint64_t a = v[r], b = v[r | rs];
            const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
            v[r] = asc ? lo : hi;
            v[r | rs] = asc ? hi : lo;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint64_t o = shfl_xor_u64m(v[r], stride);
          const bool asc = (((r * 64 + lane) & size) == 0);
          const bool lower = (lane & stride) == 0;
          const uint64_t lo = v[r] < o ? v[r] : o, hi = v[r] < o ? o : v[r];
          v[r] = (lower == asc) ? lo : hi;
        }
      }
    }
  }
}

template <int R>
__device__ __forceinline__ uint64_t wave_pick_m(const uint64_t (&v)[R], uint32_t pos) {
  uint64_t sel = 0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (static_cast<uint32_t>(r) == (pos >> 6)) sel = v[r];
  return shfl_u64m(sel, static_cast<int>(pos & 63u));
}

// Ascending bitonic sort of 32 floats held by one 16-lane segment, element
// e = reg * 16 + (lane & 15).
__device__ __forceinline__ void seg32_sort(float &v0, float &v1, int lane) {
  const int li = lane & 15;
#pragma unroll
  for (int size = 2; size <= 32; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride == 16) {
        const bool asc0 = ((li & size) == 0);  // element li vs 16 + li (size 32: ascending)
        const float lo = fminf(v0, v1), hi = fmaxf(v0, v1);
        v0 = asc0 ? lo : hi;
        v1 = asc0 ? hi : lo;
      } else {
        const float o0 = __shfl_xor(v0, stride, 64), o1 = __shfl_xor(v1, stride, 64);
        const bool lower = (li & stride) == 0;
        const bool asc0 = ((li & size) == 0), asc1 = (((16 + li) & size) == 0);
        v0 = (lower == asc0) ? fminf(v0, o0) : fmaxf(v0, o0);
        v1 = (lower == asc1) ? fminf(v1, o1) : fmaxf(v1, o1);
      }
    }
  }
}

__device__ __forceinline__ uint16_t bf16_rn(float f) {
  const uint32_t b = __float_as_uint(f);
  const uint32_t r = b + 0x7fffu + ((b >> 16) & 1u);  // finite inputs only
  return static_cast<uint16_t>(r >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// ---------------------------------------------------------------------------
// prep: centred bf16 hi/lo fragments, a = |c|^2 (1 - kappa), finiteness flag
// ---------------------------------------------------------------------------
// Fragment order: [row tile t][32-feature chunk kc][lane l][8 values], with
// row = 16 t + (l & 15) and feature = 32 kc + 8 (l >> 4) + e.  One wave-wide
// 16-byte load per lane fetches an A or B operand of v_mfma_f32_16x16x32_bf16.
__global__ void knn_frag_kernel(const float *__restrict__ x, uint64_t n, uint32_t d,
                                const double *__restrict__ mean, uint32_t kch, uint64_t ntiles,
                                const uint32_t *__restrict__ perm, uint4 *__restrict__ fhi,
                                uint4 *__restrict__ flo) {
  const uint64_t total = ntiles * kch * 64;
  for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < total; g += gridDim.x * 256ull) {
    const uint32_t l = static_cast<uint32_t>(g & 63);
    const uint64_t tk = g >> 6;
    const uint32_t kc = static_cast<uint32_t>(tk % kch);
    const uint64_t t = tk / kch;
    const uint64_t row = t * 16 + (l & 15);
    const uint64_t src = row < n ? perm[row] : 0;
    const uint32_t f0 = kc * 32 + 8 * (l >> 4);
    uint16_t h[8], lo[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float c = 0.0f;
      const uint32_t f = f0 + e;
      if (row < n && f < d) c = static_cast<float>(static_cast<double>(x[src * d + f]) - mean[f]);
      h[e] = bf16_rn(c);
      lo[e] = bf16_rn(c - bf16_to_f(h[e]));
    }
    fhi[g] = make_uint4(h[0] | (static_cast<uint32_t>(h[1]) << 16), h[2] | (static_cast<uint32_t>(h[3]) << 16),
                        h[4] | (static_cast<uint32_t>(h[5]) << 16), h[6] | (static_cast<uint32_t>(h[7]) << 16));
    if (flo) flo[g] = make_uint4(lo[0] | (static_cast<uint32_t>(lo[1]) << 16), lo[2] | (static_cast<uint32_t>(lo[3]) << 16),
                        lo[4] | (static_cast<uint32_t>(lo[5]) << 16), lo[6] | (static_cast<uint32_t>(lo[7]) << 16));
  }
}

// f32 -> OCP FP8 E4M3 (bias 7, 3 mantissa bits, max 448, no infinities),
// round to nearest even, saturating at +-448.  Finite inputs only.
__device__ __forceinline__ uint32_t f32_to_e4m3(float v) {
  const uint32_t sign = (__float_as_uint(v) >> 24) & 0x80u;
  const float x = fabsf(v);
  uint32_t code;
  if (x < 0.015625f) {  // below 2^-6: subnormal steps of 2^-9 (8 -> smallest normal)
    code = static_cast<uint32_t>(rintf(x * 512.0f));
  } else {
    const uint32_t b = __float_as_uint(x);
    const uint32_t r = b + 0x7ffffu + ((b >> 20) & 1u);  // round the 20 dropped mantissa bits
    const int e = static_cast<int>(r >> 23) - 127;
    const uint32_t m = (r >> 20) & 7u;
    code = (e > 8 || (e == 8 && m == 7u)) ? 0x7eu : (static_cast<uint32_t>(e + 7) << 3) | m;
  }
  return sign | code;
}

// FP8 fragments of the centred rows scaled by 1 / rinv[row] (|values| <= 448):
// [row tile t][128-feature chunk kc][lane l][32 bytes], row = 16 t + (l & 15),
// byte e = feature 128 kc + 32 (l >> 4) + e; bytes 0..15 go to fa, 16..31 to
// fb (the two halves of a v_mfma_*_16x16x128_f8f6f4 operand).
__global__ void knn_frag8_kernel(const float *__restrict__ x, uint64_t n, uint32_t d,
                                 const double *__restrict__ mean, uint32_t kch, uint64_t ntiles,
                                 const float *__restrict__ rinv, const uint32_t *__restrict__ perm,
                                 uint4 *__restrict__ fa, uint4 *__restrict__ fb) {
  const uint64_t total = ntiles * kch * 64;
  for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < total; g += gridDim.x * 256ull) {
    const uint32_t l = static_cast<uint32_t>(g & 63);
    const uint64_t tk = g >> 6;
    const uint32_t kc = static_cast<uint32_t>(tk % kch);
    const uint64_t t = tk / kch;
    const uint64_t row = t * 16 + (l & 15);
    const uint64_t src = row < n ? perm[row] : 0;
    const uint32_t f0 = kc * 128 + 32 * (l >> 4);
    const float ri = row < n ? rinv[row] : 0.0f;
    const float sc = ri > 0.0f ? 1.0f / ri : 0.0f;
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      uint32_t word = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t f = f0 + 4 * q + e;
        float c = 0.0f;
        if (row < n && f < d) c = static_cast<float>(static_cast<double>(x[src * d + f]) - mean[f]);
        word |= f32_to_e4m3(c * sc) << (8 * e);
      }
      w[q] = word;
    }
    fa[g] = make_uint4(w[0], w[1], w[2], w[3]);
    fb[g] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}

// Position r of the candidate order holds row perm[r]: arow[r] (and
// arow_orig[perm[r]], the same value by row) = |c|^2 (1 - kappa).  L lanes
// share a row and read it in coalesced runs; the double sum of the squares
// is reduced over the L lanes before its one rounding to float.
template <int L>
__global__ __launch_bounds__(256) void knn_rownorm_kernel(const float *__restrict__ x, uint64_t n, uint32_t d,
                                                          const double *__restrict__ mean, float one_minus_kappa,
                                                          uint64_t rows_padded, const uint32_t *__restrict__ perm,
                                                          float *__restrict__ arow, float *__restrict__ arow_orig,
                                                          uint32_t *__restrict__ nonfinite, float *__restrict__ rinv) {
  const uint64_t row = blockIdx.x * static_cast<uint64_t>(256 / L) + threadIdx.x / L;
  const uint32_t j = threadIdx.x % L;
  if (row >= rows_padded) return;  // uniform over the row's L lanes
  if (row >= n) {  // padding rows: NaN makes every lower bound against them fail
    if (j == 0) {
      arow[row] = __builtin_nanf("");
      if (rinv) rinv[row] = 0.0f;
    }
    return;
  }
  const uint64_t src = perm[row];
  const float *xr = x + src * d;
  double s = 0.0;
  float mx = 0.0f;
  bool bad = false;
  for (uint32_t f = j; f < d; f += L) {
    const float xv = xr[f];
    bad = bad || !isfinite(xv);
    const float c = static_cast<float>(static_cast<double>(xv) - mean[f]);
    s += static_cast<double>(c) * c;
    mx = fmaxf(mx, fabsf(c));
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, L);
    mx = fmaxf(mx, __shfl_xor(mx, o, L));
    bad = bad || __shfl_xor(static_cast<int>(bad), o, L) != 0;
  }
  if (j != 0) return;
  const float a = static_cast<float>(s * static_cast<double>(one_minus_kappa));
  arow[row] = a;
  arow_orig[src] = a;
  if (rinv) rinv[row] = mx / 448.0f;
  if (bad) atomicOr(nonfinite, 1u);
}

// ---------------------------------------------------------------------------
// candidate order and tile-pair pruning
// ---------------------------------------------------------------------------
// The candidate kernels work on the rows in the order of the Morton code of
// their first kKnnOrderDims coordinates (each quantised to kKnnOrderBits
// bits between its minimum and maximum), so that the 16 rows of a tile lie
// close together there.  A tile's box is the range of each of its first
// kKnnBoxDims coordinates.  The squared gaps between two boxes, summed,
// bound the squared distance of every pair of their rows from below: a
// candidate tile whose bound to a query group exceeds the largest tau of
// both is skipped by the symmetric pass, as its tests would all fail.
constexpr int kKnnOrderDims = 6;
constexpr int kKnnOrderBits = 30 / kKnnOrderDims;
constexpr int kKnnBoxDims = 8;

// bit b of v to bit b kKnnOrderDims
__device__ __forceinline__ uint32_t morton_spread(uint32_t v) {
  uint32_t out = 0;
#pragma unroll
  for (int b = 0; b < kKnnOrderBits; ++b) out |= ((v >> b) & 1u) << (b * kKnnOrderDims);
  return out;
}

// mm: kKnnOrderDims minimum keys, then as many maximum keys (tkey_m order)
__global__ void knn_order_init_kernel(uint32_t *mm) {
  const int t = threadIdx.x;
  if (t < kKnnOrderDims) {
    mm[t] = 0xffffffffu;
    mm[kKnnOrderDims + t] = 0u;
  }
}

__global__ __launch_bounds__(256) void knn_order_minmax_kernel(const float *__restrict__ x, uint64_t n, uint32_t d,
                                                               uint32_t kd, uint32_t *__restrict__ mm) {
  uint32_t lo[kKnnOrderDims], hi[kKnnOrderDims];
#pragma unroll
  for (int j = 0; j < kKnnOrderDims; ++j) {
    lo[j] = 0xffffffffu;
    hi[j] = 0u;
  }
  for (uint64_t r = blockIdx.x * 256ull + threadIdx.x; r < n; r += gridDim.x * 256ull)
#pragma unroll
    for (int j = 0; j < kKnnOrderDims; ++j)
      if (static_cast<uint32_t>(j) < kd) {
        const uint32_t k = tkey_m(x[r * d + j]);
        lo[j] = min(lo[j], k);
        hi[j] = max(hi[j], k);
      }
  // the workgroup's extremes first (LDS), then one atomic per coordinate
  __shared__ uint32_t sl[4][2 * kKnnOrderDims];
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kKnnOrderDims; ++j) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[j] = min(lo[j], static_cast<uint32_t>(__shfl_xor(static_cast<int>(lo[j]), o, 64)));
      hi[j] = max(hi[j], static_cast<uint32_t>(__shfl_xor(static_cast<int>(hi[j]), o, 64)));
    }
    if ((threadIdx.x & 63) == 0) {
      sl[wave][j] = lo[j];
      sl[wave][kKnnOrderDims + j] = hi[j];
    }
  }
  __syncthreads();
  if (threadIdx.x < static_cast<uint32_t>(kKnnOrderDims) && threadIdx.x < kd) {
    const int j = threadIdx.x;
    atomicMin(&mm[j], min(min(sl[0][j], sl[1][j]), min(sl[2][j], sl[3][j])));
    atomicMax(&mm[kKnnOrderDims + j],
              max(max(sl[0][kKnnOrderDims + j], sl[1][kKnnOrderDims + j]),
                  max(sl[2][kKnnOrderDims + j], sl[3][kKnnOrderDims + j])));
  }
}

__global__ void knn_order_key_kernel(const float *__restrict__ x, uint64_t n, uint32_t d, uint32_t kd,
                                     const uint32_t *__restrict__ mm, uint32_t *__restrict__ key,
                                     uint32_t *__restrict__ idx) {
  const uint64_t r = blockIdx.x * 256ull + threadIdx.x;
  if (r >= n) return;
  uint32_t code = 0;
#pragma unroll
  for (int j = 0; j < kKnnOrderDims; ++j) {
    if (static_cast<uint32_t>(j) >= kd) break;
    const float lo = from_tkey_m(mm[j]), hi = from_tkey_m(mm[kKnnOrderDims + j]);
    const float v = x[r * d + j];
    constexpr float levels = static_cast<float>(1u << kKnnOrderBits);
    const float u = hi > lo ? (v - lo) / (hi - lo) * levels : 0.0f;
    const uint32_t q = u >= levels - 1.0f ? (1u << kKnnOrderBits) - 1u : (u > 0.0f ? static_cast<uint32_t>(u) : 0u);
    code |= morton_spread(q) << j;
  }
  key[r] = code;
  idx[r] = static_cast<uint32_t>(r);
}

// box: per tile 2 kKnnBoxDims floats, the minima then the maxima of its
// rows' first mb coordinates; one thread per (tile, coordinate)
__global__ void knn_tilebox_kernel(const float *__restrict__ x, uint64_t n, uint32_t d, uint32_t mb,
                                   const uint32_t *__restrict__ perm, uint64_t ntiles, float *__restrict__ box) {
  const uint64_t g = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t t = g / mb;
  const uint32_t j = static_cast<uint32_t>(g % mb);
  if (t >= ntiles) return;
  float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
  for (uint32_t i = 0; i < 16; ++i) {
    const uint64_t r = t * 16 + i;
    if (r < n) {
      const float v = x[static_cast<uint64_t>(perm[r]) * d + j];
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
  }
  box[t * (2 * kKnnBoxDims) + j] = lo;
  box[t * (2 * kKnnBoxDims) + kKnnBoxDims + j] = hi;
}

// After the tau sample: the largest tau of each tile, and tau by row
// (gtau_orig[perm[r]] = gtau[r]) for the rerank.
__global__ void knn_tiletau_kernel(const float *__restrict__ gtau, const uint32_t *__restrict__ perm, uint64_t n,
                                   uint64_t ntiles, float *__restrict__ tmax, float *__restrict__ gtau_orig) {
  const uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t >= ntiles) return;
  float m = -__builtin_huge_valf();
  for (uint32_t i = 0; i < 16; ++i) {
    const uint64_t r = t * 16 + i;
    if (r < n) {
      const float v = gtau[r];
      m = fmaxf(m, v);
      gtau_orig[perm[r]] = v;
    }
  }
  tmax[t] = m;
}

// Per query group (gt tiles): the union of its tiles' boxes and its
// largest tau (gbox: 2 kKnnBoxDims + 1 floats per group).
__global__ void knn_groupbox_kernel(const float *__restrict__ box, const float *__restrict__ tmax, uint64_t ntiles,
                                    uint32_t groups, uint32_t gt, uint32_t mb, float *__restrict__ gbox) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  if (g >= groups) return;
  float lo[kKnnBoxDims], hi[kKnnBoxDims];
  float m = -__builtin_huge_valf();
#pragma unroll
  for (int j = 0; j < kKnnBoxDims; ++j) {
    lo[j] = __builtin_huge_valf();
    hi[j] = -__builtin_huge_valf();
  }
  for (uint32_t i = 0; i < gt; ++i) {
    const uint64_t t = static_cast<uint64_t>(g) * gt + i;
    if (t >= ntiles) break;
#pragma unroll
    for (int j = 0; j < kKnnBoxDims; ++j)
      if (static_cast<uint32_t>(j) < mb) {
        lo[j] = fminf(lo[j], box[t * (2 * kKnnBoxDims) + j]);
        hi[j] = fmaxf(hi[j], box[t * (2 * kKnnBoxDims) + kKnnBoxDims + j]);
      }
    m = fmaxf(m, tmax[t]);
  }
  float *o = gbox + static_cast<size_t>(g) * (2 * kKnnBoxDims + 1);
#pragma unroll
  for (int j = 0; j < kKnnBoxDims; ++j) {
    o[j] = lo[j];
    o[kKnnBoxDims + j] = hi[j];
  }
  o[2 * kKnnBoxDims] = m;
}

// Workgroup (word w, 256 query groups): bit i of bits[g][w] is set when
// tile 32 w + i, a tile after group g's own, can be skipped for the group:
// the boxes' summed squared gaps exceed the larger of the group's and the
// tile's largest tau.  Each FP32 gap is a correctly rounded difference, so
// the sum is within a relative 10^-6 of the exact bound, inside the 10^-5
// margin.  The word's 32 tile boxes are staged in LDS.
__global__ __launch_bounds__(256) void knn_prune_kernel(const float *__restrict__ box, const float *__restrict__ tmax,
                                                        const float *__restrict__ gbox, uint64_t ntiles,
                                                        uint32_t groups, uint32_t gt, uint32_t mb, uint32_t words,
                                                        uint32_t *__restrict__ bits,
                                                        unsigned long long *__restrict__ count) {
  __shared__ float sb[32][2 * kKnnBoxDims + 1];
  __shared__ uint32_t scount;
  if (threadIdx.x == 0) scount = 0;
  const uint32_t w = blockIdx.x, t0 = w * 32u;
  for (uint32_t i = threadIdx.x; i < 32u * (2 * kKnnBoxDims); i += 256u) {
    const uint32_t ti = i / (2 * kKnnBoxDims), j = i % (2 * kKnnBoxDims);
    const uint64_t t = static_cast<uint64_t>(t0) + ti;
    sb[ti][j] = t < ntiles ? box[t * (2 * kKnnBoxDims) + j] : 0.0f;
  }
  if (threadIdx.x < 32u) {
    const uint64_t t = static_cast<uint64_t>(t0) + threadIdx.x;
    sb[threadIdx.x][2 * kKnnBoxDims] = t < ntiles ? tmax[t] : 0.0f;
  }
  __syncthreads();
  const uint32_t g = blockIdx.y * 256u + threadIdx.x;
  uint32_t out = 0;
  const uint64_t first = (static_cast<uint64_t>(g) + 1) * gt;  // the first tile after the group's own
  if (g < groups && static_cast<uint64_t>(t0) + 32 > first) {
    const float *gb = gbox + static_cast<size_t>(g) * (2 * kKnnBoxDims + 1);
    float glo[kKnnBoxDims], ghi[kKnnBoxDims];
#pragma unroll
    for (int j = 0; j < kKnnBoxDims; ++j) {
      glo[j] = gb[j];
      ghi[j] = gb[kKnnBoxDims + j];
    }
    const float gm = gb[2 * kKnnBoxDims];
    for (uint32_t i = 0; i < 32; ++i) {
      const uint64_t t = static_cast<uint64_t>(t0) + i;
      if (t >= ntiles) break;
      if (t < first) continue;
      float b = 0.0f;
#pragma unroll
      for (int j = 0; j < kKnnBoxDims; ++j)
        if (static_cast<uint32_t>(j) < mb) {
          const float gap = fmaxf(0.0f, fmaxf(sb[i][j] - ghi[j], glo[j] - sb[i][kKnnBoxDims + j]));
          b = fmaf(gap, gap, b);
        }
      if (b * (1.0f - 1e-5f) > fmaxf(gm, sb[i][2 * kKnnBoxDims])) out |= 1u << i;
    }
  }
  if (g < groups) bits[static_cast<size_t>(g) * words + w] = out;
  // skipped pairs, for the timings: one atomic per workgroup
  uint32_t c = static_cast<uint32_t>(__popc(out));
  for (int o = 32; o > 0; o >>= 1) c += static_cast<uint32_t>(__shfl_xor(static_cast<int>(c), o, 64));
  if ((threadIdx.x & 63) == 0) atomicAdd(&scount, c);
  __syncthreads();
  if (threadIdx.x == 0 && scount) atomicAdd(count, static_cast<unsigned long long>(scount));
}

// Item b of the symmetric pass (see knn_sym_plan): windows of L = m gt
// candidate tiles (gt = 4 kCandRT tiles per query group); window j < W - 1
// holds the (j + 1) m groups whose first tile lies before its end, the last
// window all G groups, so m j (j + 1) / 2 items precede window j.  Group g
// of window j scans tiles [max(j L, g gt), min((j + 1) L, ntiles)).
__device__ __forceinline__ void knn_sym_item(const KnnSymPlan &p, uint32_t ntiles, uint32_t b, uint32_t &g,
                                             uint32_t &lo, uint32_t &hi) {
  const uint64_t m = p.m, last = m * (p.W - 1) * p.W / 2;
  uint64_t j;
  if (b >= last) {
    j = p.W - 1;
    g = static_cast<uint32_t>(b - last);
  } else {
    j = static_cast<uint64_t>((sqrt(8.0 * static_cast<double>(b) / static_cast<double>(m) + 1.0) - 1.0) * 0.5);
    while (m * (j + 1) * (j + 2) / 2 <= b) ++j;
    while (m * j * (j + 1) / 2 > b) --j;
    g = static_cast<uint32_t>(b - m * j * (j + 1) / 2);
  }
  const uint64_t w0 = j * p.L, w1 = min(w0 + p.L, static_cast<uint64_t>(ntiles));
  const uint64_t first = static_cast<uint64_t>(g) * p.gt;
  lo = static_cast<uint32_t>(w0 > first ? w0 : first);
  hi = static_cast<uint32_t>(w1);
}

// ---------------------------------------------------------------------------
// candidate kernel
// ---------------------------------------------------------------------------
// Block = 4 waves; each wave owns RT query tiles (16 RT rows) with their A
// fragments in registers and streams candidate tiles staged in LDS (CT tiles
// per stage, shared by the 4 waves, double buffered: the next stage is
// fetched into registers while the current one is multiplied, so one barrier
// per stage remains).  Lane l holds, per query tile rt and r = 0..3, the
// value for query row 16 rt + 4 (l >> 4) + r and candidate 16 t + (l & 15).
//
// PHASE 0 scans a 1/ss sample of candidate tiles and writes tau per row.
// PHASE 1 scans the candidate tiles of split blockIdx.y (of gridDim.y) and
// appends every candidate with lb <= tau to the row's list of that split
// (MODE 1: entries are (candidate, dot) pairs).
// PHASE 2 computes each pair of tiles once: the block's item (items[3 b ..])
// names a group g of 4 RT query tiles and a range of candidate tiles from
// the upper triangle (tiles >= 4 RT g).  On the group's own tiles only the
// query rows' lists are appended (every ordered pair of the group is
// computed there); on later tiles a pair (q, c) is tested against both
// tau_q and tau_c and appended to both rows' lists.  Appends are staged per
// wave in LDS and flushed with one batch of global atomics on the rows'
// counters (gcnt[row], lists of capg entries, nsplit = 1), so a list's order
// depends on timing; the rerank sorts by (distance, index) and its pruning
// bound is order-free, so the result does not.
//
// MODE 0: one bf16 product per multiply-add; MODE 1: split bf16 (hi hi + hi lo
// + lo hi); MODE 2: FP8 E4M3 rows scaled by 1 / rinv (one 16x16x128 f8f6f4
// MFMA per 128-feature chunk), the product rescaled by rinv_q rinv_c.
// (split phase 1 with 4 chunks: held to 3 waves per SIMD, 168 VGPRs, which
// it fits without spilling)
template <int KCH, int RT, int MODE, int PHASE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 1 && KCH == 4 ? 3 : 1)))
void knn_mfma_kernel(
    const uint4 *__restrict__ fhi, const uint4 *__restrict__ flo, const float *__restrict__ arow,
    const float *__restrict__ rinv, uint32_t n, uint32_t ntiles, uint32_t ss, uint32_t capg,
    uint32_t *__restrict__ gcnt, uint32_t *__restrict__ gcand, float *__restrict__ gtau,
    KnnSymPlan sp, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ prune, uint32_t pwords,
    KnnSampleSet smp) {
  static_assert(PHASE != 2 || RT <= 2, "phase 2 holds a lane's tests of two tiles in 32 bits");
  constexpr bool SPLIT = MODE == 1;
  constexpr bool FP8 = MODE == 2;
  constexpr int CT = (8 / KCH) > 0 ? (8 / KCH) : 1;  // candidate tiles per LDS stage
  constexpr int NH = MODE != 0 ? 2 : 1;
  constexpr int PER = CT * KCH * NH * 64;            // uint4 per stage
  constexpr int PF = (PER + 255) / 256;              // prefetch registers per thread
  __shared__ uint4 sB[2][CT][KCH][NH][64];
  __shared__ float sA[2][CT][16];
  __shared__ float sI[2][CT][16];  // FP8: candidate rinv
  __shared__ uint32_t sCnt[4][16 * RT];
  // rows by position (perm[position], ~0 past n): the wave's query rows and
  // each stage's candidates
  __shared__ uint32_t sQo[4][16 * RT];
  __shared__ uint32_t sP[2][CT][16];
  // PHASE 2: (a_c - tau_c) / 2 of the stage's candidates, and each wave's
  // staged appends (row, candidate, dot)
  constexpr uint32_t kApp = 128;
  __shared__ float sT[PHASE == 2 ? 2 : 1][CT][16];
  __shared__ uint32_t sApp[PHASE == 2 ? 4 : 1][3][PHASE == 2 ? kApp : 1];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t group = blockIdx.x, item_lo = 0, item_hi = 0;
  if (PHASE == 2) knn_sym_item(sp, ntiles, blockIdx.x, group, item_lo, item_hi);
  const uint32_t gtile0 = group * 4 * RT;  // the group's first query tile
  const uint32_t qtile0 = (group * 4 + wave) * RT;
  const int rsub = lane >> 4, col = lane & 15;
  const uint32_t split = PHASE == 1 ? blockIdx.y : 0, nsplit = PHASE == 1 ? gridDim.y : 1;

  bf16x8 ahi[RT][KCH], alo[RT][KCH];
  float aq[RT][4], iq[RT][4];
  uint32_t qidx[RT][4];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    const uint32_t tile = qtile0 + rt;
#pragma unroll
    for (int kc = 0; kc < KCH; ++kc) {
      uint4 h = make_uint4(0, 0, 0, 0), lo = make_uint4(0, 0, 0, 0);
      if (tile < ntiles) {
        h = fhi[(static_cast<size_t>(tile) * KCH + kc) * 64 + lane];
        if (NH == 2) lo = flo[(static_cast<size_t>(tile) * KCH + kc) * 64 + lane];
      }
      ahi[rt][kc] = __builtin_bit_cast(bf16x8, h);
      alo[rt][kc] = __builtin_bit_cast(bf16x8, lo);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t q = tile * 16 + 4 * rsub + r;
      qidx[rt][r] = q;
      aq[rt][r] = q < n ? arow[q] : __builtin_nanf("");
      iq[rt][r] = (FP8 && q < n) ? rinv[q] : 1.0f;
    }
  }
  for (int i = threadIdx.x; i < 4 * 16 * RT; i += 256) (&sCnt[0][0])[i] = 0;
  if (lane < 16 * RT) {
    const uint32_t q = qtile0 * 16 + static_cast<uint32_t>(lane);
    sQo[wave][lane] = q < n ? perm[q] : 0xffffffffu;
  }

  // Phase 1 tests lb = aq + ac - 2 dot <= tau as dot >= hq + ac / 2 with
  // hq = (aq - tau) / 2 per row (held in tau[]): two operations per pair.
  // Its roundings (a few 2^-24 of aq + ac) are far inside kappa's margin, so
  // an unlisted pair still has a squared distance above tau.
  float best0[RT][4], best1[RT][4], tau[RT][4];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      best0[rt][r] = __builtin_huge_valf();
      best1[rt][r] = __builtin_huge_valf();
      const uint32_t q = qidx[rt][r];
      tau[rt][r] = (PHASE != 0 && q < n) ? (aq[rt][r] - gtau[q]) * 0.5f : __builtin_huge_valf();
    }

  // candidate slots: sample tiles (phase 0), this split's range (phase 1)
  // or the item's range (phase 2)
  // PHASE 0 scans the sample set's tiles (rows drawn across the whole
  // input, so that a sample tile's rows are not neighbours), every other
  // phase the query order's own tiles
  const uint4 *__restrict__ cfhi = PHASE == 0 ? smp.fhi : fhi;
  const uint4 *__restrict__ cflo = PHASE == 0 ? smp.flo : flo;
  const float *__restrict__ carow = PHASE == 0 ? smp.arow : arow;
  const float *__restrict__ crinv = PHASE == 0 ? smp.rinv : rinv;
  const uint32_t ctiles = PHASE == 0 ? smp.tiles : ntiles;
  const uint32_t nsample = ctiles;
  const uint32_t slot_lo = PHASE == 0   ? 0
                           : PHASE == 2 ? item_lo
                                        : static_cast<uint32_t>(static_cast<uint64_t>(ntiles) * split / nsplit);
  const uint32_t slot_hi = PHASE == 0   ? nsample
                           : PHASE == 2 ? item_hi
                                        : static_cast<uint32_t>(static_cast<uint64_t>(ntiles) * (split + 1) / nsplit);
  auto slot_tile = [&](uint32_t slot) -> uint32_t { return slot; };
  const uint32_t nstage = (slot_hi - slot_lo + CT - 1) / CT;

  uint4 pf[PF];
  float pa = 0.0f, pi = 0.0f, pt = 0.0f;
  uint32_t pp = 0xffffffffu;
  // phase 2: the stage's candidate tiles can all be skipped for this group
  // (see knn_prune_kernel; a uniform condition)
  auto stage_pruned = [&](uint32_t st) -> bool {
    if (PHASE != 2 || pwords == 0) return false;
    const uint32_t t0 = slot_lo + st * CT;
    const uint32_t nv = min(static_cast<uint32_t>(CT), slot_hi - t0);
    const uint32_t word = prune[static_cast<size_t>(group) * pwords + (t0 >> 5)];
    const uint32_t mask = ((nv >= 32u ? 0u : (1u << nv)) - 1u) << (t0 & 31u);
    return (word & mask) == mask;
  };
  // When a stage's half (hi or lo fragments) is a whole number of 256-thread
  // rounds, the stage is fetched through buffer resources: slots past the
  // range point past the fragment arrays and read zeros with no branch, and
  // each round reads one half.  Otherwise one guarded load per element.
  constexpr int PERH = CT * KCH * 64;
  constexpr bool ROUNDS = PERH % 256 == 0;
  constexpr int PFH = ROUNDS ? PERH / 256 : 1;
  const uint32_t frag_bytes = ctiles * static_cast<uint32_t>(KCH) * 64u * 16u;
  const __amdgpu_buffer_rsrc_t rsh =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4 *>(cfhi), 0, static_cast<int>(frag_bytes), 0x00020000);
  const __amdgpu_buffer_rsrc_t rsl = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint4 *>(NH == 2 ? cflo : cfhi), 0, static_cast<int>(frag_bytes), 0x00020000);
  // stage st into LDS buffer b (ROUNDS: straight to LDS, a wave's 64 lanes
  // filling one 1 KB fragment block; otherwise into registers for stash)
  auto fetch = [&](uint32_t st, int b) {
    const uint32_t s0 = slot_lo + st * CT;
    if (stage_pruned(st)) return;
    if constexpr (ROUNDS) {
#pragma unroll
      for (int h = 0; h < NH; ++h)
#pragma unroll
        for (int j = 0; j < PFH; ++j) {
          const int i = static_cast<int>(threadIdx.x) + j * 256;
          const int l = i & 63, kc = (i >> 6) % KCH, ct = i / (64 * KCH);
          const uint32_t slot = s0 + ct;
          const uint32_t tile = slot_tile(slot);
          const uint32_t o = (slot < slot_hi && tile < ctiles)
                                 ? ((tile * static_cast<uint32_t>(KCH) + kc) * 64u + l) * 16u
                                 : 0xfffffff0u;
          // reads past the fragments (slots past the range) write zeros
          __builtin_amdgcn_raw_ptr_buffer_load_lds(h == 0 ? rsh : rsl,
                                                   (__attribute__((address_space(3))) void *)&sB[b][ct][kc][h][0], 16,
                                                   static_cast<int>(o), 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        const int i = static_cast<int>(threadIdx.x) + j * 256;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (i < PER) {
          const int l = i & 63, h = (i >> 6) % NH, kc = (i / (64 * NH)) % KCH, ct = i / (64 * NH * KCH);
          const uint32_t slot = s0 + ct;
          const uint32_t tile = slot_tile(slot);
          if (slot < slot_hi && tile < ctiles) {
            const size_t o = (static_cast<size_t>(tile) * KCH + kc) * 64 + l;
            v = h == 0 ? cfhi[o] : cflo[o];
          }
        }
        pf[j] = v;
      }
    }
    if (threadIdx.x < CT * 16) {
      const uint32_t slot = s0 + threadIdx.x / 16;
      const uint32_t tile = slot_tile(slot);
      pa = (slot < slot_hi && tile < ctiles) ? carow[tile * 16 + (threadIdx.x & 15)] : 0.0f;
      if (FP8) pi = (slot < slot_hi && tile < ctiles) ? crinv[tile * 16 + (threadIdx.x & 15)] : 0.0f;
      {
        const uint32_t c = tile * 16 + (threadIdx.x & 15);
        pp = (PHASE != 0 && slot < slot_hi && tile < ntiles && c < n) ? perm[c] : 0xffffffffu;
      }
      if (PHASE == 2) {
        // rows past n carry a = NaN, so their half-threshold is NaN too
        const uint32_t c = tile * 16 + (threadIdx.x & 15);
        pt = (slot < slot_hi && tile < ntiles) ? (pa - (c < n ? gtau[c] : 0.0f)) * 0.5f : __builtin_nanf("");
      }
    }
  };
  auto stash = [&](int buf) {
    if constexpr (ROUNDS) {
      // this wave's fragment loads into LDS are complete before the barrier
      // that publishes the stage
      __builtin_amdgcn_s_waitcnt(0);
    } else {
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        const int i = static_cast<int>(threadIdx.x) + j * 256;
        if (i < PER) (&sB[buf][0][0][0][0])[i] = pf[j];
      }
    }
    if (threadIdx.x < CT * 16) {
      (&sA[buf][0][0])[threadIdx.x] = pa;
      (&sP[buf][0][0])[threadIdx.x] = pp;
      if (FP8) (&sI[buf][0][0])[threadIdx.x] = pi;
      if (PHASE == 2) (&sT[buf & (PHASE == 2 ? 1 : 0)][0][0])[threadIdx.x] = pt;
    }
  };

  // PHASE 2 appends: this wave's staged count (uniform) and the flush of
  // its staged entries, every atomic of the batch issued before any slot
  // is used
  uint32_t napp = 0;
  // the flush: every slot of a round is reserved by an atomic on its row's
  // counter before any is written; slots past capg are dropped (the row's
  // count then exceeds capg and the rerank rescans it)
  auto flush = [&]() __attribute__((always_inline)) {
    if constexpr (PHASE == 2) {
#pragma unroll 1
      for (uint32_t e = static_cast<uint32_t>(lane); e < napp; e += 64) {
        const uint32_t row = sApp[wave][0][e], cnd = sApp[wave][1][e], dv = sApp[wave][2][e];
        const uint32_t pos = atomicAdd(&gcnt[row], 1u);
        if (pos < capg) {
          const size_t slot = static_cast<size_t>(row) * capg + pos;
          if (SPLIT)
            reinterpret_cast<uint2 *>(gcand)[slot] = make_uint2(cnd, dv);
          else
            gcand[slot] = cnd;
        }
      }
      napp = 0;
    }
  };
  if (nstage) {
    fetch(0, 0);
    stash(0);
  }
  __syncthreads();
  for (uint32_t st = 0; st < nstage; ++st) {
    const int buf = st & 1;
    if (st + 1 < nstage) fetch(st + 1, buf ^ 1);  // in flight while this stage computes
    // products of candidate tile ct into acc (3 bf16 MFMAs per chunk when
    // split, one FP8 or bf16 MFMA otherwise)
    auto products = [&](int ct, floatx4 (&acc)[RT]) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) acc[rt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < KCH; ++kc) {
        if (FP8) {
          const uint4 b0 = sB[buf][ct][kc][0][lane], b1 = sB[buf][ct][kc][NH - 1][lane];
          const intx8 bv = {static_cast<int>(b0.x), static_cast<int>(b0.y), static_cast<int>(b0.z),
                            static_cast<int>(b0.w), static_cast<int>(b1.x), static_cast<int>(b1.y),
                            static_cast<int>(b1.z), static_cast<int>(b1.w)};
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            const uint4 a0 = __builtin_bit_cast(uint4, ahi[rt][kc]), a1 = __builtin_bit_cast(uint4, alo[rt][kc]);
            const intx8 av = {static_cast<int>(a0.x), static_cast<int>(a0.y), static_cast<int>(a0.z),
                              static_cast<int>(a0.w), static_cast<int>(a1.x), static_cast<int>(a1.y),
                              static_cast<int>(a1.z), static_cast<int>(a1.w)};
            // format 0 (FP8 E4M3) for both operands, unit block scales (E8M0 127)
            acc[rt] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av, bv, acc[rt], 0, 0, 0, 127, 0, 127);
          }
          continue;
        }
        const bf16x8 bh = __builtin_bit_cast(bf16x8, sB[buf][ct][kc][0][lane]);
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rt][kc], bh, acc[rt], 0, 0, 0);
        if (SPLIT) {
          const bf16x8 bl = __builtin_bit_cast(bf16x8, sB[buf][ct][kc][NH - 1][lane]);
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rt][kc], bl, acc[rt], 0, 0, 0);
            acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[rt][kc], bh, acc[rt], 0, 0, 0);
          }
        }
      }
    };
    // both candidate tiles of a pair, chunk by chunk, so that 2 RT
    // independent accumulation chains interleave (each chain adds its
    // products in the same order as products())
    auto products2 = [&](int ct, floatx4 (&acc0)[RT], floatx4 (&acc1)[RT]) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        acc0[rt] = floatx4{0.f, 0.f, 0.f, 0.f};
        acc1[rt] = floatx4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int kc = 0; kc < KCH; ++kc) {
        if (FP8) {
          intx8 bv[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const uint4 b0 = sB[buf][ct + u][kc][0][lane], b1 = sB[buf][ct + u][kc][NH - 1][lane];
            bv[u] = intx8{static_cast<int>(b0.x), static_cast<int>(b0.y), static_cast<int>(b0.z),
                          static_cast<int>(b0.w), static_cast<int>(b1.x), static_cast<int>(b1.y),
                          static_cast<int>(b1.z), static_cast<int>(b1.w)};
          }
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            const uint4 a0 = __builtin_bit_cast(uint4, ahi[rt][kc]), a1 = __builtin_bit_cast(uint4, alo[rt][kc]);
            const intx8 av = {static_cast<int>(a0.x), static_cast<int>(a0.y), static_cast<int>(a0.z),
                              static_cast<int>(a0.w), static_cast<int>(a1.x), static_cast<int>(a1.y),
                              static_cast<int>(a1.z), static_cast<int>(a1.w)};
            acc0[rt] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av, bv[0], acc0[rt], 0, 0, 0, 127, 0, 127);
            acc1[rt] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av, bv[1], acc1[rt], 0, 0, 0, 127, 0, 127);
          }
          continue;
        }
        const bf16x8 bh0 = __builtin_bit_cast(bf16x8, sB[buf][ct][kc][0][lane]);
        const bf16x8 bh1 = __builtin_bit_cast(bf16x8, sB[buf][ct + 1][kc][0][lane]);
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          acc0[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rt][kc], bh0, acc0[rt], 0, 0, 0);
          acc1[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rt][kc], bh1, acc1[rt], 0, 0, 0);
        }
        if (SPLIT) {
          const bf16x8 bl0 = __builtin_bit_cast(bf16x8, sB[buf][ct][kc][NH - 1][lane]);
          const bf16x8 bl1 = __builtin_bit_cast(bf16x8, sB[buf][ct + 1][kc][NH - 1][lane]);
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            acc0[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rt][kc], bl0, acc0[rt], 0, 0, 0);
            acc1[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rt][kc], bl1, acc1[rt], 0, 0, 0);
          }
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            acc0[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[rt][kc], bh0, acc0[rt], 0, 0, 0);
            acc1[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[rt][kc], bh1, acc1[rt], 0, 0, 0);
          }
        }
      }
    };
    auto epilogue = [&](int ct, uint32_t tile, const floatx4 (&acc)[RT]) {
      const uint32_t c = tile * 16 + col;
      const float ac = sA[buf][ct][col];
      const float ic = FP8 ? sI[buf][ct][col] : 1.0f;
      // padded rows carry a = NaN (see knn_rownorm_kernel and aq), so their
      // tests fail every comparison: only self needs a test, and only on
      // the wave's own query tiles (a uniform condition)
      const float gc = ac * 0.5f;
      const bool diag = tile - qtile0 < static_cast<uint32_t>(RT);
      if constexpr (PHASE == 0) {
        // every lb of the tile first; one wave-wide test then ends most
        // epilogues once the two smallest have settled.  A sampled query
        // row meets itself once (lb about 0): its tau is then the 15th
        // smallest of the others, still a valid threshold
        float lbv[RT][4];
        bool any = false;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float dot = FP8 ? acc[rt][r] * (iq[rt][r] * ic) : acc[rt][r];
            lbv[rt][r] = fmaf(-2.0f, dot, aq[rt][r] + ac);
          }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int r = 0; r < 4; ++r) any |= lbv[rt][r] < best1[rt][r];
        if (__ballot(any) == 0) return;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float lb = lbv[rt][r];
            if (lb < best1[rt][r]) {
              const bool first = lb < best0[rt][r];
              best1[rt][r] = first ? best0[rt][r] : lb;
              best0[rt][r] = first ? lb : best0[rt][r];
            }
          }
        return;
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t q = qidx[rt][r];
          const float dot = FP8 ? acc[rt][r] * (iq[rt][r] * ic) : acc[rt][r];
          if (dot >= tau[rt][r] + gc && (!diag || c != q)) {
            const uint32_t pos = atomicAdd(&sCnt[wave][rt * 16 + 4 * rsub + r], 1u);
            if (pos < capg) {
              const uint32_t qo = sQo[wave][rt * 16 + 4 * rsub + r], co = sP[buf][ct][col];
              const size_t slot = (static_cast<size_t>(qo) * nsplit + split) * capg + pos;
              // split lists carry the pair's dot, for the rerank's upper bounds
              if (SPLIT)
                reinterpret_cast<uint2 *>(gcand)[slot] = make_uint2(co, __float_as_uint(dot));
              else
                gcand[slot] = co;
            }
          }
        }
    };
    // PHASE 2: the tests of candidate tiles tile0 and (two) tile0 + 1, then
    // one listing pass over both.  On the group's own tiles only the query
    // rows' lists are tested (a uniform condition).  Lane bit
    // u 8 RT + rt 4 + r lists (q, c) in q's list and bit u 8 RT + 4 RT + rt 4 + r
    // lists (c, q) in c's list, for c = 16 (tile0 + u) + col.
    auto list2 = [&](int ct, uint32_t tile0, const floatx4 (&acc0)[RT], const floatx4 (&acc1)[RT], bool two) {
      float dv[2][RT * 4];
      uint32_t bits = 0;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (u == 1 && !two) break;
        const uint32_t tile = tile0 + u;
        const float ac = sA[buf][ct + u][col];
        const float ic = FP8 ? sI[buf][ct + u][col] : 1.0f;
        const float gc = ac * 0.5f;
        const floatx4 (&acc)[RT] = u == 0 ? acc0 : acc1;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int r = 0; r < 4; ++r) dv[u][rt * 4 + r] = FP8 ? acc[rt][r] * (iq[rt][r] * ic) : acc[rt][r];
        uint32_t b = 0;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int r = 0; r < 4; ++r) b |= dv[u][rt * 4 + r] >= tau[rt][r] + gc ? 1u << (rt * 4 + r) : 0u;
        // self pairs: only on the wave's own query tiles (uniform), at
        // most one per lane
        if (tile - qtile0 < static_cast<uint32_t>(RT)) {
          const int rr = col - 4 * rsub;
          if (rr >= 0 && rr < 4) b &= ~(1u << ((tile - qtile0) * 4 + rr));
        }
        if (tile - gtile0 >= static_cast<uint32_t>(4 * RT)) {
          // tc + aq / 2: the halving is exact, so the fused form rounds
          // once, as the sum does
          const float tc = sT[buf & (PHASE == 2 ? 1 : 0)][ct + u][col];
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              b |= dv[u][rt * 4 + r] >= fmaf(aq[rt][r], 0.5f, tc) ? 1u << (4 * RT + rt * 4 + r) : 0u;
        }
        bits |= b << (u * 8 * RT);
      }
      uint64_t m = __ballot(bits != 0);
      // rounds: each lane with bits left stages its lowest one (most tile
      // pairs end here or after one round)
      while (m) {
        if (napp > kApp - 64) flush();
        if (bits) {
          const uint32_t i = __builtin_ctz(bits);
          bits &= bits - 1u;
          const uint32_t u = i / (8u * RT), k = i & (4u * RT - 1u);
          const uint32_t di = u * 4u * RT + k;  // index into dv, by a tree of selects
          float lv[RT * 8];
#pragma unroll
          for (int j = 0; j < RT * 8; ++j) lv[j] = dv[j / (RT * 4)][j % (RT * 4)];
#pragma unroll
          for (int w = RT * 4; w >= 1; w >>= 1)
#pragma unroll
            for (int j = 0; j < w; ++j) lv[j] = (di & static_cast<uint32_t>(w)) ? lv[j + w] : lv[j];
          const float dot = lv[0];
          const uint32_t c = sP[buf][ct + u][col];
          const uint32_t q = sQo[wave][(k >> 2) * 16u + 4u * rsub + (k & 3u)];
          const bool rev = (i & (4u * RT)) != 0;
          const uint32_t e = napp + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
          sApp[wave][0][e] = rev ? c : q;
          sApp[wave][1][e] = rev ? q : c;
          sApp[wave][2][e] = __float_as_uint(dot);
        }
        napp += static_cast<uint32_t>(__popcll(m));
        m = __ballot(bits != 0);
      }
    };
    // two candidate tiles at a time: their independent MFMA chains interleave
    // (slots past the range hold zero fragments; their epilogue is skipped)
    const bool skip = stage_pruned(st);
#pragma unroll 1
    for (int ct = 0; ct < (skip ? 0 : CT); ct += 2) {
      const uint32_t slot0 = slot_lo + st * CT + ct;
      const uint32_t tile0 = slot_tile(slot0);
      if (slot0 >= slot_hi || tile0 >= ctiles) break;
      floatx4 acc0[RT];
      if (ct + 1 < CT) {
        const uint32_t slot1 = slot0 + 1;
        const uint32_t tile1 = slot_tile(slot1);
        floatx4 acc1[RT];
        products2(ct, acc0, acc1);
        if constexpr (PHASE == 2) {
          list2(ct, tile0, acc0, acc1, slot1 < slot_hi && tile1 < ctiles);
        } else {
          epilogue(ct, tile0, acc0);
          if (slot1 < slot_hi && tile1 < ctiles) epilogue(ct + 1, tile1, acc1);
        }
      } else {
        products(ct, acc0);
        if constexpr (PHASE == 2)
          list2(ct, tile0, acc0, acc0, false);
        else
          epilogue(ct, tile0, acc0);
      }
    }
    if (st + 1 < nstage) stash(buf ^ 1);
    __syncthreads();
  }

  if (PHASE == 2) {
    flush();
    return;
  }
  if (PHASE == 0) {
    // tau = 16th smallest of the 32 per-row values held by the row's 16 lanes
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v0 = best0[rt][r], v1 = best1[rt][r];
        seg32_sort(v0, v1, lane);
        tau[rt][r] = __shfl(v0, (lane & ~15) | 15, 64);
      }
  }
  if (col == 0) {
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t q = qidx[rt][r];
        if (q >= n) continue;
        if (PHASE == 0)
          gtau[q] = tau[rt][r];
        else
          gcnt[static_cast<size_t>(sQo[wave][rt * 16 + 4 * rsub + r]) * nsplit + split] =
              sCnt[wave][rt * 16 + 4 * rsub + r];
      }
  }
}

// ---------------------------------------------------------------------------
// exact rerank + certificate (one wavefront per query row)
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void knn_rerank_kernel(
    const float *__restrict__ xp, uint32_t Dp, uint32_t Cp, uint32_t tail, uint32_t n, uint32_t k, uint32_t capg,
    uint32_t nsplit, const uint32_t *__restrict__ gcnt, const uint32_t *__restrict__ gcand,
    const float *__restrict__ gtau, double cert_scale, const float *__restrict__ arow, int dots,
    double ub_s1, double ub_s2, uint32_t *__restrict__ oidx, float *__restrict__ odist,
    uint32_t *__restrict__ fb_count, uint32_t *__restrict__ fb_list, unsigned long long *__restrict__ stats) {
  // per wave: the query row (Dp floats), 64 R keys, then 64 R candidate ids
  extern __shared__ __align__(16) float sQrow[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t q = blockIdx.x * 4 + wave;
  float *qv = sQrow + static_cast<size_t>(wave) * (Dp + 192 * R);
  uint64_t *keys = reinterpret_cast<uint64_t *>(qv + Dp);
  uint32_t *list = reinterpret_cast<uint32_t *>(qv + Dp + 128 * R);
  if (q < n)
    for (uint32_t f = lane; f < Dp; f += 64) qv[f] = xp[static_cast<size_t>(q) * Dp + f];
  __syncthreads();
  if (q >= n) return;
  // the row's list is the concatenation of its per-split lists: split sizes
  // in lanes 0..nsplit-1, their exclusive prefix by a wave scan
  const uint32_t myc = static_cast<uint32_t>(lane) < nsplit ? gcnt[static_cast<size_t>(q) * nsplit + lane] : 0u;
  uint32_t incl = myc;
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const uint32_t cnt = __shfl(incl, 15, 64);
  const bool overflow = __any(myc > capg);
  const bool all_listed = cnt == n - 1;
  if (overflow || cnt > capg || cnt < k) {
    if (lane == 0) fb_list[atomicAdd(fb_count, 1u)] = q;
    return;
  }
  // split lists (dots != 0) carry each pair's dot: it goes to the key area
  float *dotl = reinterpret_cast<float *>(keys);
  {
    // gather the split lists into LDS: all loads of a 64-entry round first
    uint32_t maxc = myc;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) maxc = max(maxc, __shfl_xor(maxc, o, 64));
    maxc = __shfl(maxc, 0, 64);  // lanes 0..15 hold the split sizes: the wave's bound
    const uint2 *gpair = reinterpret_cast<const uint2 *>(gcand);
    // four loads in flight per lane: four 64-entry rounds of one list, or
    // one round of four split lists
    for (uint32_t j0 = 0; j0 < maxc; j0 += (nsplit == 1 ? 256u : 64u)) {
      for (uint32_t sp0 = 0; sp0 < nsplit; sp0 += 4) {
        uint32_t v[4], dv[4], dst[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t sp = nsplit == 1 ? 0u : sp0 + u;
          const uint32_t j = j0 + lane + (nsplit == 1 ? 64u * u : 0u);
          const uint32_t csp = __shfl(myc, static_cast<int>(sp & 15u), 64);
          const uint32_t start = __shfl(incl, static_cast<int>(sp & 15u), 64) - csp;
          const size_t slot = (static_cast<size_t>(q) * nsplit + sp) * capg + j;
          v[u] = 0u;
          dv[u] = 0u;
          dst[u] = ~0u;
          if (sp < nsplit && j < csp) {
            dst[u] = start + j;
            if (dots) {
              const uint2 pr = gpair[slot];
              v[u] = pr.x;
              dv[u] = pr.y;
            } else {
              v[u] = gcand[slot];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (dst[u] != ~0u) {
            list[dst[u]] = v[u];
            if (dots) dotl[dst[u]] = __uint_as_float(dv[u]);
          }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  // Pruning by upper bounds (certified split lists: dots != 0, ub_s1 > 0).
  // With a = |c|^2 (1 - kappa) and the split-bf16 dot, every listed pair has
  //   lb = fma(-2, dot, aq + ac) (the candidate kernel's value), and
  //   d_ref^2 >= lb cert_scale (the certificate's own bound), and
  //   d_ref^2 <= U = ((aq + ac) ub_s1 - 2 dot) ub_s2,
  // ub_s1 = (1 + kappa) / ((1 - kappa)(1 - 2^-23)) bounding the product and
  // norm errors as kappa does for lb, ub_s2 the reference's summation and
  // sqrt rounding.  If theta is the k-th smallest U, at least k listed
  // candidates have d_ref^2 <= theta, so a candidate with lb cert_scale >
  // theta is strictly farther than the k-th and needs no exact distance.
  uint32_t cnt2 = cnt;
  if (dots && ub_s1 > 0.0) {
    const float aq = arow[q];
    // theta: the k-th smallest of the rounded-up U values, found by
    // bisection over their order-preserving 32-bit keys (32 rounds of one
    // ballot count per register; no sort)
    auto theta_of = [&](auto rr) -> double {
      constexpr int R2 = decltype(rr)::value;
      uint32_t v[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const uint32_t e = r * 64 + lane;
        v[r] = ~0u;
        if (e < cnt) {
          const float dot = dotl[e], ac = arow[list[e]];
          const float lb = fmaf(-2.0f, dot, aq + ac);
          const double u =
              fmax((static_cast<double>(aq) + static_cast<double>(ac)) * ub_s1 - 2.0 * static_cast<double>(dot), 0.0) *
              ub_s2;
          dotl[e] = lb;  // the dot is not needed again
          v[r] = tkey_m(__double2float_ru(u));
        }
      }
      // smallest t with #{v <= t} >= k; k <= cnt keys are finite, so t < ~0
      uint32_t lo = 0u, hi = ~0u;
      while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < R2; ++r) c += static_cast<uint32_t>(__popcll(__ballot(v[r] <= mid)));
        if (c >= k)
          hi = mid;
        else
          lo = mid + 1u;
      }
      return static_cast<double>(from_tkey_m(lo));
    };
    double theta;
    if (R >= 2 && cnt <= 128)
      theta = theta_of(std::integral_constant<int, (R >= 2 ? 2 : R)>{});
    else if (R >= 4 && cnt <= 256)
      theta = theta_of(std::integral_constant<int, (R >= 4 ? 4 : R)>{});
    else
      theta = theta_of(std::integral_constant<int, R>{});
    // keep the candidates with lb cert_scale <= theta, compacted in place
    // (a chunk is read before any of it is overwritten)
    uint32_t total = 0;
    for (uint32_t e0 = 0; e0 < cnt; e0 += 64) {
      const uint32_t e = e0 + lane;
      uint32_t c = 0;
      bool keep = false;
      if (e < cnt) {
        c = list[e];
        keep = static_cast<double>(dotl[e]) * cert_scale <= theta;
      }
      const uint64_t m = __ballot(keep);
      if (keep) list[total + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)))] = c;
      total += static_cast<uint32_t>(__popcll(m));
    }
    cnt2 = total;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (cnt2 < k) {  // cannot happen with sound bounds: rescan rather than guess
      if (lane == 0) fb_list[atomicAdd(fb_count, 1u)] = q;
      return;
    }
  }
  if (lane == 0) {  // listed and exactly reranked candidates, for the timings (256 slots)
    atomicAdd(&stats[2 * (q & 255u)], static_cast<unsigned long long>(cnt));
    atomicAdd(&stats[2 * (q & 255u) + 1], static_cast<unsigned long long>(cnt2));
  }
  // Eight lanes per candidate: lane l sums segment l (features 8c + l) in
  // order of c, the eight partials are added in lane order, then the tail
  // (d mod 8 features; the row's padding beyond it is zero in both rows and
  // adds nothing), as distance.rs:50-75 does; a group reads one packed row
  // contiguously.  A candidate's loads are issued together, up to four
  // 16-byte segment loads and its tail feature (lane l < tail holds feature
  // 8 Cp + l), before any is used.
  const int grp = lane >> 3, l = lane & 7, g0 = lane & ~7;
  const float *qseg = qv + static_cast<size_t>(l) * Cp;
  const uint32_t segEnd = 8u * Cp, nf4 = Cp / 4;
  const bool tail_lane = static_cast<uint32_t>(l) < tail;
  const float qtail = tail_lane ? qv[segEnd + l] : 0.0f;
  // two candidates per group at a time (e and e + 8): twice the gathers in flight
  for (uint32_t e = grp; e < cnt2; e += 16) {
    const bool two = e + 8 < cnt2;  // the same in the group's eight lanes
    uint32_t c[2];
    c[0] = list[e];
    c[1] = two ? list[e + 8] : c[0];
    const float4 *cseg[2];
    float ctail[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float *cr = xp + static_cast<size_t>(c[h]) * Dp;
      cseg[h] = reinterpret_cast<const float4 *>(cr + static_cast<size_t>(l) * Cp);
      ctail[h] = tail_lane ? cr[segEnd + l] : 0.0f;
    }
    const float4 *qs4 = reinterpret_cast<const float4 *>(qseg);
    float acc[2] = {0.0f, 0.0f};
    auto add4 = [&](int h, const float4 &cv, const float4 &qq) __attribute__((always_inline)) {
      const float cvs[4] = {cv.x, cv.y, cv.z, cv.w}, qvs[4] = {qq.x, qq.y, qq.z, qq.w};
#pragma unroll
      for (int e2 = 0; e2 < 4; ++e2) {
        const float df = qvs[e2] - cvs[e2];
        const float sq = df * df;
        acc[h] = acc[h] + sq;
      }
    };
    for (uint32_t f4 = 0; f4 < nf4; f4 += 4) {
      float4 cv[2][4];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          cv[h][u] = f4 + u < nf4 ? cseg[h][f4 + u] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (f4 + u < nf4) {
          const float4 qq = qs4[f4 + u];
          add4(0, cv[0][u], qq);
          add4(1, cv[1][u], qq);
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float sd = 0.0f;
#pragma unroll
      for (int w = 0; w < 8; ++w) sd = sd + __shfl(acc[h], g0 | w, 64);
      const float dft = qtail - ctail[h];
      const float sqt = dft * dft;
      for (uint32_t f = 0; f < tail; ++f) sd = sd + __shfl(sqt, g0 | static_cast<int>(f), 64);
      if (l == 0 && (h == 0 || two)) keys[e + 8 * h] = (static_cast<uint64_t>(tkey_m(sqrtf(sd))) << 32) | c[h];
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  // up to 128 keys: each key's rank is the number of smaller keys (keys are
  // distinct: a candidate is listed once per row), counted against the
  // keys read back from LDS; the first k go to their rank's slot
  if (cnt2 <= 128) {
    uint64_t v[2];
    uint32_t rank[2] = {0u, 0u};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint32_t e = r * 64 + lane;
      v[r] = e < cnt2 ? keys[e] : ~0ull;
    }
    uint32_t j = 0;
    for (; j + 1 < cnt2; j += 2) {
      const uint64_t a = keys[j], b = keys[j + 1];
#pragma unroll
      for (int r = 0; r < 2; ++r) rank[r] += (a < v[r] ? 1u : 0u) + (b < v[r] ? 1u : 0u);
    }
    if (j < cnt2) {
      const uint64_t a = keys[j];
#pragma unroll
      for (int r = 0; r < 2; ++r) rank[r] += a < v[r] ? 1u : 0u;
    }
    // the k-th smallest key is held where the rank is k - 1
    uint64_t kth = 0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint64_t m = __ballot(r * 64 + lane < static_cast<int>(cnt2) && rank[r] == k - 1);
      if (m) kth = shfl_u64m(v[r], static_cast<int>(__builtin_ctzll(m)));
    }
    const float dk = from_tkey_m(static_cast<uint32_t>(kth >> 32));
    const float t = gtau[q];
    const bool cert = all_listed || cert_scale < 0.0 ||
                      (t > 0.0f && sqrt(static_cast<double>(t) * cert_scale) > static_cast<double>(dk));
    if (!cert) {
      if (lane == 0) fb_list[atomicAdd(fb_count, 1u)] = q;
      return;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (r * 64 + lane < static_cast<int>(cnt2) && rank[r] < k) {
        oidx[static_cast<size_t>(q) * k + rank[r]] = static_cast<uint32_t>(v[r]);
        odist[static_cast<size_t>(q) * k + rank[r]] = from_tkey_m(static_cast<uint32_t>(v[r] >> 32));
      }
    }
    return;
  }
  // sort only as many keys as remain (k <= cnt2 <= 64 R2)
  auto finish = [&](auto rr) {
    constexpr int R2 = decltype(rr)::value;
    uint64_t v[R2];
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      const uint32_t e = r * 64 + lane;
      v[r] = e < cnt2 ? keys[e] : ~0ull;
    }
    wave_bitonic_m<R2>(v, lane);
    const float dk = from_tkey_m(static_cast<uint32_t>(wave_pick_m<R2>(v, k - 1) >> 32));
    const float t = gtau[q];
    // every unlisted candidate has a true squared distance > tau (up to the
    // rounding factor folded in cert_scale), hence a reference distance > dk
    const bool cert = all_listed || cert_scale < 0.0 ||
                      (t > 0.0f && sqrt(static_cast<double>(t) * cert_scale) > static_cast<double>(dk));
    if (!cert) {
      if (lane == 0) fb_list[atomicAdd(fb_count, 1u)] = q;
      return;
    }
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      const uint32_t e = r * 64 + lane;
      if (e < k) {
        oidx[static_cast<size_t>(q) * k + e] = static_cast<uint32_t>(v[r]);
        odist[static_cast<size_t>(q) * k + e] = from_tkey_m(static_cast<uint32_t>(v[r] >> 32));
      }
    }
  };
  if (R >= 4 && cnt2 <= 256)
    finish(std::integral_constant<int, (R >= 4 ? 4 : R)>{});
  else
    finish(std::integral_constant<int, R>{});
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t launch_knn_frag(const float *x, uint64_t n, uint32_t d, const double *mean, uint32_t kch,
                           uint64_t ntiles, const uint32_t *perm, void *fhi, void *flo, hipStream_t s) {
  const uint64_t total = ntiles * kch * 64;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  knn_frag_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(x, n, d, mean, kch, ntiles, perm,
                                                                static_cast<uint4 *>(fhi),
                                                                static_cast<uint4 *>(flo));
  return hipGetLastError();
}

hipError_t launch_knn_frag8(const float *x, uint64_t n, uint32_t d, const double *mean, uint32_t kch,
                            uint64_t ntiles, const float *rinv, const uint32_t *perm, void *fa, void *fb,
                            hipStream_t s) {
  const uint64_t total = ntiles * kch * 64;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  knn_frag8_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(x, n, d, mean, kch, ntiles, rinv, perm,
                                                                 static_cast<uint4 *>(fa), static_cast<uint4 *>(fb));
  return hipGetLastError();
}

hipError_t launch_knn_rownorm(const float *x, uint64_t n, uint32_t d, const double *mean,
                              float one_minus_kappa, uint64_t rows_padded, const uint32_t *perm, float *arow,
                              float *arow_orig, uint32_t *nonfinite, float *rinv, hipStream_t s) {
  if (d >= 256)
    knn_rownorm_kernel<64><<<static_cast<uint32_t>((rows_padded + 3) / 4), 256, 0, s>>>(
        x, n, d, mean, one_minus_kappa, rows_padded, perm, arow, arow_orig, nonfinite, rinv);
  else
    knn_rownorm_kernel<16><<<static_cast<uint32_t>((rows_padded + 15) / 16), 256, 0, s>>>(
        x, n, d, mean, one_minus_kappa, rows_padded, perm, arow, arow_orig, nonfinite, rinv);
  return hipGetLastError();
}

constexpr int kCandRT = 2;  // query tiles per wave in the candidate kernels

uint32_t knn_cand_groups(uint32_t ntiles) { return (ntiles + 4 * kCandRT - 1) / (4 * kCandRT); }

// 0 when the query fails (the sort then reports the error)
size_t knn_order_temp_bytes(uint64_t n) {
  size_t bytes = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<const uint32_t *>(nullptr),
                                         static_cast<uint32_t *>(nullptr), static_cast<const uint32_t *>(nullptr),
                                         static_cast<uint32_t *>(nullptr), static_cast<int>(n), 0,
                                         kKnnOrderBits * kKnnOrderDims) != hipSuccess)
    return 0;
  return bytes;
}

hipError_t launch_knn_order(const float *x, uint64_t n, uint32_t d, uint32_t *mm, uint32_t *key, uint32_t *key_sorted,
                            uint32_t *idx, uint32_t *perm, void *temp, size_t temp_bytes, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t kd = d < static_cast<uint32_t>(kKnnOrderDims) ? d : static_cast<uint32_t>(kKnnOrderDims);
  knn_order_init_kernel<<<1, 64, 0, s>>>(mm);
  const uint64_t blocks = (n + 255) / 256;
  knn_order_minmax_kernel<<<static_cast<uint32_t>(blocks < 512 ? blocks : 512), 256, 0, s>>>(x, n, d, kd, mm);
  knn_order_key_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(x, n, d, kd, mm, key, idx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, key, key_sorted, idx, perm, static_cast<int>(n), 0,
                                            kKnnOrderBits * kKnnOrderDims, s);
}

__global__ void knn_sample_rows_kernel(uint64_t n, uint64_t m, uint32_t stride, uint32_t *__restrict__ rows) {
  const uint64_t p = blockIdx.x * 256ull + threadIdx.x;
  if (p < m) rows[p] = static_cast<uint32_t>((static_cast<uint64_t>(stride) * p) % n);
}

uint32_t knn_sample_stride(uint64_t n) {
  if (n < 3) return 1u;
  auto gcd = [](uint64_t a, uint64_t b) {
    while (b) {
      const uint64_t t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  uint64_t a = static_cast<uint64_t>(0.6180339887498949 * static_cast<double>(n)) | 1u;
  while (gcd(a, n) != 1) a += 2;
  return static_cast<uint32_t>(a % n);
}

hipError_t launch_knn_sample_rows(uint64_t n, uint64_t m, uint32_t stride, uint32_t *rows, hipStream_t s) {
  if (m == 0) return hipSuccess;
  knn_sample_rows_kernel<<<static_cast<uint32_t>((m + 255) / 256), 256, 0, s>>>(n, m, stride, rows);
  return hipGetLastError();
}

uint32_t knn_box_dims(uint32_t d) { return d < static_cast<uint32_t>(kKnnBoxDims) ? d : static_cast<uint32_t>(kKnnBoxDims); }

hipError_t launch_knn_tilebox(const float *x, uint64_t n, uint32_t d, const uint32_t *perm, uint64_t ntiles,
                              float *box, hipStream_t s) {
  const uint32_t mb = knn_box_dims(d);
  const uint64_t total = ntiles * mb;
  if (total == 0) return hipSuccess;
  knn_tilebox_kernel<<<static_cast<uint32_t>((total + 255) / 256), 256, 0, s>>>(x, n, d, mb, perm, ntiles, box);
  return hipGetLastError();
}

hipError_t launch_knn_tiletau(const float *gtau, const uint32_t *perm, uint64_t n, uint64_t ntiles, float *tmax,
                              float *gtau_orig, hipStream_t s) {
  if (ntiles == 0) return hipSuccess;
  knn_tiletau_kernel<<<static_cast<uint32_t>((ntiles + 255) / 256), 256, 0, s>>>(gtau, perm, n, ntiles, tmax,
                                                                                gtau_orig);
  return hipGetLastError();
}

hipError_t launch_knn_prune(const float *box, const float *tmax, uint64_t ntiles, uint32_t d, uint32_t words,
                            float *gbox, uint32_t *bits, unsigned long long *count, hipStream_t s) {
  const uint32_t groups = knn_cand_groups(static_cast<uint32_t>(ntiles));
  if (groups == 0 || words == 0) return hipSuccess;
  knn_groupbox_kernel<<<(groups + 255) / 256, 256, 0, s>>>(box, tmax, ntiles, groups, 4 * kCandRT, knn_box_dims(d),
                                                           gbox);
  knn_prune_kernel<<<dim3(words, (groups + 255) / 256), 256, 0, s>>>(box, tmax, gbox, ntiles, groups, 4 * kCandRT,
                                                                     knn_box_dims(d), words, bits, count);
  return hipGetLastError();
}


template <int KCH, int MODE>
static hipError_t cand_dispatch(int phase, const void *fhi, const void *flo, const float *arow,
                                const float *rinv, uint32_t n, uint32_t ntiles, uint32_t ss, uint32_t capg,
                                uint32_t nsplit, uint32_t *gcnt, uint32_t *gcand, float *gtau,
                                const KnnSymPlan &sp, const KnnOrder &o, hipStream_t s) {
  const uint32_t groups = knn_cand_groups(ntiles);
  const uint4 *h = static_cast<const uint4 *>(fhi), *l = static_cast<const uint4 *>(flo);
  if (phase == 0)
    knn_mfma_kernel<KCH, kCandRT, MODE, 0><<<dim3(groups, 1), 256, 0, s>>>(
        h, l, arow, rinv, n, ntiles, ss, capg, gcnt, gcand, gtau, KnnSymPlan{}, o.perm, nullptr, 0u, o.sample);
  else if (phase == 1)
    knn_mfma_kernel<KCH, kCandRT, MODE, 1><<<dim3(groups, nsplit), 256, 0, s>>>(
        h, l, arow, rinv, n, ntiles, ss, capg, gcnt, gcand, gtau, KnnSymPlan{}, o.perm, nullptr, 0u, KnnSampleSet{});
  else if (sp.items)
    knn_mfma_kernel<KCH, kCandRT, MODE, 2><<<dim3(sp.items, 1), 256, 0, s>>>(
        h, l, arow, rinv, n, ntiles, ss, capg, gcnt, gcand, gtau, sp, o.perm, o.prune, o.prune ? o.pwords : 0u,
        KnnSampleSet{});
  return hipGetLastError();
}

template <int KCH>
static hipError_t cand_dispatch_mode(int mode, int phase, const void *fhi, const void *flo,
                                     const float *arow, const float *rinv, uint32_t n, uint32_t ntiles,
                                     uint32_t ss, uint32_t capg, uint32_t nsplit, uint32_t *gcnt,
                                     uint32_t *gcand, float *gtau, const KnnSymPlan &sp, const KnnOrder &o,
                                     hipStream_t s) {
  switch (mode) {
    case 0: return cand_dispatch<KCH, 0>(phase, fhi, flo, arow, rinv, n, ntiles, ss, capg, nsplit, gcnt, gcand, gtau, sp, o, s);
    case 1: return cand_dispatch<KCH, 1>(phase, fhi, flo, arow, rinv, n, ntiles, ss, capg, nsplit, gcnt, gcand, gtau, sp, o, s);
    case 2: return cand_dispatch<KCH, 2>(phase, fhi, flo, arow, rinv, n, ntiles, ss, capg, nsplit, gcnt, gcand, gtau, sp, o, s);
    default: return hipErrorInvalidValue;
  }
}

// Plan of the symmetric pass: the candidate tiles are cut into windows of
// L tiles (a multiple of the query group's gt tiles); group g takes, in
// each window that reaches past its own first tile, the window's tiles from
// that first tile on.  Items run window by window, so the workgroups
// resident together stream the same candidate tiles and share them in L2.
// L gives about target_items items and is at most max_window tiles.
KnnSymPlan knn_sym_plan(uint32_t ntiles, uint32_t target_items, uint32_t max_window) {
  KnnSymPlan p;
  p.gt = 4 * kCandRT;
  const uint64_t groups = (ntiles + p.gt - 1) / p.gt;
  uint64_t tri = 0;
  for (uint64_t g = 0; g < groups; ++g) tri += ntiles - g * p.gt;
  uint64_t len = (tri + target_items - 1) / target_items;
  len = std::min<uint64_t>(len, max_window);
  len = std::max<uint64_t>(len / p.gt, 2) * p.gt;
  p.L = static_cast<uint32_t>(len);
  p.m = p.L / p.gt;
  p.W = static_cast<uint32_t>((ntiles + len - 1) / len);
  if (ntiles == 0) p.W = 0;
  const uint64_t items = p.W ? static_cast<uint64_t>(p.m) * (p.W - 1) * p.W / 2 + groups : 0;
  p.items = items < (1ull << 31) ? static_cast<uint32_t>(items) : 0;  // 0: the caller takes the split pass
  return p;
}

hipError_t launch_knn_cand(int phase, const void *fhi, const void *flo, const float *arow, const float *rinv,
                           uint32_t n, uint32_t kch, uint32_t ntiles, uint32_t ss, uint32_t capg,
                           uint32_t nsplit, uint32_t *gcnt, uint32_t *gcand, float *gtau,
                           int mode, hipStream_t s, const KnnSymPlan &sp, const KnnOrder &o) {
  if (mode == 2 && rinv == nullptr) return hipErrorInvalidValue;
  if (o.perm == nullptr) return hipErrorInvalidValue;
  if (phase == 2 && sp.items && (sp.L == 0 || sp.L % sp.gt != 0 || sp.m * sp.gt != sp.L || sp.W == 0 ||
                                 sp.gt != 4 * kCandRT))
    return hipErrorInvalidValue;
#define PM_CAND(K) \
  case K: return cand_dispatch_mode<K>(mode, phase, fhi, flo, arow, rinv, n, ntiles, ss, capg, nsplit, gcnt, gcand, gtau, sp, o, s);
  switch (kch) {
    PM_CAND(1)
    PM_CAND(2)
    PM_CAND(3)
    PM_CAND(4)
    PM_CAND(5)
    PM_CAND(6)
    PM_CAND(7)
    PM_CAND(8)
    default: return hipErrorInvalidValue;
  }
#undef PM_CAND
}

hipError_t launch_knn_rerank(const float *xp, const KnnLayout &L, uint32_t n, uint32_t k,
                             uint32_t capg, uint32_t nsplit, const uint32_t *gcnt,
                             const uint32_t *gcand, const float *gtau, double cert_scale,
                             const float *arow, bool dots, double ub_s1, double ub_s2,
                             uint32_t *idx, float *dist, uint32_t *fb_count, uint32_t *fb_list,
                             unsigned long long *stats, hipStream_t s) {
  const uint32_t blocks = (n + 3) / 4;
  const uint32_t R = capg / 64;
  const size_t shm = static_cast<size_t>(4) * (L.Dp + 192 * R) * sizeof(float);
  if (nsplit > 16) return hipErrorInvalidValue;
  if (capg == 512)
    knn_rerank_kernel<8><<<blocks, 256, shm, s>>>(xp, L.Dp, L.Cp, L.R, n, k, capg, nsplit, gcnt, gcand, gtau,
                                                  cert_scale, arow, dots ? 1 : 0, ub_s1, ub_s2, idx, dist,
                                                  fb_count, fb_list, stats);
  else if (capg == 1024)
    knn_rerank_kernel<16><<<blocks, 256, shm, s>>>(xp, L.Dp, L.Cp, L.R, n, k, capg, nsplit, gcnt, gcand, gtau,
                                                   cert_scale, arow, dots ? 1 : 0, ub_s1, ub_s2, idx, dist,
                                                   fb_count, fb_list, stats);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace pm
