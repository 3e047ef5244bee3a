// grad_device.h -- device helpers shared by the gradient kernels
// (kernels.hip, grad_paged.hip): the per-pair term of gradient.rs:118-148,
// its per-type constants, and the staging copy of Y into LDS.  Every
// gradient kernel ends in pair_term_ck, so all of them round each term
// identically.
#pragma once

#include "internal.h"

namespace pm {

constexpr uint32_t kLdsPerCu = 163840;
// |yr - far|^2 <= 8 (1e18)^2 stays finite; 1 / (c + D) >= 1.2e-37 is a
// normal number; its square is far below the smallest denormal.  Padding
// entries point at a staged point with this value in every coordinate, so
// their term is a signed zero and no loss.
constexpr float kSellFar = 1.0e18f;

// 1/x from v_rcp_f32 refined by one Newton step (x >= 1 here): within an ulp
// of the correctly rounded quotient, at a quarter of the instructions of an
// IEEE division.
__device__ __forceinline__ float recip_nr(float x) {
  const float r0 = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r0, 1.0f);
  return fmaf(r0, e, r0);
}

// Per-type constants of the pair terms (gradient.rs:127-147), one set per
// iteration: k = w * numerator constant with the far-pair sign folded in (so
// every entry adds k rc^2 u), wt = w for the loss.  Kept as separate scalars:
// selecting between members of an aggregate can turn into an indexed memory
// load per entry.
struct PairConsts {
  float k_nn, k_mn, k_fp;
  float w_nn, w_mn, w_fp;
};
__device__ __forceinline__ PairConsts pair_consts(const float (&w)[3]) {
  PairConsts q;
  q.k_nn = w[0] * 20.0f;
  q.k_mn = w[1] * 20000.0f;
  q.k_fp = -(w[2] * 2.0f);
  q.w_nn = w[0];
  q.w_mn = w[1];
  q.w_fp = w[2];
  return q;
}

// One CSR entry's contribution (gradient.rs:118-148) with the pair type's
// constants already selected: w = entry word, c = 10 / 10000 / 1 and
// k = its numerator constant (NN / MN / FP), yr = the row's point, yo = the
// other point.  Every gradient kernel below ends here, so all of them round
// each term identically.  D = 1 + |yr - yo|^2 is accumulated in the
// reference's order with fused multiply-adds, as is the gradient (one
// rounding per step instead of the reference's two); the coefficient is
// k / (c + D)^2 with the reciprocal from recip_nr.
template <int DIMS, bool LOSS>
__device__ __forceinline__ void pair_term_ck(uint32_t w, float c, float k, const float (&yr)[DIMS],
                                             const float *yo, float w_nn, float w_mn, float w_fp,
                                             float (&g)[DIMS], float &loss) {
  float u[DIMS];
  float D = 1.0f;  // gradient.rs:121-125
#pragma unroll
  for (int d = 0; d < DIMS; ++d) {
    u[d] = yr[d] - yo[d];
    D = fmaf(u[d], u[d], D);
  }
  const float rc = recip_nr(c + D);
  const float w1 = k * (rc * rc);  // gradient.rs:130 / 140 (sign folded into k)
#pragma unroll
  for (int d = 0; d < DIMS; ++d) g[d] = fmaf(w1, u[d], g[d]);
  if (LOSS && (w & kOwnerBit)) {
    const bool nn = w < (1u << 30), fp = w >= (2u << 30);
    const float wt = nn ? w_nn : (fp ? w_fp : w_mn);
    const float num = fp ? 1.0f : D;  // gradient.rs:129 / 139
    loss = loss + wt * (num * rc);
  }
}

// pair_term_ck with the type read from the word's top bits (0 NN, 1 MN,
// 2 and 3 FP).
template <int DIMS, bool LOSS>
__device__ __forceinline__ void pair_term(uint32_t w, const float (&yr)[DIMS], const float *yo, float k_nn,
                                          float k_mn, float k_fp, float w_nn, float w_mn, float w_fp,
                                          float (&g)[DIMS], float &loss) {
  const bool nn = w < (1u << 30), fp = w >= (2u << 30);
  const float c = nn ? 10.0f : (fp ? 1.0f : 10000.0f);
  const float k = nn ? k_nn : (fp ? k_fp : k_mn);
  pair_term_ck<DIMS, LOSS>(w, c, k, yr, yo, w_nn, w_mn, w_fp, g, loss);
}

// Copies count floats from global memory to LDS with a T-thread
// workgroup: 16-byte copies (both addresses 16-byte aligned), 8 in flight
// per thread, then the remainder and the last < 4 floats.
template <uint32_t T>
__device__ __forceinline__ void stage_copy(float *__restrict__ dst, const float *__restrict__ src, uint32_t count,
                                           uint32_t t) {
  const uint32_t n4 = count / 4;
  const float4 *s4 = reinterpret_cast<const float4 *>(src);
  float4 *d4 = reinterpret_cast<float4 *>(dst);
  uint32_t i0 = t;
  for (; i0 + 7 * T < n4; i0 += 8 * T) {
    const float4 v0 = s4[i0], v1 = s4[i0 + T], v2 = s4[i0 + 2 * T], v3 = s4[i0 + 3 * T];
    const float4 v4 = s4[i0 + 4 * T], v5 = s4[i0 + 5 * T], v6 = s4[i0 + 6 * T], v7 = s4[i0 + 7 * T];
    d4[i0] = v0;
    d4[i0 + T] = v1;
    d4[i0 + 2 * T] = v2;
    d4[i0 + 3 * T] = v3;
    d4[i0 + 4 * T] = v4;
    d4[i0 + 5 * T] = v5;
    d4[i0 + 6 * T] = v6;
    d4[i0 + 7 * T] = v7;
  }
  for (; i0 < n4; i0 += T) d4[i0] = s4[i0];
  if (t < count % 4) dst[n4 * 4 + t] = src[n4 * 4 + t];
}

// Range b's segment of a row in the tb_partition layout: [beg, end) relative
// to the row start (off: kTargetBlocks x n u16 range starts).
__device__ __forceinline__ uint32_t tb_seg_beg(const uint16_t *__restrict__ off, uint32_t n, uint32_t b,
                                               uint32_t row) {
  return b == 0 ? 0u : off[static_cast<size_t>(b) * n + row];
}
__device__ __forceinline__ uint32_t tb_seg_end(const uint32_t *__restrict__ row_ptr,
                                               const uint16_t *__restrict__ off, uint32_t n, uint32_t b,
                                               uint32_t row) {
  return b + 1 == kTargetBlocks ? row_ptr[row + 1] - row_ptr[row] : off[static_cast<size_t>(b + 1) * n + row];
}

}  // namespace pm
