// mlp.hip -- parametric encoder d -> H -> H -> dims (GELU), trained with the
// PaCMAP pair loss.  BF16 weights and activations feed v_mfma_f32_16x16x32_bf16
// with FP32 accumulation; FP32 master weights and Adam state.
//
// A batch of B pairs becomes M = 2B rows (row p: first endpoint of pair p,
// row B + p: second endpoint).  Every reduction (split-K partials, bias sums,
// loss) is summed in a fixed order, so training is deterministic.
#include "mlp_internal.h"

#include <hip/hip_bf16.h>

#include <mutex>

namespace pm {

typedef __bf16 bf16x8m __attribute__((ext_vector_type(8)));
typedef float floatx4m __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  return __bfloat16_as_ushort(__float2bfloat16(f));  // round to nearest even
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }

// GELU with the exact (erf) form and its derivative
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// ---------------------------------------------------------------------------
// Batch rows: pair p of the batch is pair index perm(start + p) of the typed
// list (NN rows first, then MN, then FP); perm(pos) = (a pos + c) mod P.
// ---------------------------------------------------------------------------
// One block per 64 batch rows: the bf16 tile goes through LDS so that both
// xb (rows) and xbT (features x rows) are written with contiguous stores.
__global__ __launch_bounds__(256) void mlp_gather_kernel(const float *__restrict__ x, uint32_t d, uint32_t Kp,
                                                         const uint32_t *__restrict__ pi,
                                                         const uint32_t *__restrict__ pj, uint64_t P,
                                                         uint64_t start, uint32_t count, uint64_t pa, uint64_t pc,
                                                         uint32_t B, uint32_t M, uint16_t *__restrict__ xb,
                                                         uint16_t *__restrict__ xbT, uint32_t *__restrict__ pidx) {
  extern __shared__ uint16_t tile[];  // 64 x (Kp + 8)
  __shared__ uint32_t src[64];
  __shared__ bool valid[64];
  const uint32_t ld = Kp + 8, r0 = blockIdx.x * 64, t = threadIdx.x;
  if (t < 64) {
    const uint32_t r = r0 + t;
    uint32_t sidx = 0;
    bool ok = false;
    if (r < 2 * B) {
      const uint32_t p = r < B ? r : r - B;
      if (p < count) {
        const uint64_t pos = start + p;
        const uint64_t e = pa == 0 ? pos : (pa * pos + pc) % P;  // pa == 0: identity order
        sidx = r < B ? pi[e] : pj[e];
        ok = true;
        if (r < B) pidx[p] = static_cast<uint32_t>(e);
      } else if (r < B) {
        pidx[p] = 0xffffffffu;
      }
      ok = ok || (r < 2 * B);  // rows of padding pairs read point 0
    }
    src[t] = sidx;
    valid[t] = ok;
  }
  __syncthreads();
  for (uint32_t e = t; e < 64 * Kp; e += 256) {
    const uint32_t rr = e / Kp, k = e % Kp;
    const uint16_t v = (valid[rr] && k < d) ? f2bf(x[static_cast<size_t>(src[rr]) * d + k]) : 0;
    tile[rr * ld + k] = v;
    xb[static_cast<size_t>(r0 + rr) * Kp + k] = v;
  }
  __syncthreads();
  // xbT: for feature k, the 64 rows are 128 contiguous bytes (8 uint4 per k)
  for (uint32_t e = t; e < Kp * 8; e += 256) {
    const uint32_t k = e / 8, q = e % 8;
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      w[j] = tile[(8 * q + 2 * j) * ld + k] | (static_cast<uint32_t>(tile[(8 * q + 2 * j + 1) * ld + k]) << 16);
    *reinterpret_cast<uint4 *>(xbT + static_cast<size_t>(k) * M + r0 + 8 * q) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// Transform rows: x rows [row0, row0 + count) in order, padded with zeros.
__global__ void mlp_rows_kernel(const float *__restrict__ x, uint32_t d, uint32_t Kp, uint64_t row0,
                                uint32_t count, uint16_t *__restrict__ xb) {
  const uint32_t r = blockIdx.x;
  for (uint32_t k = threadIdx.x; k < Kp; k += blockDim.x) {
    const uint16_t v = (r < count && k < d) ? f2bf(x[(row0 + r) * d + k]) : 0;
    xb[static_cast<size_t>(r) * Kp + k] = v;
  }
}

// ---------------------------------------------------------------------------
// C (M x N) = A (M x K) B^T with A, B row-major bf16 (K contiguous), FP32
// accumulation.  64 x 64 block tile, 4 waves of 32 x 32; each k-step of 32
// loads the MFMA operands straight from global memory (16 B per lane).
// Requires M, N multiples of 64 and K (and kchunk) multiples of 32.
// ---------------------------------------------------------------------------
struct MlpEpi {
  const float *bias = nullptr;   // EPI_FWD: bias[n]
  float *z = nullptr;            // EPI_FWD: pre-activation out (M x N)
  uint16_t *hb = nullptr;        // EPI_FWD: bf16 activation (M x N)
  uint16_t *hT = nullptr;        // EPI_FWD: bf16 activation transposed (N x M)
  float *hf = nullptr;           // EPI_FWD: FP32 activation (M x N), optional
  const float *zin = nullptr;    // EPI_DH: pre-activation of the layer below (M x N)
  uint16_t *dzT = nullptr;       // EPI_DH: bf16 dz transposed (N x M)
  float *part = nullptr;         // EPI_PART: partial slab z at part + z * M * N
};
enum { EPI_FWD = 0, EPI_DH = 1, EPI_PART = 2 };

template <int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const uint16_t *__restrict__ A, uint32_t lda,
                                                       const uint16_t *__restrict__ B, uint32_t ldb,
                                                       uint32_t M, uint32_t N, uint32_t K, uint32_t kchunk,
                                                       MlpEpi e) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t m0 = blockIdx.x * 64 + (wave >> 1) * 32;
  const uint32_t n0 = blockIdx.y * 64 + (wave & 1) * 32;
  const uint32_t kbeg = blockIdx.z * kchunk;
  const uint32_t kend = kbeg + kchunk < K ? kbeg + kchunk : K;
  floatx4m acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4m{0.f, 0.f, 0.f, 0.f};
  const uint32_t lr = lane & 15, lk = 8 * (lane >> 4);
  const uint16_t *a0 = A + static_cast<size_t>(m0 + lr) * lda + lk;
  const uint16_t *a1 = A + static_cast<size_t>(m0 + 16 + lr) * lda + lk;
  const uint16_t *b0 = B + static_cast<size_t>(n0 + lr) * ldb + lk;
  const uint16_t *b1 = B + static_cast<size_t>(n0 + 16 + lr) * ldb + lk;
  // up to 8 k-steps (256 values of k) of operands are loaded before any MFMA
  constexpr int KS = 8;
  for (uint32_t k0 = kbeg; k0 < kend; k0 += 32 * KS) {
    uint4 ra0[KS], ra1[KS], rb0[KS], rb1[KS];
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      const uint32_t k = k0 + 32 * q;
      if (k < kend) {
        ra0[q] = *reinterpret_cast<const uint4 *>(a0 + k);
        ra1[q] = *reinterpret_cast<const uint4 *>(a1 + k);
        rb0[q] = *reinterpret_cast<const uint4 *>(b0 + k);
        rb1[q] = *reinterpret_cast<const uint4 *>(b1 + k);
      }
    }
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      if (k0 + 32 * q >= kend) break;
      const bf16x8m va0 = __builtin_bit_cast(bf16x8m, ra0[q]), va1 = __builtin_bit_cast(bf16x8m, ra1[q]);
      const bf16x8m vb0 = __builtin_bit_cast(bf16x8m, rb0[q]), vb1 = __builtin_bit_cast(bf16x8m, rb1[q]);
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb1, acc[1][1], 0, 0, 0);
    }
  }
  // lane holds rows 4 (lane >> 4) + r (r < 4) of column lane & 15 of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t mb = m0 + 16 * i + 4 * (lane >> 4);
      const uint32_t n = n0 + 16 * j + lr;
      if (EPI == EPI_PART) {
        float *P = e.part + static_cast<size_t>(blockIdx.z) * M * N;
#pragma unroll
        for (int r = 0; r < 4; ++r) P[static_cast<size_t>(mb + r) * N + n] = acc[i][j][r];
      } else if (EPI == EPI_FWD) {
        uint16_t hv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float zv = acc[i][j][r] + e.bias[n];
          const float h = gelu(zv);
          const size_t o = static_cast<size_t>(mb + r) * N + n;
          if (e.z) e.z[o] = zv;
          hv[r] = f2bf(h);
          if (e.hb) e.hb[o] = hv[r];
          if (e.hf) e.hf[o] = h;
        }
        if (e.hT) {
          const uint2 pk = make_uint2(hv[0] | (static_cast<uint32_t>(hv[1]) << 16),
                                      hv[2] | (static_cast<uint32_t>(hv[3]) << 16));
          *reinterpret_cast<uint2 *>(e.hT + static_cast<size_t>(n) * M + mb) = pk;
        }
      } else {  // EPI_DH: dz = (A B^T) * gelu'(z)
        uint16_t dv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const size_t o = static_cast<size_t>(mb + r) * N + n;
          dv[r] = f2bf(acc[i][j][r] * gelu_grad(e.zin[o]));
        }
        const uint2 pk = make_uint2(dv[0] | (static_cast<uint32_t>(dv[1]) << 16),
                                    dv[2] | (static_cast<uint32_t>(dv[3]) << 16));
        *reinterpret_cast<uint2 *>(e.dzT + static_cast<size_t>(n) * M + mb) = pk;
      }
    }
}

// ---------------------------------------------------------------------------
// Output layer (FP32): Y[m][c] = sum_i h2[m][i] W3[c][i] + b3[c]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mlp_out_kernel(const float *__restrict__ h2, uint32_t H,
                                                      const float *__restrict__ W3, const float *__restrict__ b3,
                                                      uint32_t dims, uint32_t rows, float *__restrict__ Y) {
  extern __shared__ float sW[];  // dims x H
  for (uint32_t e = threadIdx.x; e < dims * H; e += blockDim.x) sW[e] = W3[e];
  __syncthreads();
  // 4 lanes per row, lane q takes the columns q*4 + 16 j (float4 loads)
  const uint32_t m = blockIdx.x * 64 + (threadIdx.x >> 2), q = threadIdx.x & 3;
  float s[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) s[c] = 0.0f;
  if (m < rows) {
    const float *hr = h2 + static_cast<size_t>(m) * H;
    for (uint32_t i = 4 * q; i < H; i += 16) {
      const float4 h = *reinterpret_cast<const float4 *>(hr + i);
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (static_cast<uint32_t>(c) < dims) {
          const float *w = sW + c * H + i;
          s[c] = fmaf(h.x, w[0], fmaf(h.y, w[1], fmaf(h.z, w[2], fmaf(h.w, w[3], s[c]))));
        }
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    s[c] += __shfl_xor(s[c], 1, 64);
    s[c] += __shfl_xor(s[c], 2, 64);
  }
  if (m < rows && q == 0)
    for (uint32_t c = 0; c < dims; ++c) Y[static_cast<size_t>(m) * dims + c] = s[c] + b3[c];
}

// ---------------------------------------------------------------------------
// Pair loss (gradient.rs:98-151 terms on the encoder outputs): for pair p
// with u = Y[p] - Y[B + p], D = 1 + |u|^2: NN w D/(10 + D), MN w D/(1e4 + D),
// FP w/(1 + D); dY = +-w1 u / count as in process_pairs.  Self pairs and
// padding rows contribute nothing.  Loss partials per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mlp_pair_loss_kernel(const float *__restrict__ Y, uint32_t dims,
                                                            uint32_t B, const uint32_t *__restrict__ pidx,
                                                            const uint32_t *__restrict__ pi,
                                                            const uint32_t *__restrict__ pj, uint64_t nn_end,
                                                            uint64_t mn_end, float w_nn, float w_mn, float w_fp,
                                                            float inv_count, float *__restrict__ dY,
                                                            float *__restrict__ loss_part) {
  __shared__ float red[256];
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  float loss = 0.0f;
  if (p < B) {
    const uint32_t e = pidx[p];
    float w1 = 0.0f;
    float u[8];
    float D = 1.0f;
    for (uint32_t c = 0; c < dims; ++c) {
      u[c] = Y[static_cast<size_t>(p) * dims + c] - Y[static_cast<size_t>(B + p) * dims + c];
      D += u[c] * u[c];
    }
    if (e != 0xffffffffu && pi[e] != pj[e]) {
      if (e < nn_end) {
        loss = w_nn * (D / (10.0f + D));
        w1 = w_nn * (20.0f / ((10.0f + D) * (10.0f + D)));
      } else if (e < mn_end) {
        loss = w_mn * (D / (10000.0f + D));
        w1 = w_mn * (20000.0f / ((10000.0f + D) * (10000.0f + D)));
      } else {
        loss = w_fp / (1.0f + D);
        w1 = -w_fp * (2.0f / ((1.0f + D) * (1.0f + D)));
      }
    }
    for (uint32_t c = 0; c < dims; ++c) {
      const float g = w1 * u[c] * inv_count;
      dY[static_cast<size_t>(p) * dims + c] = g;
      dY[static_cast<size_t>(B + p) * dims + c] = -g;
    }
  }
  red[threadIdx.x] = loss;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Output-layer backward: dh2 = dY W3, dz2 = dh2 gelu'(z2) (bf16, row-major
// and transposed); per-block partials of dW3 = dY^T h2 and db3 = sum dY.
// One block per 64 rows, thread per hidden unit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mlp_out_bwd_kernel(const float *__restrict__ dY, uint32_t dims,
                                                          const float *__restrict__ W3, const float *__restrict__ h2,
                                                          const float *__restrict__ z2, uint32_t H, uint32_t M,
                                                          uint16_t *__restrict__ dz2b, uint16_t *__restrict__ dz2T,
                                                          float *__restrict__ w3_part, float *__restrict__ b3_part) {
  constexpr uint32_t RB = 16;  // rows per block
  const uint32_t m0 = blockIdx.x * RB;
  __shared__ float sdY[RB * 8];
  if (threadIdx.x < RB * dims) sdY[threadIdx.x] = dY[static_cast<size_t>(m0) * dims + threadIdx.x];
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) {
    float w[8], gw[8];
    for (uint32_t c = 0; c < dims; ++c) {
      w[c] = W3[c * H + i];
      gw[c] = 0.0f;
    }
    uint32_t pk[RB / 2];
#pragma unroll
    for (uint32_t r = 0; r < RB; ++r) {
      const size_t o = static_cast<size_t>(m0 + r) * H + i;
      const float hv = h2[o], zv = z2[o];
      float dh = 0.0f;
      for (uint32_t c = 0; c < dims; ++c) {
        const float g = sdY[r * dims + c];
        dh = fmaf(g, w[c], dh);
        gw[c] = fmaf(g, hv, gw[c]);
      }
      const uint16_t dv = f2bf(dh * gelu_grad(zv));
      dz2b[o] = dv;
      if (r & 1) pk[r / 2] |= static_cast<uint32_t>(dv) << 16;
      else pk[r / 2] = dv;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(dz2T + static_cast<size_t>(i) * M + m0);
    dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
    for (uint32_t c = 0; c < dims; ++c) w3_part[(static_cast<size_t>(blockIdx.x) * dims + c) * H + i] = gw[c];
  }
  if (threadIdx.x < dims) {
    float sb = 0.0f;
    for (uint32_t r = 0; r < RB; ++r) sb += sdY[r * dims + threadIdx.x];
    b3_part[static_cast<size_t>(blockIdx.x) * dims + threadIdx.x] = sb;
  }
}

// bias gradients: out[i] = sum_m dzT[i][m] (bf16 values, FP32 sum, fixed order)
__global__ __launch_bounds__(256) void mlp_rowsum_kernel(const uint16_t *__restrict__ dzT, uint32_t M,
                                                         float *__restrict__ out) {
  __shared__ float red[256];
  const uint4 *r = reinterpret_cast<const uint4 *>(dzT + static_cast<size_t>(blockIdx.x) * M);
  float s = 0.0f;
  for (uint32_t m = threadIdx.x; m < M / 8; m += 256) {  // M is a multiple of 64
    const uint4 v = r[m];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) s += bf2f(static_cast<uint16_t>(w[j] & 0xffffu)) + bf2f(static_cast<uint16_t>(w[j] >> 16));
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// out[o] = sum over z < S of part[z * slab + o] for o < count: one block per
// output, 256 strided partial sums, then a fixed-order tree
__global__ __launch_bounds__(256) void mlp_tree_sum_kernel(const float *__restrict__ part, uint32_t S, uint64_t slab,
                                                           float *__restrict__ out) {
  __shared__ float red[256];
  const uint32_t o = blockIdx.x;
  float s = 0.0f;
  for (uint32_t z = threadIdx.x; z < S; z += 256) s += part[z * slab + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[o] = red[0];
}

// out[r][c] (r < rows, c < cols, row stride ldo) = sum over s < S of
// part[s][r][c] (partial row stride ldp), in order of s
__global__ void mlp_sum_parts_kernel(const float *__restrict__ part, uint32_t S, uint32_t rows, uint32_t cols,
                                     uint32_t ldp, uint64_t slab, float *__restrict__ out, uint32_t ldo) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<uint64_t>(rows) * cols) return;
  const uint32_t r = static_cast<uint32_t>(e / cols), c = static_cast<uint32_t>(e % cols);
  float s = 0.0f;
  for (uint32_t z = 0; z < S; ++z) s += part[z * slab + static_cast<size_t>(r) * ldp + c];
  out[static_cast<size_t>(r) * ldo + c] = s;
}

// ---------------------------------------------------------------------------
// Adam on the FP32 master parameters, then the bf16 copies used by the GEMMs:
// W1b (H x Kp, zero padded), W2b (H x H) and W2tb (H x H, transposed).
// ---------------------------------------------------------------------------
__global__ void mlp_adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                float *__restrict__ v, uint64_t count, float lr_t, float b1, float b2,
                                float eps) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const float ge = g[e];
  const float me = b1 * m[e] + (1.0f - b1) * ge;
  const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  p[e] = p[e] - lr_t * me / (sqrtf(ve) + eps);
}

__global__ void mlp_pack_kernel(const float *__restrict__ p, uint32_t d, uint32_t Kp, uint32_t H,
                                uint16_t *__restrict__ w1b, uint16_t *__restrict__ w2b,
                                uint16_t *__restrict__ w2tb) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t n1 = static_cast<uint64_t>(H) * Kp, n2 = static_cast<uint64_t>(H) * H;
  if (e < n1) {
    const uint32_t o = static_cast<uint32_t>(e / Kp), k = static_cast<uint32_t>(e % Kp);
    w1b[e] = k < d ? f2bf(p[static_cast<size_t>(o) * d + k]) : 0;
  } else if (e < n1 + n2) {
    const uint64_t f = e - n1;
    const uint32_t o = static_cast<uint32_t>(f / H), i = static_cast<uint32_t>(f % H);
    const uint16_t b = f2bf(p[static_cast<size_t>(H) * d + H + f]);
    w2b[f] = b;
    if (w2tb) w2tb[static_cast<size_t>(i) * H + o] = b;
  }
}

// ---------------------------------------------------------------------------
// Fused transform (hidden width 256): one block per 64 rows keeps the bf16
// input tile and the first activation in LDS, runs both hidden layers on MFMA
// (wave w owns hidden units 64w .. 64w + 63, weights read from L2), and folds
// the FP32 output layer into the second layer's epilogue: per-wave partial
// outputs (lane butterfly over the 16 columns of a tile) are added over the
// four waves in order.  No intermediate buffer leaves the CU.
// ---------------------------------------------------------------------------
constexpr uint32_t kFusedH = 256;
// 16-row tiles per transform block: 2 (32 rows) keeps the accumulators at 32
// registers, so two waves share a SIMD
constexpr int kTransformTiles = 2;

template <int DIMS, int IT>
__global__ __launch_bounds__(256) void mlp_transform_kernel(const float *__restrict__ x, uint64_t n, uint32_t d,
                                                            uint32_t Kp, const uint16_t *__restrict__ W1b,
                                                            const float *__restrict__ b1,
                                                            const uint16_t *__restrict__ W2b,
                                                            const float *__restrict__ b2,
                                                            const float *__restrict__ W3,
                                                            const float *__restrict__ b3, float *__restrict__ Y) {
  constexpr uint32_t dims = DIMS;
  constexpr uint32_t H = kFusedH, LH = H + 8;
  constexpr uint32_t ROWS = 16 * IT;  // rows per block
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t LX = Kp + 8;
  uint16_t *sX = reinterpret_cast<uint16_t *>(smem);  // ROWS x LX
  uint16_t *sH = sX + ROWS * LX;  // ROWS x LH
  float *sY = reinterpret_cast<float *>(sH + ROWS * LH);  // 4 waves x ROWS rows x dims
  const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const uint64_t r0 = static_cast<uint64_t>(blockIdx.x) * ROWS;
  for (uint32_t e = t; e < ROWS * Kp; e += 256) {
    const uint32_t rr = e / Kp, k = e % Kp;
    sX[rr * LX + k] = (r0 + rr < n && k < d) ? f2bf(x[(r0 + rr) * d + k]) : 0;
  }
  __syncthreads();
  const uint32_t lr = lane & 15, lk = 8 * (lane >> 4), c0 = wave * 64;
  floatx4m acc[IT][4];
  auto layer = [&](const uint16_t *sA, uint32_t la, const uint16_t *W, uint32_t K) {
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = floatx4m{0.f, 0.f, 0.f, 0.f};
    for (uint32_t k0 = 0; k0 < K; k0 += 32) {
      bf16x8m a[IT], b[4];
#pragma unroll
      for (int i = 0; i < IT; ++i)
        a[i] = __builtin_bit_cast(bf16x8m, *reinterpret_cast<const uint4 *>(sA + (16 * i + lr) * la + k0 + lk));
#pragma unroll
      for (int j = 0; j < 4; ++j)
        b[j] = __builtin_bit_cast(bf16x8m, *reinterpret_cast<const uint4 *>(W + static_cast<size_t>(c0 + 16 * j + lr) * K + k0 + lk));
#pragma unroll
      for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };
  // layer 1 -> sH (bf16)
  layer(sX, LX, W1b, Kp);
#pragma unroll
  for (int i = 0; i < IT; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t col = c0 + 16 * j + lr;
      const float bias = b1[col];
#pragma unroll
      for (int r = 0; r < 4; ++r) sH[(16 * i + 4 * (lane >> 4) + r) * LH + col] = f2bf(gelu(acc[i][j][r] + bias));
    }
  __syncthreads();
  // layer 2 with the output layer folded in
  layer(sH, LH, W2b, H);
  float py[4 * IT][DIMS];
#pragma unroll
  for (int q = 0; q < 4 * IT; ++q)
#pragma unroll
    for (int c = 0; c < DIMS; ++c) py[q][c] = 0.0f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t col = c0 + 16 * j + lr;
    const float bias = b2[col];
    float w3[DIMS];
#pragma unroll
    for (int c = 0; c < DIMS; ++c) w3[c] = W3[c * H + col];
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float h = gelu(acc[i][j][r] + bias);
#pragma unroll
        for (int c = 0; c < DIMS; ++c) py[4 * i + r][c] = fmaf(h, w3[c], py[4 * i + r][c]);
      }
  }
#pragma unroll
  for (int q = 0; q < 4 * IT; ++q)
#pragma unroll
    for (int c = 0; c < DIMS; ++c) {
      float v = py[q][c];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      py[q][c] = v;
    }
  if (lr == 0) {
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < DIMS; ++c)
          sY[(wave * ROWS + 16 * i + 4 * (lane >> 4) + r) * dims + c] = py[4 * i + r][c];
  }
  __syncthreads();
  for (uint32_t e = t; e < ROWS * dims; e += 256) {
    const uint32_t rr = e / dims, c = e % dims;
    if (r0 + rr >= n) continue;
    float v = sY[(0 * ROWS + rr) * dims + c];
    v += sY[(1 * ROWS + rr) * dims + c];
    v += sY[(2 * ROWS + rr) * dims + c];
    v += sY[(3 * ROWS + rr) * dims + c];
    Y[(r0 + rr) * dims + c] = v + b3[c];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_mlp_gather(const float *x, uint32_t d, uint32_t Kp, const uint32_t *pi, const uint32_t *pj,
                             uint64_t P, uint64_t start, uint32_t count, uint64_t pa, uint64_t pc, uint32_t B,
                             uint32_t M, uint16_t *xb, uint16_t *xbT, uint32_t *pidx, hipStream_t s) {
  if (M % 64 || Kp % 8) return hipErrorInvalidValue;
  const size_t shm = static_cast<size_t>(64) * (Kp + 8) * sizeof(uint16_t);
  mlp_gather_kernel<<<M / 64, 256, shm, s>>>(x, d, Kp, pi, pj, P, start, count, pa, pc, B, M, xb, xbT, pidx);
  return hipGetLastError();
}

hipError_t launch_mlp_rows(const float *x, uint32_t d, uint32_t Kp, uint64_t row0, uint32_t count, uint32_t M,
                           uint16_t *xb, hipStream_t s) {
  mlp_rows_kernel<<<M, 128, 0, s>>>(x, d, Kp, row0, count, xb);
  return hipGetLastError();
}

static hipError_t gemm_check(uint32_t M, uint32_t N, uint32_t K, uint32_t kchunk) {
  if (M % 64 || N % 64 || K % 32 || kchunk % 32 || kchunk == 0) return hipErrorInvalidValue;
  return hipSuccess;
}

hipError_t launch_mlp_fwd(const uint16_t *A, uint32_t lda, const uint16_t *W, uint32_t ldw, uint32_t M,
                          uint32_t N, uint32_t K, const float *bias, float *z, uint16_t *hb, uint16_t *hT,
                          float *hf, hipStream_t s) {
  if (hipError_t e = gemm_check(M, N, K, K); e != hipSuccess) return e;
  MlpEpi ep;
  ep.bias = bias;
  ep.z = z;
  ep.hb = hb;
  ep.hT = hT;
  ep.hf = hf;
  mlp_gemm_kernel<EPI_FWD><<<dim3(M / 64, N / 64, 1), 256, 0, s>>>(A, lda, W, ldw, M, N, K, K, ep);
  return hipGetLastError();
}

hipError_t launch_mlp_dh(const uint16_t *dz, uint32_t ldd, const uint16_t *Wt, uint32_t ldw, uint32_t M,
                         uint32_t N, uint32_t K, const float *zin, uint16_t *dzT, hipStream_t s) {
  if (hipError_t e = gemm_check(M, N, K, K); e != hipSuccess) return e;
  MlpEpi ep;
  ep.zin = zin;
  ep.dzT = dzT;
  mlp_gemm_kernel<EPI_DH><<<dim3(M / 64, N / 64, 1), 256, 0, s>>>(dz, ldd, Wt, ldw, M, N, K, K, ep);
  return hipGetLastError();
}

hipError_t launch_mlp_wgrad(const uint16_t *AT, const uint16_t *BT, uint32_t Mo, uint32_t No, uint32_t K,
                            uint32_t S, float *part, float *out, uint32_t out_cols, uint32_t ldo, hipStream_t s) {
  const uint32_t kchunk = K / S;
  if (K % S) return hipErrorInvalidValue;
  if (hipError_t e = gemm_check(Mo, No, K, kchunk); e != hipSuccess) return e;
  MlpEpi ep;
  ep.part = part;
  mlp_gemm_kernel<EPI_PART><<<dim3(Mo / 64, No / 64, S), 256, 0, s>>>(AT, K, BT, K, Mo, No, K, kchunk, ep);
  const uint64_t total = static_cast<uint64_t>(Mo) * out_cols;
  mlp_sum_parts_kernel<<<static_cast<uint32_t>((total + 255) / 256), 256, 0, s>>>(
      part, S, Mo, out_cols, No, static_cast<uint64_t>(Mo) * No, out, ldo);
  return hipGetLastError();
}

hipError_t launch_mlp_out(const float *h2, uint32_t H, const float *W3, const float *b3, uint32_t dims,
                          uint32_t rows, float *Y, hipStream_t s) {
  if (H % 16 || dims > 8) return hipErrorInvalidValue;
  mlp_out_kernel<<<(rows + 63) / 64, 256, dims * H * sizeof(float), s>>>(h2, H, W3, b3, dims, rows, Y);
  return hipGetLastError();
}

hipError_t launch_mlp_pair_loss(const float *Y, uint32_t dims, uint32_t B, const uint32_t *pidx,
                                const uint32_t *pi, const uint32_t *pj, uint64_t nn_end, uint64_t mn_end,
                                const float w[3], float inv_count, float *dY, float *loss_part, hipStream_t s) {
  mlp_pair_loss_kernel<<<(B + 255) / 256, 256, 0, s>>>(Y, dims, B, pidx, pi, pj, nn_end, mn_end, w[0], w[1],
                                                       w[2], inv_count, dY, loss_part);
  return hipGetLastError();
}

hipError_t launch_mlp_out_bwd(const float *dY, uint32_t dims, const float *W3, const float *h2, const float *z2,
                              uint32_t H, uint32_t M, uint16_t *dz2b, uint16_t *dz2T, float *w3_part,
                              float *b3_part, hipStream_t s) {
  if (M % 64 || dims > 8) return hipErrorInvalidValue;
  mlp_out_bwd_kernel<<<M / 16, 256, 0, s>>>(dY, dims, W3, h2, z2, H, M, dz2b, dz2T, w3_part, b3_part);
  return hipGetLastError();
}

hipError_t launch_mlp_rowsum(const uint16_t *dzT, uint32_t rows, uint32_t M, float *out, hipStream_t s) {
  mlp_rowsum_kernel<<<rows, 256, 0, s>>>(dzT, M, out);
  return hipGetLastError();
}

hipError_t launch_mlp_sum_parts(const float *part, uint32_t S, uint32_t rows, uint32_t cols, uint32_t ldp,
                                uint64_t slab, float *out, uint32_t ldo, hipStream_t s) {
  const uint64_t total = static_cast<uint64_t>(rows) * cols;
  if (total == 0) return hipSuccess;
  mlp_sum_parts_kernel<<<static_cast<uint32_t>((total + 255) / 256), 256, 0, s>>>(part, S, rows, cols, ldp, slab,
                                                                                 out, ldo);
  return hipGetLastError();
}

// LDS of one fused transform block: the bf16 input tile, the bf16 first
// activation and the four waves' partial outputs.
static size_t mlp_transform_lds_bytes(uint32_t Kp, uint32_t dims) {
  constexpr uint32_t kRows = 16 * kTransformTiles;
  return static_cast<size_t>(kRows) * (Kp + 8) * 2 + static_cast<size_t>(kRows) * (kFusedH + 8) * 2 +
         static_cast<size_t>(4) * kRows * dims * sizeof(float);
}

bool mlp_transform_fused_ok(uint32_t H, uint32_t dims, uint32_t Kp) {
  return H == kFusedH && dims >= 1 && dims <= 8 && Kp % 32 == 0 && mlp_transform_lds_bytes(Kp, dims) <= kLdsBytes;
}

hipError_t launch_mlp_transform(const float *x, uint64_t n, uint32_t d, uint32_t Kp, const uint16_t *W1b,
                                const float *b1, const uint16_t *W2b, const float *b2, const float *W3,
                                const float *b3, uint32_t dims, float *Y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (!mlp_transform_fused_ok(kFusedH, dims, Kp)) return hipErrorInvalidValue;
  constexpr uint32_t kRows = 16 * kTransformTiles;
  const size_t shm = mlp_transform_lds_bytes(Kp, dims);
  const uint32_t g = static_cast<uint32_t>((n + kRows - 1) / kRows);
  switch (dims) {
#define PM_TR(D) \
  case D: mlp_transform_kernel<D, kTransformTiles><<<g, 256, shm, s>>>(x, n, d, Kp, W1b, b1, W2b, b2, W3, b3, Y); break;
    PM_TR(1) PM_TR(2) PM_TR(3) PM_TR(4) PM_TR(5) PM_TR(6) PM_TR(7) PM_TR(8)
#undef PM_TR
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mlp_tree_sum(const float *part, uint32_t S, uint64_t slab, uint32_t count, float *out,
                               hipStream_t s) {
  if (count == 0) return hipSuccess;
  mlp_tree_sum_kernel<<<count, 256, 0, s>>>(part, S, slab, out);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float *p, const float *g, float *m, float *v, uint64_t count, float lr_t, float b1,
                           float b2, float eps, hipStream_t s) {
  mlp_adam_kernel<<<static_cast<uint32_t>((count + 255) / 256), 256, 0, s>>>(p, g, m, v, count, lr_t, b1, b2, eps);
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const float *p, uint32_t d, uint32_t Kp, uint32_t H, uint16_t *w1b, uint16_t *w2b,
                           uint16_t *w2tb, hipStream_t s) {
  const uint64_t total = static_cast<uint64_t>(H) * Kp + static_cast<uint64_t>(H) * H;
  mlp_pack_kernel<<<static_cast<uint32_t>((total + 255) / 256), 256, 0, s>>>(p, d, Kp, H, w1b, w2b, w2tb);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Workspace of the allocation-free transform: the bf16 copies of W1 (H x Kp)
// and W2 (H x H) for H = 256 and Kp <= kNoAllocMaxKp, in a file-scope device
// array.  Each device holds its own copy of the array, at its own address,
// for the life of the process; the host looks that address up once per device
// and passes it to the kernels (the array's name is not a device address in
// host code).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocHalfs = static_cast<size_t>(kFusedH) * kNoAllocMaxKp + static_cast<size_t>(kFusedH) * kFusedH;
static_assert(kNoAllocHalfs * sizeof(uint16_t) == kNoAllocScratchBytes, "scratch size");
static __device__ __align__(256) uint16_t scratch[kNoAllocHalfs];

hipError_t mlp_noalloc_scratch(uint16_t **w1b, uint16_t **w2b) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[dev] = p;
  }
  *w1b = static_cast<uint16_t *>(addr[dev]);
  *w2b = *w1b + static_cast<size_t>(kFusedH) * kNoAllocMaxKp;
  return hipSuccess;
}

}  // namespace pm
