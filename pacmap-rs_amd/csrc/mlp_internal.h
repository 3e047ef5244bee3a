// mlp_internal.h -- launchers of the encoder kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pm {

// batch rows from the typed pair list (first endpoints, then second
// endpoints); pidx[p] = pair index of batch pair p (0xffffffff: padding)
hipError_t launch_mlp_gather(const float *x, uint32_t d, uint32_t Kp, const uint32_t *pi, const uint32_t *pj,
                             uint64_t P, uint64_t start, uint32_t count, uint64_t pa, uint64_t pc, uint32_t B,
                             uint32_t M, uint16_t *xb, uint16_t *xbT, uint32_t *pidx, hipStream_t s);
// consecutive rows [row0, row0 + count) of x, zero padded to M rows
hipError_t launch_mlp_rows(const float *x, uint32_t d, uint32_t Kp, uint64_t row0, uint32_t count, uint32_t M,
                           uint16_t *xb, hipStream_t s);
// hidden layer: z = A W^T + bias, h = gelu(z); outputs optional
hipError_t launch_mlp_fwd(const uint16_t *A, uint32_t lda, const uint16_t *W, uint32_t ldw, uint32_t M,
                          uint32_t N, uint32_t K, const float *bias, float *z, uint16_t *hb, uint16_t *hT,
                          float *hf, hipStream_t s);
// dz (transposed, bf16) = (dz_above W_above) * gelu'(z)
hipError_t launch_mlp_dh(const uint16_t *dz, uint32_t ldd, const uint16_t *Wt, uint32_t ldw, uint32_t M,
                         uint32_t N, uint32_t K, const float *zin, uint16_t *dzT, hipStream_t s);
// out (Mo x out_cols, row stride ldo) = AT BT^T summed over K in S split-K slabs
hipError_t launch_mlp_wgrad(const uint16_t *AT, const uint16_t *BT, uint32_t Mo, uint32_t No, uint32_t K,
                            uint32_t S, float *part, float *out, uint32_t out_cols, uint32_t ldo, hipStream_t s);
hipError_t launch_mlp_out(const float *h2, uint32_t H, const float *W3, const float *b3, uint32_t dims,
                          uint32_t rows, float *Y, hipStream_t s);
hipError_t launch_mlp_pair_loss(const float *Y, uint32_t dims, uint32_t B, const uint32_t *pidx,
                                const uint32_t *pi, const uint32_t *pj, uint64_t nn_end, uint64_t mn_end,
                                const float w[3], float inv_count, float *dY, float *loss_part, hipStream_t s);
// per 16-row block: dz2 (bf16, row-major and transposed) and the partials of
// dW3 (M/16 slabs of dims x H) and db3 (M/16 slabs of dims)
hipError_t launch_mlp_out_bwd(const float *dY, uint32_t dims, const float *W3, const float *h2, const float *z2,
                              uint32_t H, uint32_t M, uint16_t *dz2b, uint16_t *dz2T, float *w3_part,
                              float *b3_part, hipStream_t s);
hipError_t launch_mlp_rowsum(const uint16_t *dzT, uint32_t rows, uint32_t M, float *out, hipStream_t s);
hipError_t launch_mlp_sum_parts(const float *part, uint32_t S, uint32_t rows, uint32_t cols, uint32_t ldp,
                                uint64_t slab, float *out, uint32_t ldo, hipStream_t s);
// LDS per workgroup on gfx950
constexpr size_t kLdsBytes = 160 * 1024;
// Fused transform Y = encoder(x) for hidden width 256 when its block's LDS
// (input tile of Kp bf16 columns, first activation, partial outputs) fits
bool mlp_transform_fused_ok(uint32_t H, uint32_t dims, uint32_t Kp);
hipError_t launch_mlp_transform(const float *x, uint64_t n, uint32_t d, uint32_t Kp, const uint16_t *W1b,
                                const float *b1, const uint16_t *W2b, const float *b2, const float *W3,
                                const float *b3, uint32_t dims, float *Y, hipStream_t s);
// out[o] = sum over z < S of part[z * slab + o], o < count (fixed-order tree)
hipError_t launch_mlp_tree_sum(const float *part, uint32_t S, uint64_t slab, uint32_t count, float *out,
                               hipStream_t s);
hipError_t launch_mlp_adam(float *p, const float *g, float *m, float *v, uint64_t count, float lr_t, float b1,
                           float b2, float eps, hipStream_t s);
// bf16 copies of W1 (H x Kp, zero padded), W2 and (w2tb non-null) W2^T
hipError_t launch_mlp_pack(const float *p, uint32_t d, uint32_t Kp, uint32_t H, uint16_t *w1b, uint16_t *w2b,
                           uint16_t *w2tb, hipStream_t s);
// Allocation-free transform: the bf16 W1 (256 x Kp, Kp <= kNoAllocMaxKp) and
// W2 (256 x 256) live in a static device array of kNoAllocScratchBytes per
// device; this returns the current device's addresses of both parts.
constexpr uint32_t kNoAllocMaxKp = 2048;
constexpr size_t kNoAllocScratchBytes = (static_cast<size_t>(256) * kNoAllocMaxKp + 256 * 256) * 2;
hipError_t mlp_noalloc_scratch(uint16_t **w1b, uint16_t **w2b);

}  // namespace pm
