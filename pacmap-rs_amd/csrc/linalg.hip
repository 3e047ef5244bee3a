// linalg.hip -- small dense FP64 linear algebra for the PCA initialisation.
//
// The principal axes come from the FP64 covariance (d x d).  For d <= 128 a
// Jacobi eigensolver runs on it directly; for larger d a block subspace
// iteration (CholeskyQR2 re-orthonormalisation) narrows it to b <= 128
// directions and the Jacobi solver diagonalises the b x b Rayleigh-Ritz
// matrix.  Every step is a kernel on the caller's stream, so the whole
// eigen-step runs without host round trips.
#include "internal.h"

#include <cstdint>

namespace pm {

// ---------------------------------------------------------------------------
// C (M x N) = op(A) (M x K) * op(B) (K x N), row-major FP64, 32 x 32 tiles
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gemm64_kernel(uint32_t M, uint32_t N, uint32_t K,
                                                     const double *__restrict__ A, uint32_t lda, int tA,
                                                     const double *__restrict__ B, uint32_t ldb, int tB,
                                                     double *__restrict__ C, uint32_t ldc) {
  constexpr int TB = 32, KB = 32;
  __shared__ double sA[KB][TB + 1];
  __shared__ double sB[KB][TB + 1];
  const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
  const uint32_t r0 = blockIdx.x * TB, c0 = blockIdx.y * TB;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (uint32_t k0 = 0; k0 < K; k0 += KB) {
    __syncthreads();
    for (int li = t; li < TB * KB; li += 256) {
      const int a = li / KB, b = li % KB;  // tile row / k
      const uint32_t r = r0 + a, k = k0 + b;
      double v = 0.0;
      if (r < M && k < K) v = tA ? A[static_cast<size_t>(k) * lda + r] : A[static_cast<size_t>(r) * lda + k];
      sA[b][a] = v;
      const uint32_t kk = k0 + a, c = c0 + b;
      double w = 0.0;
      if (kk < K && c < N) w = tB ? B[static_cast<size_t>(c) * ldb + kk] : B[static_cast<size_t>(kk) * ldb + c];
      sB[a][b] = w;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < KB; ++kk) {
      const double a0 = sA[kk][ty * 2], a1 = sA[kk][ty * 2 + 1];
      const double b0 = sB[kk][tx * 2], b1 = sB[kk][tx * 2 + 1];
      acc[0][0] = fma(a0, b0, acc[0][0]);
      acc[0][1] = fma(a0, b1, acc[0][1]);
      acc[1][0] = fma(a1, b0, acc[1][0]);
      acc[1][1] = fma(a1, b1, acc[1][1]);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t r = r0 + ty * 2 + i, c = c0 + tx * 2 + j;
      if (r < M && c < N) C[static_cast<size_t>(r) * ldc + c] = acc[i][j];
    }
}

hipError_t launch_gemm64(uint32_t M, uint32_t N, uint32_t K, const double *A, uint32_t lda, bool tA,
                         const double *B, uint32_t ldb, bool tB, double *C, uint32_t ldc, hipStream_t s) {
  if (M == 0 || N == 0) return hipSuccess;
  gemm64_kernel<<<dim3((M + 31) / 32, (N + 31) / 32), 256, 0, s>>>(M, N, K, A, lda, tA ? 1 : 0, B, ldb,
                                                                   tB ? 1 : 0, C, ldc);
  return hipGetLastError();
}

constexpr int kSmallMax = 128;  // largest matrix the single-workgroup kernels hold in LDS

// ---------------------------------------------------------------------------
// G (b x b, SPD) = R^T R; writes R^{-1} (upper triangular).  A failed pivot
// retries with a growing diagonal shift; status[0] = 1 if every try fails.
// One workgroup; G is held in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void chol_inv_kernel(const double *__restrict__ G, uint32_t b,
                                                        double *__restrict__ Rinv,
                                                        uint32_t *__restrict__ status) {
  __shared__ double A[kSmallMax * kSmallMax];
  __shared__ double shift;
  __shared__ int fail;
  const int t = threadIdx.x;
  double tr = 0.0;
  for (uint32_t j = 0; j < b; ++j) tr += G[static_cast<size_t>(j) * b + j];
  if (t == 0) shift = 0.0;
  for (int attempt = 0; attempt < 6; ++attempt) {
    __syncthreads();
    for (uint32_t e = t; e < b * b; e += blockDim.x) {
      const uint32_t r = e / b, c = e % b;
      A[r * kSmallMax + c] = G[e] + (r == c ? shift : 0.0);
    }
    if (t == 0) fail = 0;
    __syncthreads();
    for (uint32_t j = 0; j < b; ++j) {  // right-looking, upper factor in rows
      if (t == 0) {
        const double dj = A[j * kSmallMax + j];
        if (!(dj > 0.0)) fail = 1;
        A[j * kSmallMax + j] = dj > 0.0 ? sqrt(dj) : 1.0;
      }
      __syncthreads();
      if (fail) break;
      const double rjj = A[j * kSmallMax + j];
      for (uint32_t i = j + 1 + t; i < b; i += blockDim.x) A[j * kSmallMax + i] /= rjj;
      __syncthreads();
      const uint32_t m = b - j - 1;
      for (uint32_t e = t; e < m * m; e += blockDim.x) {
        const uint32_t k = j + 1 + e / m, i = j + 1 + e % m;
        if (i >= k) A[k * kSmallMax + i] -= A[j * kSmallMax + k] * A[j * kSmallMax + i];
      }
      __syncthreads();
    }
    if (!fail) break;
    if (t == 0) shift = shift == 0.0 ? 1e-12 * tr : shift * 100.0;
    __syncthreads();
    if (attempt == 5 && t == 0) status[0] = 1;
  }
  __syncthreads();
  // back substitution, one thread per column c of R^{-1}
  for (uint32_t c = t; c < b; c += blockDim.x) {
    for (uint32_t i = 0; i < b; ++i)
      if (i > c) Rinv[static_cast<size_t>(i) * b + c] = 0.0;
    Rinv[static_cast<size_t>(c) * b + c] = 1.0 / A[c * kSmallMax + c];
    for (int i = static_cast<int>(c) - 1; i >= 0; --i) {
      double acc = 0.0;
      for (uint32_t k = i + 1; k <= c; ++k) acc += A[i * kSmallMax + k] * Rinv[static_cast<size_t>(k) * b + c];
      Rinv[static_cast<size_t>(i) * b + c] = -acc / A[i * kSmallMax + i];
    }
  }
}

// ---------------------------------------------------------------------------
// Jacobi eigensolver for a symmetric m x m matrix (m <= 128), one workgroup.
// Parallel (round-robin) ordering: each round applies m/2 disjoint plane
// rotations at once; m - 1 rounds form a sweep.  Output: the q eigenvectors
// of largest eigenvalue as columns of Vout (m x q), sign chosen so the
// largest-magnitude entry is positive.  Vwork: m x m scratch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void jacobi_kernel(const double *__restrict__ Ain, uint32_t m,
                                                      uint32_t q, double *__restrict__ Vwork,
                                                      double *__restrict__ Vout) {
  __shared__ double A[kSmallMax * kSmallMax];
  __shared__ double cs[kSmallMax / 2], sn[kSmallMax / 2];
  __shared__ int pp[kSmallMax / 2], qq[kSmallMax / 2];
  __shared__ double red[2][1024];
  __shared__ int rank[kSmallMax];
  const int t = threadIdx.x;
  const uint32_t M = (m + 1) & ~1u;  // even count; index m is a dummy when m is odd
  const uint32_t half = M / 2;
  for (uint32_t e = t; e < m * m; e += blockDim.x) {
    const uint32_t r = e / m, c = e % m;
    A[r * kSmallMax + c] = 0.5 * (Ain[e] + Ain[static_cast<size_t>(c) * m + r]);  // symmetrise
    Vwork[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (uint32_t e = t; e < m * m; e += blockDim.x) {
      const uint32_t r = e / m, c = e % m;
      const double v = A[r * kSmallMax + c];
      if (r == c) dg += v * v;
      else if (r < c) off += v * v;
    }
    red[0][t] = off;
    red[1][t] = dg;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
      if (t < w) {
        red[0][t] += red[0][t + w];
        red[1][t] += red[1][t + w];
      }
      __syncthreads();
    }
    const double offs = red[0][0], dgs = red[1][0];
    __syncthreads();
    if (offs <= 1e-30 * dgs || offs == 0.0) break;
    for (uint32_t round = 0; round + 1 < M; ++round) {
      if (static_cast<uint32_t>(t) < half) {
        auto pos = [&](uint32_t j) -> uint32_t { return j == 0 ? 0 : 1 + (j - 1 + round) % (M - 1); };
        uint32_t p = pos(t), r2 = pos(M - 1 - t);
        if (p > r2) {
          const uint32_t tmp = p;
          p = r2;
          r2 = tmp;
        }
        double c = 1.0, s = 0.0;
        if (r2 < m) {
          const double apq = A[p * kSmallMax + r2];
          if (apq != 0.0) {
            const double theta = (A[r2 * kSmallMax + r2] - A[p * kSmallMax + p]) / (2.0 * apq);
            const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            s = tt * c;
          }
        }
        pp[t] = static_cast<int>(p);
        qq[t] = static_cast<int>(r2);
        cs[t] = c;
        sn[t] = s;
      }
      __syncthreads();
      // columns: A <- A J
      for (uint32_t e = t; e < m * half; e += blockDim.x) {
        const uint32_t k = e / half, i = e % half;
        const int p = pp[i], r2 = qq[i];
        if (static_cast<uint32_t>(r2) >= m) continue;
        const double akp = A[k * kSmallMax + p], akq = A[k * kSmallMax + r2];
        A[k * kSmallMax + p] = cs[i] * akp - sn[i] * akq;
        A[k * kSmallMax + r2] = sn[i] * akp + cs[i] * akq;
      }
      __syncthreads();
      // rows: A <- J^T A; eigenvectors: V <- V J
      for (uint32_t e = t; e < m * half; e += blockDim.x) {
        const uint32_t k = e / half, i = e % half;
        const int p = pp[i], r2 = qq[i];
        if (static_cast<uint32_t>(r2) >= m) continue;
        const double apk = A[p * kSmallMax + k], aqk = A[r2 * kSmallMax + k];
        A[p * kSmallMax + k] = cs[i] * apk - sn[i] * aqk;
        A[r2 * kSmallMax + k] = sn[i] * apk + cs[i] * aqk;
        double &vkp = Vwork[static_cast<size_t>(k) * m + p], &vkq = Vwork[static_cast<size_t>(k) * m + r2];
        const double x = vkp, y = vkq;
        vkp = cs[i] * x - sn[i] * y;
        vkq = sn[i] * x + cs[i] * y;
      }
      __syncthreads();
      if (static_cast<uint32_t>(t) < half && static_cast<uint32_t>(qq[t]) < m) {
        A[pp[t] * kSmallMax + qq[t]] = 0.0;
        A[qq[t] * kSmallMax + pp[t]] = 0.0;
      }
      __syncthreads();
    }
  }
  // rank by eigenvalue (descending, ties by index)
  for (uint32_t j = t; j < m; j += blockDim.x) {
    const double wj = A[j * kSmallMax + j];
    int r = 0;
    for (uint32_t i = 0; i < m; ++i) {
      const double wi = A[i * kSmallMax + i];
      r += (wi > wj || (wi == wj && i < j)) ? 1 : 0;
    }
    rank[j] = r;
  }
  __syncthreads();
  for (uint32_t j = t; j < m; j += blockDim.x) {
    const uint32_t col = static_cast<uint32_t>(rank[j]);
    if (col >= q) continue;
    double best = -1.0, sg = 1.0;
    for (uint32_t k = 0; k < m; ++k) {
      const double v = Vwork[static_cast<size_t>(k) * m + j];
      if (fabs(v) > best) {
        best = fabs(v);
        sg = v < 0 ? -1.0 : 1.0;
      }
    }
    for (uint32_t k = 0; k < m; ++k) Vout[static_cast<size_t>(k) * q + col] = sg * Vwork[static_cast<size_t>(k) * m + j];
  }
}

// Column signs of V (d x q): largest-magnitude entry positive; also writes
// the FP32 copy used by the projection kernel.
__global__ void sign_fix_kernel(double *__restrict__ V, uint32_t d, uint32_t q, float *__restrict__ Vf) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= q) return;
  double best = -1.0, sg = 1.0;
  for (uint32_t r = 0; r < d; ++r) {
    const double v = V[static_cast<size_t>(r) * q + c];
    if (fabs(v) > best) {
      best = fabs(v);
      sg = v < 0 ? -1.0 : 1.0;
    }
  }
  for (uint32_t r = 0; r < d; ++r) {
    const double v = sg * V[static_cast<size_t>(r) * q + c];
    V[static_cast<size_t>(r) * q + c] = v;
    Vf[static_cast<size_t>(r) * q + c] = static_cast<float>(v);
  }
}

hipError_t launch_chol_inv(const double *G, uint32_t b, double *Rinv, uint32_t *status, hipStream_t s) {
  if (b > kSmallMax) return hipErrorInvalidValue;
  chol_inv_kernel<<<1, 1024, 0, s>>>(G, b, Rinv, status);
  return hipGetLastError();
}

hipError_t launch_jacobi(const double *A, uint32_t m, uint32_t q, double *Vwork, double *Vout, hipStream_t s) {
  if (m > kSmallMax || q > m) return hipErrorInvalidValue;
  jacobi_kernel<<<1, 1024, 0, s>>>(A, m, q, Vwork, Vout);
  return hipGetLastError();
}

hipError_t launch_sign_fix(double *V, uint32_t d, uint32_t q, float *Vf, hipStream_t s) {
  sign_fix_kernel<<<(q + 63) / 64, 64, 0, s>>>(V, d, q, Vf);
  return hipGetLastError();
}

}  // namespace pm
