// linalg.hip -- small dense FP64 linear algebra for the PCA initialisation.
//
// The principal axes come from the FP64 covariance (d x d).  For d <= 112 a
// Jacobi eigensolver runs on it directly; for larger d a block subspace
// iteration (CholeskyQR2 re-orthonormalisation) narrows it to b <= 112
// directions and the Jacobi solver diagonalises the b x b Rayleigh-Ritz
// matrix.  Every step is a kernel on the caller's stream, so the whole
// eigen-step runs without host round trips.
#include "internal.h"

#include <cstdint>

namespace pm {

// ---------------------------------------------------------------------------
// C (M x N) = op(A) (M x K) * op(B) (K x N), row-major FP64, 32 x 32 tiles.
// blockIdx.z takes the K range [z * kchunk, (z + 1) * kchunk) and writes the
// z-th M x ldc slab of C (split-K partials, summed by sum_parts_kernel).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gemm64_kernel(uint32_t M, uint32_t N, uint32_t K,
                                                     const double *__restrict__ A, uint32_t lda, int tA,
                                                     const double *__restrict__ B, uint32_t ldb, int tB,
                                                     double *__restrict__ C, uint32_t ldc, uint32_t kchunk) {
  constexpr int TB = 32, KB = 32;
  __shared__ double sA[KB][TB + 1];
  __shared__ double sB[KB][TB + 1];
  const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
  const uint32_t r0 = blockIdx.x * TB, c0 = blockIdx.y * TB;
  const uint32_t kbeg = blockIdx.z * kchunk;
  const uint32_t kend = kbeg + kchunk < K ? kbeg + kchunk : K;
  C += static_cast<size_t>(blockIdx.z) * M * ldc;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (uint32_t k0 = kbeg; k0 < kend; k0 += KB) {
    __syncthreads();
    for (int li = t; li < TB * KB; li += 256) {
      const int a = li / KB, b = li % KB;  // tile row / k
      const uint32_t r = r0 + a, k = k0 + b;
      double v = 0.0;
      if (r < M && k < kend) v = tA ? A[static_cast<size_t>(k) * lda + r] : A[static_cast<size_t>(r) * lda + k];
      sA[b][a] = v;
      const uint32_t kk = k0 + a, c = c0 + b;
      double w = 0.0;
      if (kk < kend && c < N) w = tB ? B[static_cast<size_t>(c) * ldb + kk] : B[static_cast<size_t>(kk) * ldb + c];
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

// C[r][c] = sum over z = 0..ks-1 (in order) of part[z][r][c]
__global__ void sum_parts_kernel(const double *__restrict__ part, uint32_t ks, uint32_t M, uint32_t N,
                                 double *__restrict__ C, uint32_t ldc) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t mn = static_cast<uint64_t>(M) * N;
  if (e >= mn) return;
  double acc = part[e];
  for (uint32_t z = 1; z < ks; ++z) acc += part[z * mn + e];
  C[(e / N) * ldc + e % N] = acc;
}

hipError_t launch_gemm64(uint32_t M, uint32_t N, uint32_t K, const double *A, uint32_t lda, bool tA,
                         const double *B, uint32_t ldb, bool tB, double *C, uint32_t ldc, hipStream_t s,
                         double *part, uint64_t part_cap) {
  if (M == 0 || N == 0) return hipSuccess;
  const uint32_t tiles = ((M + 31) / 32) * ((N + 31) / 32);
  uint64_t ks = (256 + tiles - 1) / tiles;
  ks = ks < (K + 63) / 64 ? ks : (K + 63) / 64;
  const uint64_t mn = static_cast<uint64_t>(M) * N;
  if (part == nullptr || part_cap < 2 * mn) ks = 1;
  else if (ks * mn > part_cap) ks = part_cap / mn;
  if (ks < 1) ks = 1;
  const uint32_t kchunk = static_cast<uint32_t>((((K + ks - 1) / ks) + 31) / 32 * 32);
  ks = (K + kchunk - 1) / kchunk;  // no empty slabs
  if (ks < 1) ks = 1;
  const dim3 grid((M + 31) / 32, (N + 31) / 32, static_cast<uint32_t>(ks));
  if (ks == 1) {
    gemm64_kernel<<<grid, 256, 0, s>>>(M, N, K, A, lda, tA ? 1 : 0, B, ldb, tB ? 1 : 0, C, ldc,
                                       K > 0 ? K : 1);
    return hipGetLastError();
  }
  gemm64_kernel<<<grid, 256, 0, s>>>(M, N, K, A, lda, tA ? 1 : 0, B, ldb, tB ? 1 : 0, part, N, kchunk);
  sum_parts_kernel<<<static_cast<uint32_t>((mn + 255) / 256), 256, 0, s>>>(part, static_cast<uint32_t>(ks), M, N,
                                                                           C, ldc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// G (b x b, SPD, b <= 128) = R^T R; writes R^{-1} (upper triangular, b x b).
//
// Elimination on the augmented matrix [G | I]: step j scales row j by
// 1/sqrt(pivot) and subtracts f_r * (scaled row j) from every row r > j,
// where f_r is entry r of the scaled row (the Schur complement is
// symmetric).  The left half ends as R, the right half as E with E G = R,
// i.e. E = R^{-T}; R^{-1} is its transpose.  The 128 x 256 array lives in
// registers: thread (rg, cg) of a 16 x 32 grid holds rows rg + 16 a and
// columns cg + 32 e (a, e < 8); the pivot row goes through a double-buffered
// LDS row, one barrier per step.  Row j is final after step j, so its right
// half is written out then.  A non-positive pivot restarts with a diagonal
// shift (1e-12 trace, then x100); status[0] = 1 if six tries fail.
// ---------------------------------------------------------------------------
constexpr int kCholMax = 128;

__global__ __launch_bounds__(512) void chol_inv_kernel(const double *__restrict__ G, uint32_t b,
                                                       double *__restrict__ Rinv,
                                                       uint32_t *__restrict__ status) {
  __shared__ double row[2][2 * kCholMax];
  const int t = threadIdx.x, rg = t >> 5, cg = t & 31;
  double tr = 0.0;
  for (uint32_t j = 0; j < b; ++j) tr += G[static_cast<size_t>(j) * b + j];
  double shift = 0.0;
  double M[8][8];
  bool ok = false;
  for (int attempt = 0; attempt < 6 && !ok; ++attempt) {
    __syncthreads();  // every thread has left the previous attempt's reads
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const uint32_t r = rg + 16 * a;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const uint32_t c = cg + 32 * e;
        double v = 0.0;
        if (r < b) {
          if (c < kCholMax) {
            if (c < b) v = G[static_cast<size_t>(r) * b + c] + (r == c ? shift : 0.0);
          } else {
            v = (c - kCholMax == r) ? 1.0 : 0.0;
          }
        }
        M[a][e] = v;
      }
    }
    ok = true;
#pragma unroll
    for (int aj = 0; aj < 8; ++aj) {
      for (uint32_t jj = 0; jj < 16; ++jj) {
        const uint32_t j = 16 * aj + jj;
        if (j >= b) break;
        double *rb = row[j & 1];
        if (static_cast<uint32_t>(rg) == jj) {
#pragma unroll
          for (int e = 0; e < 8; ++e) rb[cg + 32 * e] = M[aj][e];
        }
        __syncthreads();
        const double p = rb[j];
        if (!(p > 0.0)) {  // same value in every thread: a uniform exit
          ok = false;
          break;
        }
        const double inv = 1.0 / sqrt(p);
        double rc[8], f[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) rc[e] = rb[cg + 32 * e] * inv;
        // row j is final once scaled (later steps only change rows below
        // it): its right half goes out now, and rows <= j take f = 0 (the
        // fma then leaves them as they are)
        if (static_cast<uint32_t>(rg) == jj) {
#pragma unroll
          for (int e = 4; e < 8; ++e) {
            const uint32_t c = cg + 32 * e - kCholMax;
            if (c < b) Rinv[static_cast<size_t>(c) * b + j] = rc[e];
          }
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          const uint32_t r = rg + 16 * a;
          f[a] = r > j ? rb[r] * inv : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
          for (int e = 0; e < 8; ++e) M[a][e] = fma(-f[a], rc[e], M[a][e]);
      }
      if (!ok) break;
    }
    if (!ok) shift = shift == 0.0 ? 1e-12 * tr : shift * 100.0;
  }
  if (!ok && t == 0) status[0] = 1;
}

// ---------------------------------------------------------------------------
// Jacobi eigensolver for a symmetric m x m matrix (m <= 112).
//
// Rotations (one workgroup): the matrix is kept in LDS as its packed upper
// triangle.  Parallel (round-robin) ordering: each round applies the
// h = ceil(m/2) disjoint rotations of that round at once, as J^T A J: for
// every two rotation pairs i < j the 2 x 2 block of A between them is
// updated with both rotations, and each pair's own 2 x 2 diagonal block with
// the closed form (app - t apq, aqq + t apq, 0).  The per-thread work list is
// the same every round, so it is decoded once.  Sweeps stop when the
// off-diagonal square sum is below 1e-24 of the diagonal one (off-diagonal
// norm 1e-12 relative).  Every round's rotations (p, q, c, s) are recorded.
//
// Eigenvectors (one wave per component): V^T starts as the identity and every
// recorded rotation turns its rows p and q.  Column k of V^T only ever mixes
// with itself, so each wave replays all rounds on one column held in LDS,
// with the rounds' records loaded ahead.  The rotation workgroup is left
// with A alone, which takes 1/5 of the LDS traffic of updating both.
//
// Output: the q eigenvectors of largest eigenvalue (ties by index) as columns
// of Vout (m x q), each with its largest-magnitude entry positive.
// ---------------------------------------------------------------------------
constexpr int kJacMax = 112;
constexpr int kJacPairs = kJacMax / 2;
constexpr int kJacThreads = 1024;
constexpr int kJacSlots = 2;  // work slots per thread per round
constexpr int kJacSweeps = 40;
static_assert(kJacPairs * (kJacPairs - 1) / 2 + kJacPairs <= kJacThreads * kJacSlots,
              "Jacobi work list does not fit the slots");

__device__ __forceinline__ uint32_t packed_idx(uint32_t r, uint32_t c, uint32_t m) {
  if (r > c) {
    const uint32_t x = r;
    r = c;
    c = x;
  }
  return r * m - (r * (r - 1)) / 2 + (c - r);  // row r of the upper triangle starts at r*m - r(r-1)/2
}

struct JacobiWork {
  double2 *cs;        // rounds x h rotations (c, s)
  uint32_t *pq;       // rounds x h: p | q << 16 (q >= m: the dummy index, no rotation)
  double *evals;      // m
  double *vt;         // m x m, V^T
  uint32_t *rounds;   // 1
};

__global__ __launch_bounds__(kJacThreads) void jacobi_rot_kernel(const double *__restrict__ Ain, uint32_t m,
                                                                 JacobiWork w) {
  __shared__ double Ap[kJacMax * (kJacMax + 1) / 2];
  __shared__ double cs[kJacPairs], sn[kJacPairs], tn[kJacPairs];
  __shared__ int pp[kJacPairs], qq[kJacPairs];
  __shared__ double red[2][kJacThreads / 64];
  const int t = threadIdx.x;
  const uint32_t M = (m + 1) & ~1u;  // even count; index m is a dummy when m is odd
  const uint32_t h = M / 2;
  for (uint32_t e = t; e < m * m; e += kJacThreads) {
    const uint32_t r = e / m, c = e % m;
    if (r <= c) Ap[packed_idx(r, c, m)] = 0.5 * (Ain[e] + Ain[static_cast<size_t>(c) * m + r]);
  }
  // Work list: blocks (i < j) enumerated as (x, (x + s) mod h), then the h
  // diagonal blocks.  Code: type << 24 | a << 12 | b.
  const uint32_t hs = h / 2;
  const uint32_t full = (h & 1) ? hs : (hs > 0 ? hs - 1 : 0);
  const uint32_t nb = h * (h - 1) / 2;
  const uint32_t total = nb + h;
  uint32_t slot[kJacSlots];
#pragma unroll
  for (int k = 0; k < kJacSlots; ++k) {
    const uint32_t wi = t + k * kJacThreads;
    uint32_t code = 0xffffffffu;
    if (wi < nb) {
      uint32_t sft, x;
      if (wi < full * h) {
        sft = 1 + wi / h;
        x = wi % h;
      } else {
        sft = hs;
        x = wi - full * h;
      }
      uint32_t i = x, j = (x + sft) % h;
      if (i > j) {
        const uint32_t y = i;
        i = j;
        j = y;
      }
      code = (0u << 24) | (i << 12) | j;
    } else if (wi < total) {
      code = (1u << 24) | ((wi - nb) << 12);
    }
    slot[k] = code;
  }
  __syncthreads();
  uint32_t grnd = 0;  // rounds recorded
  for (int sweep = 0; sweep < kJacSweeps; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (uint32_t r = 0; r < m; ++r) {  // thread t takes columns c = t, t + 1024, ... of every row
      const uint32_t base = r * m - (r * (r - 1)) / 2;
      for (uint32_t c = r + t; c < m; c += kJacThreads) {
        const double v = Ap[base + (c - r)];
        if (c == r) dg += v * v;
        else off += v * v;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o);
      dg += __shfl_xor(dg, o);
    }
    if ((t & 63) == 0) {
      red[0][t >> 6] = off;
      red[1][t >> 6] = dg;
    }
    __syncthreads();
    double offs = 0.0, dgs = 0.0;
    for (int wv = 0; wv < kJacThreads / 64; ++wv) {
      offs += red[0][wv];
      dgs += red[1][wv];
    }
    if (offs == 0.0 || offs <= 1e-24 * dgs) break;  // uniform: every thread summed the same values
    for (uint32_t round = 0; round + 1 < M; ++round, ++grnd) {
      if (static_cast<uint32_t>(t) < h) {
        auto pos = [&](uint32_t j) -> uint32_t { return j == 0 ? 0 : 1 + (j - 1 + round) % (M - 1); };
        uint32_t p = pos(t), r2 = pos(M - 1 - t);
        if (p > r2) {
          const uint32_t tmp = p;
          p = r2;
          r2 = tmp;
        }
        double c = 1.0, s = 0.0, tt = 0.0;
        if (r2 < m) {
          const double apq = Ap[packed_idx(p, r2, m)];
          if (apq != 0.0) {
            // tan of the rotation angle, the smaller root of t^2 + 2 theta t - 1 = 0
            // with theta = (aqq - app) / (2 apq), written as 2 apq sgn(dd) / (|dd| +
            // sqrt(dd^2 + 4 apq^2)) to save a division
            const double dd = Ap[packed_idx(r2, r2, m)] - Ap[packed_idx(p, p, m)];
            const double a2 = 2.0 * apq;
            const double den = fabs(dd) + sqrt(fma(dd, dd, a2 * a2));
            tt = (dd >= 0.0 ? a2 : -a2) / den;
            c = rsqrt(fma(tt, tt, 1.0));
            s = tt * c;
          }
        }
        pp[t] = static_cast<int>(p);
        qq[t] = static_cast<int>(r2);
        cs[t] = c;
        sn[t] = s;
        tn[t] = tt;
        const size_t rec = static_cast<size_t>(grnd) * h + t;
        w.cs[rec] = make_double2(c, s);
        w.pq[rec] = p | (r2 << 16);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kJacSlots; ++k) {
        const uint32_t code = slot[k];
        if (code == 0xffffffffu) continue;
        const uint32_t type = code >> 24, a = (code >> 12) & 0xfff, bb = code & 0xfff;
        if (type == 0) {
          const uint32_t p1 = pp[a], q1 = qq[a], p2 = pp[bb], q2 = qq[bb];
          const double c1 = cs[a], s1 = sn[a], c2 = cs[bb], s2 = sn[bb];
          const bool h1 = q1 < m, h2 = q2 < m;  // false: the pair holds the dummy index
          const uint32_t i00 = packed_idx(p1, p2, m);
          const uint32_t i01 = h2 ? packed_idx(p1, q2, m) : 0;
          const uint32_t i10 = h1 ? packed_idx(q1, p2, m) : 0;
          const uint32_t i11 = (h1 && h2) ? packed_idx(q1, q2, m) : 0;
          const double b00 = Ap[i00], b01 = h2 ? Ap[i01] : 0.0;
          const double b10 = h1 ? Ap[i10] : 0.0, b11 = (h1 && h2) ? Ap[i11] : 0.0;
          // rows (rotation a), then columns (rotation bb)
          const double r00 = c1 * b00 - s1 * b10, r01 = c1 * b01 - s1 * b11;
          const double r10 = s1 * b00 + c1 * b10, r11 = s1 * b01 + c1 * b11;
          Ap[i00] = c2 * r00 - s2 * r01;
          if (h2) Ap[i01] = s2 * r00 + c2 * r01;
          if (h1) Ap[i10] = c2 * r10 - s2 * r11;
          if (h1 && h2) Ap[i11] = s2 * r10 + c2 * r11;
        } else {
          const uint32_t p = pp[a], r2 = qq[a];
          const double tt = tn[a];
          if (r2 < m && tt != 0.0) {
            const uint32_t ipq = packed_idx(p, r2, m), ipp = packed_idx(p, p, m), iqq = packed_idx(r2, r2, m);
            const double apq = Ap[ipq];
            Ap[ipp] = Ap[ipp] - tt * apq;
            Ap[iqq] = Ap[iqq] + tt * apq;
            Ap[ipq] = 0.0;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t j = t; j < m; j += kJacThreads) w.evals[j] = Ap[packed_idx(j, j, m)];
  if (t == 0) w.rounds[0] = grnd;
}

// One wave per column k of V^T: replays every recorded round on it.
constexpr int kJacChunk = 8;  // rounds whose records are loaded ahead
__global__ __launch_bounds__(64) void jacobi_vec_kernel(uint32_t m, JacobiWork w) {
  __shared__ double x[kJacMax + 1];
  const uint32_t k = blockIdx.x, lane = threadIdx.x;
  const uint32_t M = (m + 1) & ~1u, h = M / 2;
  for (uint32_t j = lane; j <= m; j += 64) x[j] = (j == k) ? 1.0 : 0.0;
  const uint32_t R = w.rounds[0];
  const uint32_t last = R * h;  // records
  const bool act = lane < h;
  struct Chunk {
    double2 cs[kJacChunk];
    uint32_t pq[kJacChunk];
  };
  auto load = [&](uint32_t r0, Chunk &ch) {
#pragma unroll
    for (int i = 0; i < kJacChunk; ++i) {
      const uint32_t rec = min((r0 + i) * h + lane, last ? last - 1 : 0u);  // clamped; unused past R
      ch.cs[i] = w.cs[rec];
      ch.pq[i] = w.pq[rec];
    }
  };
  auto run = [&](uint32_t r0, const Chunk &ch) {
#pragma unroll
    for (int i = 0; i < kJacChunk; ++i) {
      if (r0 + i >= R) break;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (act) {
        const uint32_t p = ch.pq[i] & 0xffffu, q = ch.pq[i] >> 16;
        if (q < m) {
          const double c = ch.cs[i].x, s = ch.cs[i].y;
          const double xp = x[p], xq = x[q];
          x[p] = c * xp - s * xq;
          x[q] = s * xp + c * xq;
        }
      }
    }
  };
  Chunk ca, cb;
  if (R) load(0, ca);
  for (uint32_t r0 = 0; r0 < R;) {
    load(r0 + kJacChunk, cb);
    run(r0, ca);
    r0 += kJacChunk;
    if (r0 >= R) break;
    load(r0 + kJacChunk, ca);
    run(r0, cb);
    r0 += kJacChunk;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t j = lane; j < m; j += 64) w.vt[static_cast<size_t>(j) * m + k] = x[j];
}

// rank by eigenvalue (descending, ties by index), sign, and write Vout
__global__ __launch_bounds__(128) void jacobi_out_kernel(uint32_t m, uint32_t q, JacobiWork w,
                                                         double *__restrict__ Vout) {
  const uint32_t j = threadIdx.x;
  if (j >= m) return;
  const double wj = w.evals[j];
  uint32_t col = 0;
  for (uint32_t i = 0; i < m; ++i) {
    const double wi = w.evals[i];
    col += (wi > wj || (wi == wj && i < j)) ? 1u : 0u;
  }
  if (col >= q) return;
  const double *v = w.vt + static_cast<size_t>(j) * m;
  double best = -1.0, sg = 1.0;
  for (uint32_t k = 0; k < m; ++k) {
    const double e = v[k];
    if (fabs(e) > best) {
      best = fabs(e);
      sg = e < 0 ? -1.0 : 1.0;
    }
  }
  for (uint32_t k = 0; k < m; ++k) Vout[static_cast<size_t>(k) * q + col] = sg * v[k];
}

// Column signs of V (d x q): largest-magnitude entry positive (first such
// entry on ties); also writes the FP32 copy used by the projection kernel,
// and the FP32 copy of the mean (rounded to nearest).  One 256-thread block
// per column (at least one block).
__global__ __launch_bounds__(256) void sign_fix_kernel(double *__restrict__ V, uint32_t d, uint32_t q,
                                                       float *__restrict__ Vf, const double *__restrict__ mean,
                                                       float *__restrict__ meanf) {
  for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < d; r += gridDim.x * 256)
    meanf[r] = static_cast<float>(mean[r]);
  if (blockIdx.x >= q) return;
  __shared__ double bv[256];
  __shared__ uint32_t bi[256];
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  double best = -1.0;
  uint32_t bidx = 0xffffffffu;
  for (uint32_t r = t; r < d; r += 256) {
    const double v = fabs(V[static_cast<size_t>(r) * q + c]);
    if (v > best) {  // rows visited in increasing order: keeps the first maximum
      best = v;
      bidx = r;
    }
  }
  bv[t] = best;
  bi[t] = bidx;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (t < w) {
      const double o = bv[t + w];
      const uint32_t oi = bi[t + w];
      if (o > bv[t] || (o == bv[t] && oi < bi[t])) {
        bv[t] = o;
        bi[t] = oi;
      }
    }
    __syncthreads();
  }
  const uint32_t rmax = bi[0];
  const double sg = (rmax < d && V[static_cast<size_t>(rmax) * q + c] < 0) ? -1.0 : 1.0;
  for (uint32_t r = t; r < d; r += 256) {
    const double v = sg * V[static_cast<size_t>(r) * q + c];
    V[static_cast<size_t>(r) * q + c] = v;
    Vf[static_cast<size_t>(r) * q + c] = static_cast<float>(v);
  }
}

uint32_t jacobi_max_size() { return kJacMax; }

hipError_t launch_chol_inv(const double *G, uint32_t b, double *Rinv, uint32_t *status, hipStream_t s) {
  if (b == 0 || b > kCholMax) return hipErrorInvalidValue;
  chol_inv_kernel<<<1, 512, 0, s>>>(G, b, Rinv, status);
  return hipGetLastError();
}

static uint64_t jacobi_records(uint32_t m) {
  const uint64_t M = (m + 1) & ~1u;
  return static_cast<uint64_t>(kJacSweeps) * (M - 1) * (M / 2);
}
uint64_t jacobi_work_doubles(uint32_t m) {
  return 2 * jacobi_records(m) + (jacobi_records(m) + 1) / 2 + static_cast<uint64_t>(m) * m + m + 1;
}

hipError_t launch_jacobi(const double *A, uint32_t m, uint32_t q, double *Vout, double *work, hipStream_t s) {
  if (m == 0 || m > kJacMax || q > m) return hipErrorInvalidValue;
  const uint64_t nr = jacobi_records(m);
  JacobiWork w;
  w.cs = reinterpret_cast<double2 *>(work);
  w.pq = reinterpret_cast<uint32_t *>(work + 2 * nr);
  w.vt = work + 2 * nr + (nr + 1) / 2;
  w.evals = w.vt + static_cast<size_t>(m) * m;
  w.rounds = reinterpret_cast<uint32_t *>(w.evals + m);
  jacobi_rot_kernel<<<1, kJacThreads, 0, s>>>(A, m, w);
  jacobi_vec_kernel<<<m, 64, 0, s>>>(m, w);
  jacobi_out_kernel<<<1, 128, 0, s>>>(m, q, w, Vout);
  return hipGetLastError();
}

hipError_t launch_sign_fix(double *V, uint32_t d, uint32_t q, float *Vf, const double *mean, float *meanf,
                           hipStream_t s) {
  if (d == 0) return hipSuccess;
  sign_fix_kernel<<<q > 0 ? q : 1, 256, 0, s>>>(V, d, q, Vf, mean, meanf);
  return hipGetLastError();
}

}  // namespace pm
