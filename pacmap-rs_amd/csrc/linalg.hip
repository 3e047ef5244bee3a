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
// C (M x N) = op(A) (M x K) * B (K x N), row-major FP64, on the FP64 matrix
// cores: one 512-thread block per 16 x 32 tile of C (small tiles: these
// products have few tiles, and more blocks reach more CUs).  The block's 8
// waves split K into contiguous ranges of 16-deep blocks; each accumulates
// the whole tile as two v_mfma_f64_16x16x4 products, and the 8 partial tiles
// are summed in LDS in wave order (a fixed order: the result does not depend
// on timing).  Within a 16-deep block, step s takes k = 4 (l / 16) + s at
// lane l, so a lane's four A values of a row are contiguous when A is not
// transposed.  Operand lane l holds element (l % 16, l / 16) of A (16 x 4)
// and (l / 16, l % 16) of B (4 x 16); the result lane l holds rows
// l / 16 + 4 v of column l % 16.  The next block's operands are loaded into
// registers while the current one is multiplied.
// ---------------------------------------------------------------------------
constexpr int kG64Waves = 8;
typedef double doublex4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kG64Waves * 64) void gemm64_mfma_kernel(uint32_t M, uint32_t N, uint32_t K,
                                                                     const double *__restrict__ A, uint32_t lda,
                                                                     int tA, const double *__restrict__ B,
                                                                     uint32_t ldb, double *__restrict__ C,
                                                                     uint32_t ldc) {
  __shared__ double sP[kG64Waves][16 * 32];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = lane & 15, g = lane >> 4;
  const uint32_t m0 = blockIdx.x * 16, n0 = blockIdx.y * 32;
  const uint32_t nkb = (K + 15) / 16;
  const uint32_t kb0 = nkb * wave / kG64Waves, kb1 = nkb * (wave + 1) / kG64Waves;
  double pa[4], pb[2][4];
  auto fetch = [&](uint32_t kb) __attribute__((always_inline)) {
    const uint32_t kbase = kb * 16 + 4 * g;
    const uint32_t r = m0 + m;
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      const uint32_t k = kbase + s2;
      const size_t off = tA ? static_cast<size_t>(k) * lda + r : static_cast<size_t>(r) * lda + k;
      pa[s2] = (r < M && k < K) ? A[off] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t c = n0 + 16 * j + m;
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) {
        const uint32_t k = kbase + s2;
        pb[j][s2] = (c < N && k < K) ? B[static_cast<size_t>(k) * ldb + c] : 0.0;
      }
    }
  };
  doublex4 acc[2] = {doublex4{0.0, 0.0, 0.0, 0.0}, doublex4{0.0, 0.0, 0.0, 0.0}};
  if (kb0 < kb1) fetch(kb0);
  for (uint32_t kb = kb0; kb < kb1; ++kb) {
    double a[4], b[2][4];
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      a[s2] = pa[s2];
      b[0][s2] = pb[0][s2];
      b[1][s2] = pb[1][s2];
    }
    if (kb + 1 < kb1) fetch(kb + 1);
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s2], b[j][s2], acc[j], 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) sP[wave][(g + 4 * v) * 32 + 16 * j + m] = acc[j][v];
  __syncthreads();
  for (int e = t; e < 16 * 32; e += kG64Waves * 64) {
    double sum = sP[0][e];
#pragma unroll
    for (int w = 1; w < kG64Waves; ++w) sum += sP[w][e];
    const uint32_t r = m0 + e / 32, c = n0 + e % 32;
    if (r < M && c < N) C[static_cast<size_t>(r) * ldc + c] = sum;
  }
}

hipError_t launch_gemm64(uint32_t M, uint32_t N, uint32_t K, const double *A, uint32_t lda, bool tA,
                         const double *B, uint32_t ldb, double *C, uint32_t ldc, hipStream_t s) {
  if (M == 0 || N == 0) return hipSuccess;
  const dim3 grid((M + 15) / 16, (N + 31) / 32);
  gemm64_mfma_kernel<<<grid, kG64Waves * 64, 0, s>>>(M, N, K, A, lda, tA ? 1 : 0, B, ldb, C, ldc);
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
        // Only rows below j change, and only 32-column blocks that can hold
        // live entries: left columns > j (columns <= j are never read
        // again) and right columns <= j (row j of E = R^-T is zero beyond
        // its diagonal).  Those form the block range [elo, ehi].
        const uint32_t elo = (j + 1) / 32, ehi = 4 + j / 32;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          if (static_cast<uint32_t>(rg + 16 * a) <= j) continue;
#pragma unroll
          for (int e = 0; e < 8; ++e)
            if (static_cast<uint32_t>(e) >= elo && static_cast<uint32_t>(e) <= ehi)
              M[a][e] = fma(-f[a], rc[e], M[a][e]);
        }
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

// ODD: m is odd, so every round pairs one index with the dummy index m
constexpr int kJacNormSlots = (kJacMax * (kJacMax + 1) / 2 + kJacThreads - 1) / kJacThreads;
template <bool ODD>
__global__ __launch_bounds__(kJacThreads) void jacobi_rot_kernel(const double *__restrict__ Ain, uint32_t m,
                                                                 JacobiWork w) {
  __shared__ double Ap[kJacMax * (kJacMax + 1) / 2];
  // round rotations: (p, q, rowbase(p), rowbase(q)) and (c, s); rowbase(x)
  // = packed_idx(x, x, m) - x, so packed_idx(x, y, m) = rowbase(x) + y for x <= y
  __shared__ int4 rq[kJacPairs];
  __shared__ double2 rcs[kJacPairs];
  __shared__ double tn[kJacPairs];
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
  // the packed entries this thread adds into the sweep's square sums:
  // e = t + 1024 k, flagged when e is a diagonal entry
  uint32_t nidx[kJacNormSlots], nvalid = 0, ndiag = 0;
#pragma unroll
  for (int k = 0; k < kJacNormSlots; ++k) {
    const uint32_t e = t + k * kJacThreads;
    nidx[k] = 0;
    if (e < m * (m + 1) / 2) {
      nidx[k] = e;
      nvalid |= 1u << k;
      for (uint32_t r = 0; r < m; ++r)
        if (packed_idx(r, r, m) == e) ndiag |= 1u << k;
    }
  }
  __syncthreads();
  uint32_t grnd = 0;  // rounds recorded
  for (int sweep = 0; sweep < kJacSweeps; ++sweep) {
    double off = 0.0, dg = 0.0;
#pragma unroll
    for (int k = 0; k < kJacNormSlots; ++k) {
      const double v = Ap[nidx[k]];
      if (nvalid >> k & 1u) {
        if (ndiag >> k & 1u) dg += v * v;
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
        // round-robin position of slot j: 1 + (j - 1 + round) mod (M - 1),
        // where j - 1 + round < 2 (M - 1)
        auto pos = [&](uint32_t j) -> uint32_t {
          if (j == 0) return 0;
          uint32_t x = j - 1 + round;
          if (x >= M - 1) x -= M - 1;
          return 1 + x;
        };
        uint32_t p = pos(t), r2 = pos(M - 1 - t);
        if (p > r2) {
          const uint32_t tmp = p;
          p = r2;
          r2 = tmp;
        }
        const bool live = !ODD || r2 < m;  // false: r2 is the dummy index
        auto rowbase = [&](uint32_t x) { return static_cast<int>(x * m - (x * (x - 1)) / 2 - x); };
        const int bp = rowbase(p), bq = live ? rowbase(r2) : 0;
        // the three entries are read together (packed_idx(x, y) = rowbase(x) + y, x <= y)
        const double apq = live ? Ap[bp + r2] : 0.0;
        const double app = Ap[bp + p];
        const double aqq = live ? Ap[bq + r2] : 0.0;
        double c = 1.0, s = 0.0, tt = 0.0;
        if (apq != 0.0) {
          // tan of the rotation angle, the smaller root of t^2 + 2 theta t - 1 = 0
          // with theta = (aqq - app) / (2 apq), written as 2 apq sgn(dd) / (|dd| +
          // sqrt(dd^2 + 4 apq^2)) to save a division
          const double dd = aqq - app;
          const double a2 = 2.0 * apq;
          const double den = fabs(dd) + sqrt(fma(dd, dd, a2 * a2));
          tt = (dd >= 0.0 ? a2 : -a2) / den;
          c = rsqrt(fma(tt, tt, 1.0));
          s = tt * c;
        }
        rq[t] = make_int4(static_cast<int>(p), static_cast<int>(r2), bp, bq);
        rcs[t] = make_double2(c, s);
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
          const int4 ra = rq[a], rb4 = rq[bb];
          const double2 ca = rcs[a], cb = rcs[bb];
          const uint32_t p1 = ra.x, q1 = ra.y, p2 = rb4.x, q2 = rb4.y;
          const double c1 = ca.x, s1 = ca.y, c2 = cb.x, s2 = cb.y;
          const bool h1 = !ODD || q1 < m, h2 = !ODD || q2 < m;  // false: the pair holds the dummy index
          auto idx = [](uint32_t x, int bx, uint32_t y, int by) -> uint32_t {
            return x <= y ? static_cast<uint32_t>(bx) + y : static_cast<uint32_t>(by) + x;
          };
          const uint32_t i00 = idx(p1, ra.z, p2, rb4.z);
          const uint32_t i01 = h2 ? idx(p1, ra.z, q2, rb4.w) : 0;
          const uint32_t i10 = h1 ? idx(q1, ra.w, p2, rb4.z) : 0;
          const uint32_t i11 = (h1 && h2) ? idx(q1, ra.w, q2, rb4.w) : 0;
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
          const int4 ra = rq[a];
          const uint32_t p = ra.x, r2 = ra.y;  // p < r2
          const double tt = tn[a];
          if ((!ODD || r2 < m) && tt != 0.0) {
            const uint32_t ipq = ra.z + r2, ipp = ra.z + p, iqq = ra.w + r2;
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
  if (m & 1)
    jacobi_rot_kernel<true><<<1, kJacThreads, 0, s>>>(A, m, w);
  else
    jacobi_rot_kernel<false><<<1, kJacThreads, 0, s>>>(A, m, w);
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
