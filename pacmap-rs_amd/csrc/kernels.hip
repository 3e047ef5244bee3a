// kernels.hip -- hand-written CDNA4 (gfx950) kernels of the PaCMAP hot path.
//
// Built with -ffp-contract=off: the reference (Rust) never fuses a multiply
// and an add, and the per-pair arithmetic here must round exactly like it.
#include "internal.h"
#include "grad_device.h"

#include <cstdint>
#include <type_traits>

namespace pm {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
// Monotone map f32 -> u32 whose unsigned order is Rust's f32::total_cmp.
__device__ __forceinline__ uint32_t tkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_tkey(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(b);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
  lo = __shfl_xor(lo, m, 64);
  hi = __shfl_xor(hi, m, 64);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
  lo = __shfl(lo, src, 64);
  hi = __shfl(hi, src, 64);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Ascending bitonic sort of 64*R keys held by one full wavefront: element
// r*64 + lane lives in v[r] of `lane`.  Strides below 64 exchange across
// lanes, larger strides between registers of the same lane.
template <int R>
__device__ __forceinline__ void wave_bitonic(uint64_t (&v)[R], int lane) {
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
          const uint64_t o = shfl_xor_u64(v[r], stride);
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
__device__ __forceinline__ uint64_t wave_pick(const uint64_t (&v)[R], uint32_t pos) {
  uint64_t sel = 0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (static_cast<uint32_t>(r) == (pos >> 6)) sel = v[r];
  return shfl_u64(sel, static_cast<int>(pos & 63u));
}

// ---------------------------------------------------------------------------
// exact kNN (knn.rs:137-194 with distance.rs:50-75 arithmetic)
// ---------------------------------------------------------------------------
KnnLayout knn_layout(uint32_t d) {
  KnnLayout L;
  L.d = d;
  L.C = d / 8;
  L.R = d % 8;
  L.Cp = (L.C + 3u) / 4u * 4u;
  const uint32_t used = 8u * L.Cp + L.R;
  L.Dp = (used + kKnnChunk - 1u) / kKnnChunk * kKnnChunk;
  return L;
}

__global__ void knn_pack_kernel(const float *__restrict__ x, float *__restrict__ xp, uint64_t n,
                                uint32_t d, uint32_t C, uint32_t Cp, uint32_t Dp) {
  const uint64_t total = n * Dp;
  for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < total;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t row = e / Dp;
    const uint32_t f = static_cast<uint32_t>(e - row * Dp);
    float v = 0.0f;
    if (f < 8u * Cp) {
      const uint32_t l = f / Cp, c = f - l * Cp;
      if (c < C) v = x[row * d + 8u * c + l];
    } else {
      const uint32_t r = f - 8u * Cp;
      if (8u * C + r < d) v = x[row * d + 8u * C + r];
    }
    xp[e] = v;
  }
}

hipError_t launch_knn_pack(const float *x, float *xp, uint64_t n, const KnnLayout &L,
                           hipStream_t s) {
  const uint64_t total = n * L.Dp;
  if (total == 0) return hipSuccess;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  knn_pack_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(x, xp, n, L.d, L.C, L.Cp, L.Dp);
  return hipGetLastError();
}

// One block owns QB = 16*TQ queries and streams every candidate tile of 64
// rows.  Each thread computes a TQ x 4 tile of distances with two registers
// per pair (the running lane accumulator and the running total), so the
// reference's order -- lane partials over 8-chunks, lanes added 0..7, then
// the tail -- is reproduced exactly.  Candidates below the query's current
// k-th key (dist bits, index) are appended to an LDS buffer of KCAP keys; a
// wavefront bitonic sort compacts it back to k when it could overflow.
template <int TQ, int KCAP>
__global__ __launch_bounds__(256) void knn_exact_kernel(const float *__restrict__ xp, uint32_t n,
                                                        uint32_t Dp, uint32_t Cp, uint32_t k,
                                                        uint32_t *__restrict__ oidx,
                                                        float *__restrict__ odist,
                                                        const uint32_t *__restrict__ qlist,
                                                        uint32_t nq, uint32_t cchunk,
                                                        uint64_t *__restrict__ part) {
  constexpr int QB = 16 * TQ, CB = 64, KC = static_cast<int>(kKnnChunk), R = KCAP / 64;
  __shared__ __align__(16) float sQ[KC][QB];
  __shared__ __align__(16) float sC[KC][CB];
  __shared__ uint64_t sBuf[QB][KCAP];
  __shared__ uint64_t sTau[QB];
  __shared__ uint32_t sCnt[QB];

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int tq = t >> 4, tc = t & 15;
  const uint32_t qbase = blockIdx.x * QB;
  const uint32_t segEnd = 8u * Cp;
  const uint32_t thresh = KCAP - CB;
  // query rows of this block: consecutive rows, or entries of qlist
  auto qrow = [&](int ql) -> uint32_t {
    const uint32_t p = qbase + static_cast<uint32_t>(ql);
    if (qlist) return p < nq ? qlist[p] : 0xffffffffu;
    return p < n ? p : 0xffffffffu;
  };
  for (int i = t; i < QB; i += 256) {
    sCnt[i] = 0;
    sTau[i] = ~0ull;
  }
  __syncthreads();

  // candidate range of this block (blockIdx.y splits the rows 0..n-1)
  const uint32_t c_begin = blockIdx.y * cchunk;
  const uint32_t c_end = (c_begin + cchunk < n) ? c_begin + cchunk : n;
  for (uint32_t cbase = c_begin; cbase < c_end; cbase += CB) {
    float s[TQ][4], acc[TQ][4];
#pragma unroll
    for (int i = 0; i < TQ; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        s[i][j] = 0.0f;
        acc[i][j] = 0.0f;
      }

    for (uint32_t f0 = 0; f0 < Dp; f0 += KC) {
      __syncthreads();
      for (int li = t; li < QB * (KC / 4); li += 256) {
        const int q = li / (KC / 4), f4 = li % (KC / 4);
        const uint32_t row = qrow(q);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < n) v = *reinterpret_cast<const float4 *>(xp + static_cast<size_t>(row) * Dp + f0 + 4 * f4);
        sQ[4 * f4 + 0][q] = v.x;
        sQ[4 * f4 + 1][q] = v.y;
        sQ[4 * f4 + 2][q] = v.z;
        sQ[4 * f4 + 3][q] = v.w;
      }
      for (int li = t; li < CB * (KC / 4); li += 256) {
        const int c = li / (KC / 4), f4 = li % (KC / 4);
        const uint32_t row = cbase + c;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < n) v = *reinterpret_cast<const float4 *>(xp + static_cast<size_t>(row) * Dp + f0 + 4 * f4);
        sC[4 * f4 + 0][c] = v.x;
        sC[4 * f4 + 1][c] = v.y;
        sC[4 * f4 + 2][c] = v.z;
        sC[4 * f4 + 3][c] = v.w;
      }
      __syncthreads();
#pragma unroll
      for (int sc = 0; sc < KC / 4; ++sc) {
        const uint32_t fs = f0 + 4u * sc;
        if (fs >= segEnd) {
          // tail: distance.rs:69-72 adds each (a-b)^2 to the total directly
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float qv[TQ];
#pragma unroll
            for (int i = 0; i < TQ; ++i) qv[i] = sQ[4 * sc + e][tq * TQ + i];
            const float4 cv = *reinterpret_cast<const float4 *>(&sC[4 * sc + e][tc * 4]);
            const float cvs[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
            for (int i = 0; i < TQ; ++i)
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const float df = qv[i] - cvs[j];
                const float sq = df * df;
                s[i][j] = s[i][j] + sq;
              }
          }
        } else {
          // lane segment: distance.rs:61-64 accumulates per lane
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float qv[TQ];
#pragma unroll
            for (int i = 0; i < TQ; ++i) qv[i] = sQ[4 * sc + e][tq * TQ + i];
            const float4 cv = *reinterpret_cast<const float4 *>(&sC[4 * sc + e][tc * 4]);
            const float cvs[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
            for (int i = 0; i < TQ; ++i)
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const float df = qv[i] - cvs[j];
                const float sq = df * df;
                acc[i][j] = acc[i][j] + sq;
              }
          }
          if ((fs + 4u) % Cp == 0u) {
            // end of lane segment: distance.rs:66 adds lanes in order
#pragma unroll
            for (int i = 0; i < TQ; ++i)
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                s[i][j] = s[i][j] + acc[i][j];
                acc[i][j] = 0.0f;
              }
          }
        }
      }
    }

    // candidate filter against the query's current k-th key
#pragma unroll
    for (int i = 0; i < TQ; ++i) {
      const int ql = tq * TQ + i;
      const uint32_t q = qrow(ql);
      if (q >= n) continue;
      const uint64_t tau = sTau[ql];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t c = cbase + tc * 4 + j;
        if (c >= c_end || c == q) continue;
        const float dd = sqrtf(s[i][j]);  // distance.rs:74
        const uint64_t key = (static_cast<uint64_t>(tkey(dd)) << 32) | c;
        if (key < tau) {
          const uint32_t pos = atomicAdd(&sCnt[ql], 1u);
          sBuf[ql][pos] = key;
        }
      }
    }
    __syncthreads();
    for (int ql = wave; ql < QB; ql += 4) {
      const uint32_t cnt = sCnt[ql];
      if (cnt > thresh) {
        uint64_t v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t e = r * 64 + lane;
          v[r] = e < cnt ? sBuf[ql][e] : ~0ull;
        }
        wave_bitonic<R>(v, lane);
        const uint32_t keep = cnt < k ? cnt : k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t e = r * 64 + lane;
          if (e < keep) sBuf[ql][e] = v[r];
        }
        const uint64_t tau = (cnt >= k) ? wave_pick<R>(v, k - 1) : ~0ull;
        if (lane == 0) {
          sCnt[ql] = keep;
          sTau[ql] = tau;
        }
      }
    }
    __syncthreads();
  }

  // final ordering: ascending (distance, index), first k
  for (int ql = wave; ql < QB; ql += 4) {
    const uint32_t q = qrow(ql);
    if (q >= n) continue;
    const uint32_t cnt = sCnt[ql];
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t e = r * 64 + lane;
      v[r] = e < cnt ? sBuf[ql][e] : ~0ull;
    }
    wave_bitonic<R>(v, lane);
    if (part) {  // partial list of this candidate range, merged by knn_merge_kernel
      const size_t base = (static_cast<size_t>(qbase + ql) * gridDim.y + blockIdx.y) * k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t e = r * 64 + lane;
        if (e < k) part[base + e] = v[r];
      }
      continue;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t e = r * 64 + lane;
      if (e < k) {
        oidx[static_cast<size_t>(q) * k + e] = static_cast<uint32_t>(v[r]);
        odist[static_cast<size_t>(q) * k + e] = from_tkey(static_cast<uint32_t>(v[r] >> 32));
      }
    }
  }
}

// Final bitonic merge stage over 64*R keys (input bitonic, output ascending).
template <int R>
__device__ __forceinline__ void wave_bitonic_last(uint64_t (&v)[R], int lane) {
  constexpr int N = 64 * R;
#pragma unroll
  for (int stride = N >> 1; stride > 0; stride >>= 1) {
    if (stride >= 64) {
      const int rs = stride >> 6;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if ((r & rs) == 0) {
          const uint64_t a = v[r], b = v[r | rs];
          v[r] = a < b ? a : b;
          v[r | rs] = a < b ? b : a;
        }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint64_t o = shfl_xor_u64(v[r], stride);
        const bool lower = (lane & stride) == 0;
        v[r] = lower ? (v[r] < o ? v[r] : o) : (v[r] < o ? o : v[r]);
      }
    }
  }
}

// Merges the S sorted partial lists (k keys each) of one listed query row
// into its first k (one workgroup per row): the running list A and each
// sorted chunk B give min(A[i], B[N-1-i]), a bitonic sequence holding the N
// smallest of both, which the last bitonic stage sorts.  The four waves take
// every fourth chunk, then wave 0 merges their four running lists the same
// way.  Keys are unique (distance bits, index), so the result does not
// depend on the merge order.
template <int R>
__global__ __launch_bounds__(256) void knn_merge_kernel(const uint64_t *__restrict__ part,
                                                        uint32_t nq, uint32_t S, uint32_t k,
                                                        const uint32_t *__restrict__ qlist,
                                                        uint32_t *__restrict__ oidx,
                                                        float *__restrict__ odist) {
  constexpr uint32_t N = 64 * R;
  __shared__ uint64_t sRun[3][N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t slot = blockIdx.x;
  if (slot >= nq) return;
  const uint64_t *src = part + static_cast<size_t>(slot) * S * k;
  const uint32_t total = S * k;
  uint64_t run[R];
#pragma unroll
  for (int r = 0; r < R; ++r) run[r] = ~0ull;
  // the next chunk's keys are loaded while the current one is merged
  uint64_t nxt[R];
  const uint32_t first = static_cast<uint32_t>(wave) * N;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = first + r * 64 + lane;
    nxt[r] = e < total ? src[e] : ~0ull;
  }
  for (uint32_t c0 = first; c0 < total; c0 += 4 * N) {
    uint64_t b[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      b[r] = nxt[r];
      const uint32_t e = c0 + 4 * N + r * 64 + lane;
      nxt[r] = e < total ? src[e] : ~0ull;
    }
    wave_bitonic<R>(b, lane);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t rev = shfl_xor_u64(b[R - 1 - r], 63);  // element N-1-i
      run[r] = run[r] < rev ? run[r] : rev;
    }
    wave_bitonic_last<R>(run, lane);
  }
  if (wave > 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) sRun[wave - 1][r * 64 + lane] = run[r];
  }
  __syncthreads();
  if (wave != 0) return;
  for (int w = 0; w < 3; ++w) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t rev = sRun[w][N - 1 - (r * 64 + lane)];  // element N-1-i of a sorted list
      run[r] = run[r] < rev ? run[r] : rev;
    }
    wave_bitonic_last<R>(run, lane);
  }
  const uint32_t q = qlist[slot];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = r * 64 + lane;
    if (e < k) {
      oidx[static_cast<size_t>(q) * k + e] = static_cast<uint32_t>(run[r]);
      odist[static_cast<size_t>(q) * k + e] = from_tkey(static_cast<uint32_t>(run[r] >> 32));
    }
  }
}

hipError_t launch_knn_rescan(const float *xp, uint64_t n, const KnnLayout &L, uint32_t k,
                             const uint32_t *qlist, uint32_t nq, uint64_t *part, uint32_t S,
                             uint32_t *idx, float *dist, hipStream_t s) {
  if (nq == 0 || k == 0) return hipSuccess;
  const uint32_t nn = static_cast<uint32_t>(n);
  const uint32_t cchunk = ((nn + S - 1) / S + 63) / 64 * 64;
  const uint32_t blocks = (nq + 15) / 16;
  dim3 g(blocks, S);
  // 16 query rows per block keep the partial buffer indexable by list slot
  if (k <= 64)
    knn_exact_kernel<1, 128><<<g, 256, 0, s>>>(xp, nn, L.Dp, L.Cp, k, idx, dist, qlist, nq, cchunk, part);
  else if (k <= 192)
    knn_exact_kernel<1, 256><<<g, 256, 0, s>>>(xp, nn, L.Dp, L.Cp, k, idx, dist, qlist, nq, cchunk, part);
  else if (k <= kKnnMaxK)
    knn_exact_kernel<1, 512><<<g, 256, 0, s>>>(xp, nn, L.Dp, L.Cp, k, idx, dist, qlist, nq, cchunk, part);
  else
    return hipErrorInvalidValue;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint32_t mb = nq;  // one workgroup per listed row
  if (k <= 64)
    knn_merge_kernel<1><<<mb, 256, 0, s>>>(part, nq, S, k, qlist, idx, dist);
  else if (k <= 128)
    knn_merge_kernel<2><<<mb, 256, 0, s>>>(part, nq, S, k, qlist, idx, dist);
  else if (k <= 256)
    knn_merge_kernel<4><<<mb, 256, 0, s>>>(part, nq, S, k, qlist, idx, dist);
  else
    knn_merge_kernel<8><<<mb, 256, 0, s>>>(part, nq, S, k, qlist, idx, dist);
  return hipGetLastError();
}

hipError_t launch_knn_exact(const float *xp, uint64_t n, const KnnLayout &L, uint32_t k,
                            uint32_t *idx, float *dist, hipStream_t s, const uint32_t *qlist,
                            uint32_t nq) {
  if (n == 0 || k == 0) return hipSuccess;
  const uint32_t nn = static_cast<uint32_t>(n);
  const uint32_t rows = qlist ? nq : nn;
  if (rows == 0) return hipSuccess;
  if (k <= 64) {
    knn_exact_kernel<4, 128><<<(rows + 63) / 64, 256, 0, s>>>(xp, nn, L.Dp, L.Cp, k, idx, dist, qlist, nq, (nn + 63) / 64 * 64, nullptr);
  } else if (k <= 192) {
    knn_exact_kernel<2, 256><<<(rows + 31) / 32, 256, 0, s>>>(xp, nn, L.Dp, L.Cp, k, idx, dist, qlist, nq, (nn + 63) / 64 * 64, nullptr);
  } else if (k <= kKnnMaxK) {
    knn_exact_kernel<1, 512><<<(rows + 15) / 16, 256, 0, s>>>(xp, nn, L.Dp, L.Cp, k, idx, dist, qlist, nq, (nn + 63) / 64 * 64, nullptr);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// sigma and scaled distances (neighbors.rs:77-87, distance.rs:27-37)
// ---------------------------------------------------------------------------
__global__ void sigma_kernel(const float *__restrict__ dist, uint32_t n, uint32_t kp,
                             uint32_t start, uint32_t end, float *__restrict__ sig) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (end <= start) {  // empty lane: mean_axis is None -> 1e-10
    sig[i] = 1e-10f;
    return;
  }
  float s = 0.0f;  // ndarray sums a short lane sequentially from zero
  for (uint32_t c = start; c < end; ++c) s = s + dist[static_cast<size_t>(i) * kp + c];
  const float mean = s / static_cast<float>(end - start);
  sig[i] = fmaxf(mean, 1e-10f);  // f32::max returns the non-NaN operand
}

hipError_t launch_sigma(const float *dist, uint64_t n, uint32_t kp, float *sig, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint32_t start = kp >= 1 ? kp - 1 : 0;
  if (start > 3) start = 3;
  const uint32_t end = kp < 6 ? kp : 6;
  sigma_kernel<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, s>>>(
      dist, static_cast<uint32_t>(n), kp, start, end, sig);
  return hipGetLastError();
}

// One wavefront per row: scaled distances, then the row's scaled order
// (ties by column) picks the first knn neighbours (sampling.rs:251-286).
template <int R>
__global__ __launch_bounds__(256) void scale_nn_kernel(const float *__restrict__ dist,
                                                       const uint32_t *__restrict__ idx,
                                                       const float *__restrict__ sig, uint32_t n,
                                                       uint32_t kp, uint32_t knn,
                                                       float *__restrict__ scaled,
                                                       uint32_t *__restrict__ pairs) {
  const int lane = threadIdx.x & 63;
  const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const float si = sig[row];
  uint64_t v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t c = r * 64 + lane;
    v[r] = ~0ull;
    if (c < kp) {
      const size_t o = static_cast<size_t>(row) * kp + c;
      const float d = dist[o];
      const float dd = d * d;
      const float ss = si * sig[idx[o]];
      const float sc = dd / ss;
      scaled[o] = sc;
      v[r] = (static_cast<uint64_t>(tkey(sc)) << 32) | c;
    }
  }
  if (pairs == nullptr) return;
  wave_bitonic<R>(v, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t j = r * 64 + lane;
    if (j < knn) {
      const uint32_t col = static_cast<uint32_t>(v[r]);
      const size_t o = (static_cast<size_t>(row) * knn + j) * 2;
      pairs[o] = row;
      pairs[o + 1] = idx[static_cast<size_t>(row) * kp + col];
    }
  }
}

hipError_t launch_scale_nn(const float *dist, const uint32_t *idx, const float *sig, uint64_t n,
                           uint32_t kp, uint32_t knn, float *scaled, uint32_t *pairs,
                           hipStream_t s) {
  if (n == 0 || kp == 0) return hipSuccess;
  const uint32_t blocks = static_cast<uint32_t>((n + 3) / 4), nn = static_cast<uint32_t>(n);
  if (kp <= 64)
    scale_nn_kernel<1><<<blocks, 256, 0, s>>>(dist, idx, sig, nn, kp, knn, scaled, pairs);
  else if (kp <= 128)
    scale_nn_kernel<2><<<blocks, 256, 0, s>>>(dist, idx, sig, nn, kp, knn, scaled, pairs);
  else if (kp <= 256)
    scale_nn_kernel<4><<<blocks, 256, 0, s>>>(dist, idx, sig, nn, kp, knn, scaled, pairs);
  else if (kp <= 512)
    scale_nn_kernel<8><<<blocks, 256, 0, s>>>(dist, idx, sig, nn, kp, knn, scaled, pairs);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fused gradient + Adam over the symmetrised CSR (gradient.rs:98-151,
// adam.rs:51-59)
// ---------------------------------------------------------------------------
// 16 lanes own one row.  Lane l walks entries beg+l, beg+l+16, ... in order
// and accumulates s * w1(D) * (y_row - y_other) exactly as the reference
// computes each per-pair term; the 16 partials are combined by a fixed xor
// butterfly, so the result is bit-identical from run to run (no atomics).
template <int DIMS, bool ADAM, bool LOSS, int U>
__global__ __launch_bounds__(256) void grad_adam_kernel(const GradArgs a) {
  const int t = threadIdx.x, gl = t & 15, grp = t >> 4;
  const uint32_t row = blockIdx.x * kGradRowsPerBlock + grp;
  const PairConsts pc = pair_consts(a.w);
  const float kn = pc.k_nn, km = pc.k_mn, kf = pc.k_fp, wn = pc.w_nn, wm = pc.w_mn, wf = pc.w_fp;
  float g[DIMS];
  float yr[DIMS], mr[DIMS], vr[DIMS];
#pragma unroll
  for (int d = 0; d < DIMS; ++d) {
    g[d] = 0.0f;
    yr[d] = 0.0f;
    mr[d] = 0.0f;
    vr[d] = 0.0f;
  }
  float loss = 0.0f;
  if (row < a.n) {
    const size_t base = static_cast<size_t>(row) * DIMS;
    // the Adam state does not depend on the gradient: fetch it up front
    if (ADAM && gl == 0) {
#pragma unroll
      for (int d = 0; d < DIMS; ++d) {
        mr[d] = a.m[base + d];
        vr[d] = a.v[base + d];
      }
    }
#pragma unroll
    for (int d = 0; d < DIMS; ++d) yr[d] = a.y_in[base + d];
    const uint32_t beg = a.row_ptr[row], end = a.row_ptr[row + 1];
    // lane gl owns entries beg+gl, beg+gl+16, ...; U of them are loaded
    // (entry word, then the other point's y) before any is used, so one
    // round of loads covers rows of up to 16 U entries; entries are consumed
    // in the same order
    for (uint32_t e0 = beg + gl; e0 < end; e0 += 16 * U) {
      uint32_t wv[U];
      float yo[U][DIMS];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const uint32_t e = e0 + 16u * uu;
        wv[uu] = e < end ? a.ent[e] : 0xffffffffu;
      }
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        if (wv[uu] == 0xffffffffu) break;  // slots past the row end: no gather
        const uint32_t o = wv[uu] & kIdxMask;
#pragma unroll
        for (int d = 0; d < DIMS; ++d) yo[uu][d] = a.y_in[static_cast<size_t>(o) * DIMS + d];
      }
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const uint32_t w = wv[uu];
        if (w == 0xffffffffu) break;
        pair_term<DIMS, LOSS>(w, yr, yo[uu], kn, km, kf, wn, wm, wf, g, loss);
      }
    }
  }
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < DIMS; ++d) g[d] = g[d] + __shfl_xor(g[d], off, 64);
    if (LOSS) loss = loss + __shfl_xor(loss, off, 64);
  }
  if (gl == 0 && row < a.n) {
    const size_t base = static_cast<size_t>(row) * DIMS;
    if (ADAM) {
#pragma unroll
      for (int d = 0; d < DIMS; ++d) {
        float m = mr[d], v = vr[d];
        m = m + a.omb1 * (g[d] - m);          // adam.rs:57
        const float gg = g[d] * g[d];
        v = v + a.omb2 * (gg - v);            // adam.rs:58
        const float num = a.lr_t * m;
        const float den = sqrtf(v) + 1e-7f;
        const float y = yr[d] - num / den;    // adam.rs:59
        a.m[base + d] = m;
        a.v[base + d] = v;
        a.y_out[base + d] = y;
        if (a.snap) a.snap[base + d] = y;
      }
    } else {
#pragma unroll
      for (int d = 0; d < DIMS; ++d) a.grad_out[base + d] = g[d];
    }
  }
  if (LOSS) {
    __shared__ float sLoss[kGradRowsPerBlock];
    if (gl == 0) sLoss[grp] = loss;
    __syncthreads();
    if (t == 0) {
      float s = 0.0f;
      for (uint32_t i = 0; i < kGradRowsPerBlock; ++i) s = s + sLoss[i];
      a.loss_part[blockIdx.x] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Target-blocked variant for embeddings larger than one XCD's L2.  The other
// points are split into kTargetBlocks contiguous index ranges; each CSR row
// is stably partitioned by the range of its entries (tb_offsets below), and
// pass 1 runs one workgroup per (row group, range) with blockIdx.x % 8 = the
// range, so with the round-robin dispatch of workgroups over the 8 XCDs
// every XCD gathers from one eighth of Y, which stays in its L2.  Pass 1
// writes per-range partial sums; pass 2 adds them in range order and applies
// Adam.  Deterministic: fixed lane order, butterfly and range order.
// ---------------------------------------------------------------------------
constexpr uint32_t kTbLanes = 4;                  // lanes per row in pass 1
constexpr uint32_t kTbRows = 256 / kTbLanes;      // rows per pass-1 workgroup

// Stable partition of each row by target range; off[b * n + row] (b = 1..7,
// u16, relative to the row start) is where range b begins; bad[0] is set when
// a row is longer than 65535 entries (the caller then keeps the plain path).
// One wave per row, 64 entries at a time: per-range counts from ballots,
// then each entry's place = its range's start + earlier entries of the same
// range (ballot below the lane, plus earlier chunks).
__global__ __launch_bounds__(256) void tb_partition_kernel(const uint32_t *__restrict__ row_ptr,
                                                           const uint32_t *__restrict__ ent, uint32_t n,
                                                           uint32_t span, uint32_t *__restrict__ out,
                                                           uint16_t *__restrict__ off,
                                                           uint32_t *__restrict__ bad) {
  const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const uint32_t beg = row_ptr[row], end = row_ptr[row + 1];
  if (end - beg > 65535u) {
    if (lane == 0) atomicOr(bad, 1u);
    return;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t pos[kTargetBlocks];
#pragma unroll
  for (uint32_t q = 0; q < kTargetBlocks; ++q) pos[q] = 0;
  for (uint32_t e0 = beg; e0 < end; e0 += 64) {
    const uint32_t e = e0 + lane;
    const uint32_t b = e < end ? (ent[e] & kIdxMask) / span : kTargetBlocks;
#pragma unroll
    for (uint32_t q = 0; q < kTargetBlocks; ++q) pos[q] += __popcll(__ballot(b == q));
  }
  uint32_t acc = 0;
#pragma unroll
  for (uint32_t q = 0; q < kTargetBlocks; ++q) {
    const uint32_t c = pos[q];
    pos[q] = acc;
    if (q > 0 && lane == 0) off[static_cast<size_t>(q) * n + row] = static_cast<uint16_t>(acc);
    acc += c;
  }
  for (uint32_t e0 = beg; e0 < end; e0 += 64) {
    const uint32_t e = e0 + lane;
    const uint32_t w = e < end ? ent[e] : 0u;
    const uint32_t b = e < end ? (w & kIdxMask) / span : kTargetBlocks;
    uint32_t p = 0;
#pragma unroll
    for (uint32_t q = 0; q < kTargetBlocks; ++q) {
      const uint64_t m = __ballot(b == q);
      if (b == q) p = pos[q] + __popcll(m & below);
      pos[q] += __popcll(m);
    }
    if (e < end) out[beg + p] = w;
  }
}

template <int DIMS, bool LOSS>
__global__ __launch_bounds__(256) void grad_tb_kernel(const GradArgs a, const uint16_t *__restrict__ off,
                                                      float *__restrict__ part) {
  const uint32_t b = blockIdx.x % kTargetBlocks, grp = blockIdx.x / kTargetBlocks;
  const int t = threadIdx.x, gl = t & (kTbLanes - 1);
  const uint32_t row = grp * kTbRows + (t / kTbLanes);
  const PairConsts pc = pair_consts(a.w);
  const float kn = pc.k_nn, km = pc.k_mn, kf = pc.k_fp, wn = pc.w_nn, wm = pc.w_mn, wf = pc.w_fp;
  float g[DIMS], yr[DIMS];
#pragma unroll
  for (int d = 0; d < DIMS; ++d) {
    g[d] = 0.0f;
    yr[d] = 0.0f;
  }
  float loss = 0.0f;
  if (row < a.n) {
#pragma unroll
    for (int d = 0; d < DIMS; ++d) yr[d] = a.y_in[static_cast<size_t>(row) * DIMS + d];
    const uint32_t rb = a.row_ptr[row];
    const uint32_t beg = rb + (b == 0 ? 0u : off[static_cast<size_t>(b) * a.n + row]);
    const uint32_t end = b + 1 == kTargetBlocks ? a.row_ptr[row + 1] : rb + off[static_cast<size_t>(b + 1) * a.n + row];
    constexpr int U = 4;
    for (uint32_t e0 = beg + gl; e0 < end; e0 += kTbLanes * U) {
      uint32_t wv[U];
      float yo[U][DIMS];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const uint32_t e = e0 + kTbLanes * uu;
        wv[uu] = e < end ? a.ent[e] : 0xffffffffu;
      }
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        if (wv[uu] == 0xffffffffu) break;  // slots past the row end: no gather
        const uint32_t o = wv[uu] & kIdxMask;
#pragma unroll
        for (int d = 0; d < DIMS; ++d) yo[uu][d] = a.y_in[static_cast<size_t>(o) * DIMS + d];
      }
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const uint32_t w = wv[uu];
        if (w == 0xffffffffu) break;
        pair_term<DIMS, LOSS>(w, yr, yo[uu], kn, km, kf, wn, wm, wf, g, loss);
      }
    }
  }
#pragma unroll
  for (int o = kTbLanes / 2; o > 0; o >>= 1) {
#pragma unroll
    for (int d = 0; d < DIMS; ++d) g[d] = g[d] + __shfl_xor(g[d], o, 64);
    if (LOSS) loss = loss + __shfl_xor(loss, o, 64);
  }
  if (gl == 0 && row < a.n) {
#pragma unroll
    for (int d = 0; d < DIMS; ++d) part[(static_cast<size_t>(b) * a.n + row) * DIMS + d] = g[d];
  }
  if (LOSS) {
    __shared__ float sLoss[kTbRows];
    if (gl == 0) sLoss[t / kTbLanes] = loss;
    __syncthreads();
    if (t == 0) {
      float s = 0.0f;
      for (uint32_t i = 0; i < kTbRows; ++i) s = s + sLoss[i];
      a.loss_part[blockIdx.x] = s;
    }
  }
}

// Sum of a row's kTargetBlocks range partials in range order, then Adam
// (adam.rs:51-59) or the plain gradient.
template <int DIMS, bool ADAM>
__device__ __forceinline__ void combine_row(const GradArgs &a, const float *__restrict__ part, uint32_t row) {
  const size_t base = static_cast<size_t>(row) * DIMS;
  float g[DIMS];
#pragma unroll
  for (int d = 0; d < DIMS; ++d) g[d] = part[base + d];
  for (uint32_t b = 1; b < kTargetBlocks; ++b)
#pragma unroll
    for (int d = 0; d < DIMS; ++d) g[d] = g[d] + part[(static_cast<size_t>(b) * a.n + row) * DIMS + d];
  if (ADAM) {
#pragma unroll
    for (int d = 0; d < DIMS; ++d) {
      float m = a.m[base + d], v = a.v[base + d];
      m = m + a.omb1 * (g[d] - m);          // adam.rs:57
      const float gg = g[d] * g[d];
      v = v + a.omb2 * (gg - v);            // adam.rs:58
      const float num = a.lr_t * m;
      const float den = sqrtf(v) + 1e-7f;
      const float y = a.y_in[base + d] - num / den;  // adam.rs:59
      a.m[base + d] = m;
      a.v[base + d] = v;
      a.y_out[base + d] = y;
      if (a.snap) a.snap[base + d] = y;
    }
  } else {
#pragma unroll
    for (int d = 0; d < DIMS; ++d) a.grad_out[base + d] = g[d];
  }
}

template <int DIMS, bool ADAM>
__global__ __launch_bounds__(256) void grad_tb_combine_kernel(const GradArgs a, const float *__restrict__ part) {
  const uint32_t row = blockIdx.x * 256 + threadIdx.x;
  if (row < a.n) combine_row<DIMS, ADAM>(a, part, row);
}

// The same combine with one thread per coordinate (twice the waves for
// dims = 2, half the dependent work per thread).
template <bool ADAM, int RANGES>
__global__ __launch_bounds__(256) void grad_combine_elem_kernel(const GradArgs a, const float *__restrict__ part,
                                                                uint64_t count, uint32_t chunk_elems,
                                                                uint32_t per_chunk) {
  // workgroup -> (chunk c, piece j) with c % 8 = blockIdx.x % 8 (the XCD
  // that wrote the chunk's partials)
  const uint32_t x = blockIdx.x % 8, k = blockIdx.x / 8;
  const uint32_t j = k % per_chunk, c = (k / per_chunk) * 8 + x;
  const uint32_t off = j * 256 + threadIdx.x;
  if (off >= chunk_elems) return;
  const uint64_t e = static_cast<uint64_t>(c) * chunk_elems + off;
  if (e >= count) return;
  float pv[RANGES];  // all loads in flight before the ordered sum
#pragma unroll
  for (int b = 0; b < RANGES; ++b) pv[b] = part[b * count + e];
  float g = pv[0];
#pragma unroll
  for (int b = 1; b < RANGES; ++b) g = g + pv[b];
  if (ADAM) {
    float m = a.m[e], v = a.v[e];
    m = m + a.omb1 * (g - m);          // adam.rs:57
    const float gg = g * g;
    v = v + a.omb2 * (gg - v);         // adam.rs:58
    const float num = a.lr_t * m;
    const float den = sqrtf(v) + 1e-7f;
    const float y = a.y_in[e] - num / den;  // adam.rs:59
    a.m[e] = m;
    a.v[e] = v;
    a.y_out[e] = y;
    if (a.snap) a.snap[e] = y;
  } else {
    a.grad_out[e] = g;
  }
}

// Points per target range, a multiple of 4 so that every range starts on a
// 16-byte boundary of Y for any dims.
uint32_t tb_span(uint32_t n) { return ((n + kTargetBlocks - 1) / kTargetBlocks + 3u) & ~3u; }
uint32_t grad_tb_blocks(uint32_t n) { return ((n + kTbRows - 1) / kTbRows) * kTargetBlocks; }

hipError_t launch_tb_partition(const uint32_t *row_ptr, const uint32_t *ent, uint32_t n, uint32_t *out,
                               uint16_t *off, uint32_t *bad, hipStream_t s) {
  if (n == 0) return hipSuccess;
  tb_partition_kernel<<<(n + 3) / 4, 256, 0, s>>>(row_ptr, ent, n, tb_span(n), out, off, bad);
  return hipGetLastError();
}

template <int DIMS>
static hipError_t launch_tb_dims(const GradArgs &a, const uint16_t *off, float *part, bool adam, hipStream_t s) {
  const uint32_t g1 = grad_tb_blocks(a.n);
  if (g1 == 0) return hipSuccess;
  if (a.loss_part)
    grad_tb_kernel<DIMS, true><<<g1, 256, 0, s>>>(a, off, part);
  else
    grad_tb_kernel<DIMS, false><<<g1, 256, 0, s>>>(a, off, part);
  const uint32_t g2 = (a.n + 255) / 256;
  if (adam)
    grad_tb_combine_kernel<DIMS, true><<<g2, 256, 0, s>>>(a, part);
  else
    grad_tb_combine_kernel<DIMS, false><<<g2, 256, 0, s>>>(a, part);
  return hipGetLastError();
}

template <int DIMS>
static hipError_t launch_tb_combine_dims(const GradArgs &a, const float *part, bool adam, hipStream_t s) {
  const uint32_t g2 = (a.n + 255) / 256;
  if (g2 == 0) return hipSuccess;
  if (adam)
    grad_tb_combine_kernel<DIMS, true><<<g2, 256, 0, s>>>(a, part);
  else
    grad_tb_combine_kernel<DIMS, false><<<g2, 256, 0, s>>>(a, part);
  return hipGetLastError();
}

hipError_t launch_tb_combine(const GradArgs &a, const float *part, uint32_t dims, bool adam, hipStream_t s) {
  switch (dims) {
    case 1: return launch_tb_combine_dims<1>(a, part, adam, s);
    case 2: return launch_tb_combine_dims<2>(a, part, adam, s);
    case 3: return launch_tb_combine_dims<3>(a, part, adam, s);
    case 4: return launch_tb_combine_dims<4>(a, part, adam, s);
    case 5: return launch_tb_combine_dims<5>(a, part, adam, s);
    case 6: return launch_tb_combine_dims<6>(a, part, adam, s);
    case 7: return launch_tb_combine_dims<7>(a, part, adam, s);
    case 8: return launch_tb_combine_dims<8>(a, part, adam, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_grad_adam_tb(const GradArgs &a, const uint16_t *off, float *part, uint32_t dims, bool adam,
                               hipStream_t s) {
  switch (dims) {
    case 1: return launch_tb_dims<1>(a, off, part, adam, s);
    case 2: return launch_tb_dims<2>(a, off, part, adam, s);
    case 3: return launch_tb_dims<3>(a, off, part, adam, s);
    case 4: return launch_tb_dims<4>(a, off, part, adam, s);
    case 5: return launch_tb_dims<5>(a, off, part, adam, s);
    case 6: return launch_tb_dims<6>(a, off, part, adam, s);
    case 7: return launch_tb_dims<7>(a, off, part, adam, s);
    case 8: return launch_tb_dims<8>(a, off, part, adam, s);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// LDS-staged target-blocked gradient (sliced ELL).  Same ranges, per-lane
// entry order, butterfly and combine as grad_tb_kernel, so the result is
// bit-identical to it; what changes is where the other point comes from and
// how lanes are fed:
//  * a workgroup (range b = blockIdx.x % R, row chunk c = blockIdx.x / R,
//    R = 8 target ranges or 4 = the 8 merged in pairs) first copies range b
//    of Y and its chunk's own rows of Y into LDS, so every gather is an LDS
//    read instead of an L1/L2 request;
//  * the chunk's rows are sorted by their range-b entry count (descending,
//    ties by row) and cut into slices of 16 rows x 4 lanes; a slice stores its
//    entries step-major (step k holds entries 4k + lane of its 16 rows, 64
//    words, padded to the slice's longest row), so each step is
//    one coalesced 256-byte load and lanes of a wave do nearly equal work.
// Slice entries carry the target as an index relative to the range start.
// Padding words are far-pair entries without the owner bit whose target is
// a dummy point staged after the range at kSellFar in every coordinate: its
// coefficient k rc^2 underflows to zero, so a pad adds a signed zero to a
// gradient sum that is never -0 (it starts at +0, and round-to-nearest sums
// give -0 only from two -0 terms) and no loss.  Lanes therefore never
// branch on padding, and a wave's LDS reads of several steps are in flight
// together.  With R = 8 the result is bit-identical to grad_tb_kernel; with
// R = 4 the range partials are sums of two of its partials.
// ---------------------------------------------------------------------------
constexpr uint32_t kSellRows = 16;         // rows per slice (4 lanes each)
constexpr uint32_t kSellMaxChunk = 4096;   // rows per chunk (LDS bitonic sort)
constexpr uint32_t kSellThreads = 1024;
// LDS table of (c, k) by pair type (static LDS, at a fixed address): a
// shift, a mask and one broadcast LDS read replace the type's compares and
// selects.
constexpr uint32_t kSellTabBytes = 32;
// The far point and 3 spare slots after the range, so that the chunk's rows
// start on a 16-byte boundary of LDS.
constexpr uint32_t kSellFarSlots = 4;
__host__ __device__ __forceinline__ uint32_t sell_pad(const SellPlan &p) { return (2u << 30) | p.span; }

// Entries of SELL range b: tb ranges b g .. b g + g - 1 (g = 8 / R), which
// tb_partition stored contiguously.
__device__ __forceinline__ uint32_t sell_seg_beg(const uint16_t *__restrict__ off, uint32_t n, const SellPlan &p,
                                                 uint32_t b, uint32_t row) {
  return tb_seg_beg(off, n, b * p.group, row);
}
__device__ __forceinline__ uint32_t sell_seg_end(const uint32_t *__restrict__ row_ptr,
                                                 const uint16_t *__restrict__ off, uint32_t n, const SellPlan &p,
                                                 uint32_t b, uint32_t row) {
  return tb_seg_end(row_ptr, off, n, b * p.group + p.group - 1, row);
}

bool sell_plan(uint32_t n, uint32_t dims, uint32_t ranges, SellPlan &p) {
  p = SellPlan{};
  if (n < kTargetBlocks || dims == 0 || dims > 8 || (ranges != 4 && ranges != 8)) return false;
  p.ranges = ranges;
  p.group = kTargetBlocks / ranges;
  p.span = tb_span(n) * p.group;
  // two workgroups per CU when both fit in LDS, else one; chunks of at most
  // kSellMaxChunk rows (then more workgroups than CUs)
  for (uint32_t per_cu = 2; per_cu >= 1; --per_cu) {
    const uint32_t want = 256u * per_cu / ranges;  // chunks
    uint32_t rc = (n + want - 1) / want;
    rc = (rc + kSellRows - 1) / kSellRows * kSellRows;
    if (rc > kSellMaxChunk) rc = kSellMaxChunk;
    const uint64_t lds = (static_cast<uint64_t>(p.span) + kSellFarSlots + rc) * dims * 4u + kSellTabBytes;
    if (lds > kLdsPerCu / per_cu) continue;
    p.rc = rc;
    p.chunks = (n + rc - 1) / rc;
    p.smax = rc / kSellRows;
    p.lds_bytes = static_cast<uint32_t>(lds);
    return true;
  }
  return false;
}
uint64_t sell_slices(const SellPlan &p) { return static_cast<uint64_t>(p.ranges) * p.chunks * p.smax; }
uint32_t sell_blocks(const SellPlan &p) { return p.ranges * p.chunks; }
// The gradient pass runs every range of chunk c on XCD c % 8 (workgroups
// are dealt to the XCDs round-robin by index), so a chunk's range partials
// are all written to one L2, where the combine pass of the chunk (placed the
// same way) reads them.  Chunks are padded to a multiple of 8.
uint32_t sell_grad_blocks(const SellPlan &p) { return p.ranges * ((p.chunks + 7) / 8 * 8); }
__device__ __forceinline__ void sell_grad_block(const SellPlan &p, uint32_t blk, uint32_t &b, uint32_t &c) {
  const uint32_t x = blk % 8, k = blk / 8;
  b = k % p.ranges;
  c = (k / p.ranges) * 8 + x;
}

// One workgroup per (range, chunk): bitonic sort of (~len << 32 | slot) in
// LDS; writes the chunk's row order and each slice's entry count (16 x the
// longest row rounded up to 4).
__global__ __launch_bounds__(256) void sell_sort_kernel(const uint32_t *__restrict__ row_ptr,
                                                        const uint16_t *__restrict__ off, uint32_t n,
                                                        SellPlan p, uint32_t *__restrict__ perm,
                                                        uint32_t *__restrict__ ssz) {
  __shared__ uint64_t key[kSellMaxChunk];
  const uint32_t b = blockIdx.x % p.ranges, c = blockIdx.x / p.ranges;
  const uint32_t r0 = c * p.rc, rows = min(p.rc, n - r0);
  uint32_t P = kSellRows;
  while (P < rows) P <<= 1;
  for (uint32_t i = threadIdx.x; i < P; i += 256) {
    uint64_t k = ~0ull;
    if (i < rows) {
      const uint32_t row = r0 + i;
      const uint32_t len = sell_seg_end(row_ptr, off, n, p, b, row) - sell_seg_beg(off, n, p, b, row);
      k = (static_cast<uint64_t>(~len) << 32) | i;
    }
    key[i] = k;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P; i += 256) {
        const uint32_t ij = i ^ j;
        if (ij > i) {
          const uint64_t x = key[i], y = key[ij];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            key[i] = y;
            key[ij] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  uint32_t *pm = perm + static_cast<size_t>(b) * n + r0;
  for (uint32_t i = threadIdx.x; i < rows; i += 256) pm[i] = r0 + static_cast<uint32_t>(key[i]);
  uint32_t *sz = ssz + (static_cast<size_t>(b) * p.chunks + c) * p.smax;
  for (uint32_t sl = threadIdx.x; sl < p.smax; sl += 256) {
    uint32_t w = 0;
    if (sl * kSellRows < rows) {
      const uint32_t len = ~static_cast<uint32_t>(key[sl * kSellRows] >> 32);
      w = (len + 3u) & ~3u;
    }
    sz[sl] = kSellRows * w;
  }
}

// One thread per (range, row slot): copies the row's range-b segment of the
// partitioned CSR into its slice, step-major, and pads it to the slice width.
__global__ __launch_bounds__(256) void sell_fill_kernel(const uint32_t *__restrict__ row_ptr,
                                                        const uint32_t *__restrict__ tb_ent,
                                                        const uint16_t *__restrict__ off, uint32_t n,
                                                        SellPlan p, const uint32_t *__restrict__ perm,
                                                        const uint32_t *__restrict__ soff,
                                                        uint32_t *__restrict__ sent) {
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (gid >= static_cast<uint64_t>(p.ranges) * n) return;
  const uint32_t b = static_cast<uint32_t>(gid / n), slot = static_cast<uint32_t>(gid % n);
  const uint32_t c = slot / p.rc, i = slot % p.rc;
  const uint32_t row = perm[static_cast<size_t>(b) * n + slot];
  const uint32_t sl = i / kSellRows, j = i % kSellRows;
  const size_t gs = (static_cast<size_t>(b) * p.chunks + c) * p.smax + sl;
  const uint32_t base = soff[gs], width = (soff[gs + 1] - base) / kSellRows;
  const uint32_t sb = sell_seg_beg(off, n, p, b, row);
  const uint32_t len = sell_seg_end(row_ptr, off, n, p, b, row) - sb;
  const uint32_t *src = tb_ent + row_ptr[row] + sb;
  const uint32_t lo = b * p.span;
  for (uint32_t q = 0; q < width; ++q) {
    uint32_t v = sell_pad(p);
    if (q < len) {
      const uint32_t w = src[q];
      v = (w & ~kIdxMask) | ((w & kIdxMask) - lo);
    }
    sent[base + (q >> 2) * 64u + j * 4u + (q & 3u)] = v;
  }
}

hipError_t launch_sell_sort(const uint32_t *row_ptr, const uint16_t *off, uint32_t n, const SellPlan &p,
                            uint32_t *perm, uint32_t *ssz, hipStream_t s) {
  sell_sort_kernel<<<sell_blocks(p), 256, 0, s>>>(row_ptr, off, n, p, perm, ssz);
  return hipGetLastError();
}

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint64_t count, uint32_t v) {
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < count; e += gridDim.x * 256ull) p[e] = v;
}

hipError_t launch_sell_fill(const uint32_t *row_ptr, const uint32_t *tb_ent, const uint16_t *off, uint32_t n,
                            const SellPlan &p, const uint32_t *perm, const uint32_t *soff, uint32_t *sent,
                            hipStream_t s) {
  // every word starts as padding: rows missing from a chunk's last slice keep it
  if (p.words) {
    const uint64_t nb = (p.words + 255) / 256;
    fill_u32_kernel<<<static_cast<uint32_t>(nb < 16384 ? nb : 16384), 256, 0, s>>>(sent, p.words, sell_pad(p));
  }
  const uint64_t total = static_cast<uint64_t>(p.ranges) * n;
  sell_fill_kernel<<<static_cast<uint32_t>((total + 255) / 256), 256, 0, s>>>(row_ptr, tb_ent, off, n, p, perm,
                                                                              soff, sent);
  return hipGetLastError();
}

// Value of another lane of the same quad (DPP quad_perm, e.g. 0xB1 = xor 1,
// 0x4E = xor 2); every lane of the wave must be active.
template <int PERM>
__device__ __forceinline__ float quad_swap(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), PERM, 0xF, 0xF, false));
}


// U: steps of a slice loaded ahead (4 with 8 ranges, 8 with 4 ranges: about
// the longest rows of most slices)
template <int DIMS, bool LOSS, int U>
__global__ __launch_bounds__(kSellThreads) void grad_sell_kernel(const GradArgs a, const SellPlan p,
                                                                 const uint32_t *__restrict__ perm,
                                                                 const uint32_t *__restrict__ soff,
                                                                 const uint32_t *__restrict__ sent,
                                                                 float *__restrict__ part) {
  extern __shared__ __align__(16) float lds[];
  uint32_t b, c;
  sell_grad_block(p, blockIdx.x, b, c);
  if (c >= p.chunks) {  // padding workgroup
    if (LOSS && threadIdx.x == 0) a.loss_part[blockIdx.x] = 0.0f;
    return;
  }
  const uint32_t n = a.n;
  const uint32_t lo = b * p.span;
  const uint32_t nlo = lo < n ? min(p.span, n - lo) : 0u;
  const uint32_t r0 = c * p.rc, rows = min(p.rc, n - r0);
  __shared__ float2 sTab[kSellTabBytes / 8];                   // (c, k) by type
  float *sY = lds;                                                         // range b of Y, then the far point
  float *sR = lds + (static_cast<size_t>(p.span) + kSellFarSlots) * DIMS;  // the chunk's rows of Y
  const int t = threadIdx.x;
  {
    // lo and r0 are multiples of 4 points and the LDS areas start on
    // 16-byte boundaries, as stage_copy needs
    stage_copy<kSellThreads>(sY, a.y_in + static_cast<size_t>(lo) * DIMS, nlo * DIMS, t);
    stage_copy<kSellThreads>(sR, a.y_in + static_cast<size_t>(r0) * DIMS, rows * DIMS, t);
    if (t < DIMS) sY[static_cast<size_t>(p.span) * DIMS + t] = kSellFar;
  }
  const PairConsts pc = pair_consts(a.w);
  const float kn = pc.k_nn, km = pc.k_mn, kf = pc.k_fp, wn = pc.w_nn, wm = pc.w_mn, wf = pc.w_fp;
  if (t < 4) sTab[t] = t == 0 ? make_float2(10.0f, kn) : t == 1 ? make_float2(10000.0f, km) : make_float2(1.0f, kf);
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lane = t & 63, j = lane >> 2;
  const uint32_t *pm = perm + static_cast<size_t>(b) * n + r0;
  const uint32_t *so = soff + (static_cast<size_t>(b) * p.chunks + c) * p.smax;
  constexpr uint32_t kStride = kSellThreads / 64;  // slices between a wave's turns
  float loss = 0.0f;
  // software pipeline: the next slice's row, extent and first U steps of
  // entries are loaded before the current slice is computed.  Entry words
  // come through a buffer resource: the offsets are per-wave scalars plus
  // the lane, and reads past the array return 0 (steps past the slice are
  // never used: lim below is uniform over the wave).  Rows missing from a
  // chunk's last slice hold padding words.
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(sent), 0, static_cast<int>(p.words * 4u), 0x00020000);
  struct Slice {
    uint32_t row, beg, cnt;
    uint32_t w[U];
    bool valid;
  };
  const uint32_t last_sl = (rows - 1) / kSellRows;
  auto fetch = [&](uint32_t s2, Slice &q) {
    const uint32_t s2c = min(s2, last_sl);
    const uint32_t i = s2c * kSellRows + j;
    q.valid = i < rows;
    q.row = pm[min(i, rows - 1)];
    q.beg = so[s2c];
    q.cnt = s2 <= last_sl ? so[s2c + 1] - q.beg : 0u;
#pragma unroll
    for (int uu = 0; uu < U; ++uu)
      q.w[uu] = __builtin_amdgcn_raw_buffer_load_b32(rs, lane * 4, static_cast<int>((q.beg + 64u * uu) * 4u), 0);
  };
  auto compute = [&](const Slice &q) {
    float yr[DIMS], g[DIMS];
#pragma unroll
    for (int d = 0; d < DIMS; ++d) {
      yr[d] = sR[(q.row - r0) * DIMS + d];
      g[d] = 0.0f;
    }
    // steps G0 .. G0 + G - 1 of a round: their LDS reads are issued
    // together, then the terms computed in step order
    auto group = [&](const uint32_t (&ww)[U], auto g0c, auto gsize) __attribute__((always_inline)) {
      constexpr int G0 = decltype(g0c)::value, G = decltype(gsize)::value;
      float yv[G][DIMS];
      float2 ck[G];
#pragma unroll
      for (int uu = 0; uu < G; ++uu) {
        const uint32_t w = ww[G0 + uu];
        const float *yo = sY + (w & kIdxMask) * DIMS;
#pragma unroll
        for (int d = 0; d < DIMS; ++d) yv[uu][d] = yo[d];
        ck[uu] = sTab[__builtin_amdgcn_ubfe(w, 30, 2)];
      }
#pragma unroll
      for (int uu = 0; uu < G; ++uu)
        pair_term_ck<DIMS, LOSS>(ww[G0 + uu], ck[uu].x, ck[uu].y, yr, yv[uu], wn, wm, wf, g, loss);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    // steps G0 .. G0 + rest - 1 (rest < 4, uniform over the wave)
    auto tail = [&](const uint32_t (&ww)[U], auto g0c, uint32_t rest) __attribute__((always_inline)) {
      constexpr int G0 = decltype(g0c)::value;
      if (rest >= 2) {
        group(ww, g0c, I2{});
        if (rest == 3) group(ww, std::integral_constant<int, G0 + 2>{}, I1{});
      } else if (rest == 1) {
        group(ww, g0c, I1{});
      }
    };
    // the first lim steps (lim <= U): groups of 4, then 2, then 1
    auto round = [&](const uint32_t (&ww)[U], uint32_t lim) __attribute__((always_inline)) {
      if (lim < 4) {
        tail(ww, I0{}, lim);
        return;
      }
      group(ww, I0{}, I4{});
      if constexpr (U == 8) {
        if (lim == 8)
          group(ww, I4{}, I4{});
        else
          tail(ww, I4{}, lim - 4);
      }
    };
    round(q.w, min(static_cast<uint32_t>(U), q.cnt >> 6));
    for (uint32_t k0 = 64u * U; k0 < q.cnt; k0 += 64u * U) {  // rows longer than 4 U entries
      uint32_t xv[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu)
        xv[uu] = __builtin_amdgcn_raw_buffer_load_b32(rs, lane * 4, static_cast<int>((q.beg + k0 + 64u * uu) * 4u), 0);
      round(xv, min(static_cast<uint32_t>(U), (q.cnt - k0) >> 6));
    }
    // butterfly over the row's 4 lanes (a quad): xor 2, then xor 1
#pragma unroll
    for (int d = 0; d < DIMS; ++d) {
      g[d] = g[d] + quad_swap<0x4E>(g[d]);
      g[d] = g[d] + quad_swap<0xB1>(g[d]);
    }
    if ((lane & 3) == 0 && q.valid) {
#pragma unroll
      for (int d = 0; d < DIMS; ++d) part[(static_cast<size_t>(b) * n + q.row) * DIMS + d] = g[d];
    }
  };
  // two register sets in turn, so the prefetch needs no copies (and no wait
  // for the slice just requested)
  Slice qa, qb;
  // A chunk's slices are sorted by length, longest first; turn k of wave w
  // takes slice 16 k + w on even turns and 16 k + 15 - w on odd ones, so
  // the waves that draw the longest slice of one turn draw the shortest of
  // the next.  (Which wave computes a slice does not change its sums.)
  auto slice_of = [&](uint32_t k) { return k * kStride + ((k & 1u) ? kStride - 1u - wave : wave); };
  uint32_t turn = 0;
  if (slice_of(0) <= last_sl) fetch(slice_of(0), qa);
  while (slice_of(turn) <= last_sl) {
    fetch(slice_of(turn + 1), qb);
    compute(qa);
    ++turn;
    if (slice_of(turn) > last_sl) break;
    fetch(slice_of(turn + 1), qa);
    compute(qb);
    ++turn;
  }
  if (LOSS) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss = loss + __shfl_xor(loss, o, 64);
    __syncthreads();  // sY is no longer read: reuse it for the per-wave sums
    if (lane == 0) lds[wave] = loss;
    __syncthreads();
    if (t == 0) {
      float s = 0.0f;
      for (uint32_t w = 0; w < kSellThreads / 64; ++w) s = s + lds[w];
      a.loss_part[blockIdx.x] = s;
    }
  }
}

template <int DIMS>
static hipError_t launch_sell_dims(const GradArgs &a, const SellPlan &p, const uint32_t *perm,
                                   const uint32_t *soff, const uint32_t *sent, float *part, bool adam,
                                   hipStream_t s) {
  const uint32_t g1 = sell_grad_blocks(p);
  if (g1 == 0) return hipSuccess;
  const size_t shm = (static_cast<size_t>(p.span) + kSellFarSlots + p.rc) * DIMS * sizeof(float);
  const bool loss = a.loss_part != nullptr;
  if (p.ranges == 4) {
    if (loss)
      grad_sell_kernel<DIMS, true, 8><<<g1, kSellThreads, shm, s>>>(a, p, perm, soff, sent, part);
    else
      grad_sell_kernel<DIMS, false, 8><<<g1, kSellThreads, shm, s>>>(a, p, perm, soff, sent, part);
  } else {
    if (loss)
      grad_sell_kernel<DIMS, true, 4><<<g1, kSellThreads, shm, s>>>(a, p, perm, soff, sent, part);
    else
      grad_sell_kernel<DIMS, false, 4><<<g1, kSellThreads, shm, s>>>(a, p, perm, soff, sent, part);
  }
  const uint64_t count = static_cast<uint64_t>(a.n) * DIMS;
  const uint32_t chunk_elems = p.rc * DIMS, per_chunk = (chunk_elems + 255) / 256;
  const uint32_t g2 = per_chunk * ((p.chunks + 7) / 8 * 8);
  if (p.ranges == 4) {
    if (adam)
      grad_combine_elem_kernel<true, 4><<<g2, 256, 0, s>>>(a, part, count, chunk_elems, per_chunk);
    else
      grad_combine_elem_kernel<false, 4><<<g2, 256, 0, s>>>(a, part, count, chunk_elems, per_chunk);
  } else {
    if (adam)
      grad_combine_elem_kernel<true, 8><<<g2, 256, 0, s>>>(a, part, count, chunk_elems, per_chunk);
    else
      grad_combine_elem_kernel<false, 8><<<g2, 256, 0, s>>>(a, part, count, chunk_elems, per_chunk);
  }
  return hipGetLastError();
}

hipError_t launch_grad_adam_sell(const GradArgs &a, const SellPlan &p, const uint32_t *perm, const uint32_t *soff,
                                 const uint32_t *sent, float *part, uint32_t dims, bool adam, hipStream_t s) {
  switch (dims) {
    case 1: return launch_sell_dims<1>(a, p, perm, soff, sent, part, adam, s);
    case 2: return launch_sell_dims<2>(a, p, perm, soff, sent, part, adam, s);
    case 3: return launch_sell_dims<3>(a, p, perm, soff, sent, part, adam, s);
    case 4: return launch_sell_dims<4>(a, p, perm, soff, sent, part, adam, s);
    case 5: return launch_sell_dims<5>(a, p, perm, soff, sent, part, adam, s);
    case 6: return launch_sell_dims<6>(a, p, perm, soff, sent, part, adam, s);
    case 7: return launch_sell_dims<7>(a, p, perm, soff, sent, part, adam, s);
    case 8: return launch_sell_dims<8>(a, p, perm, soff, sent, part, adam, s);
    default: return hipErrorInvalidValue;
  }
}

uint32_t grad_blocks(uint32_t n) { return (n + kGradRowsPerBlock - 1) / kGradRowsPerBlock; }

template <int DIMS, int U>
static void launch_grad_u(const GradArgs &a, bool adam, uint32_t blocks, hipStream_t s) {
  const bool loss = a.loss_part != nullptr;
  if (adam) {
    if (loss)
      grad_adam_kernel<DIMS, true, true, U><<<blocks, 256, 0, s>>>(a);
    else
      grad_adam_kernel<DIMS, true, false, U><<<blocks, 256, 0, s>>>(a);
  } else {
    if (loss)
      grad_adam_kernel<DIMS, false, true, U><<<blocks, 256, 0, s>>>(a);
    else
      grad_adam_kernel<DIMS, false, false, U><<<blocks, 256, 0, s>>>(a);
  }
}

template <int DIMS>
static hipError_t launch_grad_dims(const GradArgs &a, bool adam, hipStream_t s) {
  const uint32_t blocks = grad_blocks(a.n);
  if (blocks == 0) return hipSuccess;
  // entries per lane loaded in one round: 4 when the mean row fits in 64
  // entries, else 8 (fewer idle slots on short rows)
  if (a.mean_row_len <= 64.0f)
    launch_grad_u<DIMS, 4>(a, adam, blocks, s);
  else
    launch_grad_u<DIMS, 8>(a, adam, blocks, s);
  return hipGetLastError();
}

hipError_t launch_grad_adam(const GradArgs &a, uint32_t dims, bool adam, hipStream_t s) {
  switch (dims) {
    case 1: return launch_grad_dims<1>(a, adam, s);
    case 2: return launch_grad_dims<2>(a, adam, s);
    case 3: return launch_grad_dims<3>(a, adam, s);
    case 4: return launch_grad_dims<4>(a, adam, s);
    case 5: return launch_grad_dims<5>(a, adam, s);
    case 6: return launch_grad_dims<6>(a, adam, s);
    case 7: return launch_grad_dims<7>(a, adam, s);
    case 8: return launch_grad_dims<8>(a, adam, s);
    default: return hipErrorInvalidValue;
  }
}

__global__ __launch_bounds__(256) void loss_reduce_kernel(const float *__restrict__ part,
                                                          uint32_t nb, float *__restrict__ loss) {
  __shared__ float sm[256];
  const float *p = part + static_cast<size_t>(blockIdx.x) * nb;
  float s = 0.0f;
  for (uint32_t i = threadIdx.x; i < nb; i += 256) s = s + p[i];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[blockIdx.x] = sm[0];
}

hipError_t launch_loss_reduce(const float *part, uint32_t nb, uint32_t iters, float *loss,
                              hipStream_t s) {
  if (iters == 0) return hipSuccess;
  loss_reduce_kernel<<<iters, 256, 0, s>>>(part, nb, loss);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// preprocessing (lib.rs:486-511)
// ---------------------------------------------------------------------------
__global__ void minmax_init_kernel(uint32_t *keys) {
  keys[0] = 0xffffffffu;
  keys[1] = 0u;
}

__global__ __launch_bounds__(256) void minmax_kernel(const float *__restrict__ x, uint64_t count,
                                                     uint32_t *keys) {
  uint32_t lo = 0xffffffffu, hi = 0u;
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < count; e += gridDim.x * 256ull) {
    const uint32_t k = tkey(x[e]);
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  __shared__ uint32_t slo[256], shi[256];
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + w]);
      shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMin(&keys[0], slo[0]);
    atomicMax(&keys[1], shi[0]);
  }
}

hipError_t launch_minmax(const float *x, uint64_t count, uint32_t *keys, hipStream_t s) {
  minmax_init_kernel<<<1, 1, 0, s>>>(keys);
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  minmax_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(x, count, keys);
  return hipGetLastError();
}

__global__ void normalize_kernel(const float *__restrict__ x, float *__restrict__ out,
                                 uint64_t count, const uint32_t *__restrict__ keys) {
  const float xmin = from_tkey(keys[0]), xmax = from_tkey(keys[1]);
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < count; e += gridDim.x * 256ull) {
    const float sh = x[e] - xmin;  // lib.rs:500
    out[e] = sh / xmax;            // lib.rs:503 (divides by the max, not the range)
  }
}

hipError_t launch_normalize(const float *x, float *out, uint64_t count, const uint32_t *keys,
                            hipStream_t s) {
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  normalize_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(x, out, count, keys);
  return hipGetLastError();
}

// mean_axis(Axis(0)) of a row-major array: rows added in order from zero,
// then divided by n (lib.rs:506-508).  One thread per column keeps that order;
// the column is a serial chain, so each thread issues the loads of 128 rows
// before adding them (in order) to keep enough memory requests in flight.
__global__ __launch_bounds__(64) void colmean_seq_kernel(const float *__restrict__ x, uint64_t n, uint32_t d,
                                                         float *__restrict__ mean) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d) return;
  constexpr int U = 128;
  float s = 0.0f;
  uint64_t r = 0;
  for (; r + U <= n; r += U) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = x[(r + u) * d + c];
#pragma unroll
    for (int u = 0; u < U; ++u) s = s + v[u];
  }
  for (; r < n; ++r) s = s + x[r * d + c];
  mean[c] = s / static_cast<float>(n);
}

hipError_t launch_colmean_seq(const float *x, uint64_t n, uint32_t d, float *mean, hipStream_t s) {
  if (d == 0) return hipSuccess;
  colmean_seq_kernel<<<(d + 63) / 64, 64, 0, s>>>(x, n, d, mean);
  return hipGetLastError();
}

__global__ void sub_mean_kernel(float *__restrict__ x, uint64_t count, uint32_t d,
                                const float *__restrict__ mean) {
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < count; e += gridDim.x * 256ull)
    x[e] = x[e] - mean[e % d];
}

hipError_t launch_sub_mean(float *x, uint64_t n, uint32_t d, const float *mean, hipStream_t s) {
  const uint64_t count = n * d;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  sub_mean_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(x, count, d, mean);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// PCA initialisation pieces: FP64 column means and covariance, FP32 projection
// ---------------------------------------------------------------------------
constexpr uint32_t kCovTile = 64;
// Resident slots of cov_part_kernel on MI355X (256 CUs x 4 blocks): each
// chunk's launch is given enough (upper tile, split) blocks to fill a round.
constexpr uint64_t kCovSlots = 1024;
constexpr uint64_t kCovChunkScratch = uint64_t{256} << 20;  // bytes of chunked covariance partials

// row splits of the column sums: enough blocks to cover the chip
static uint32_t colsum_splits(uint64_t n) {
  uint64_t s = (n + 127) / 128;
  if (s < 1) s = 1;
  if (s > 1024) s = 1024;
  return static_cast<uint32_t>(s);
}

uint64_t colmean_scratch_doubles(uint64_t n, uint32_t d) {
  return static_cast<uint64_t>(colsum_splits(n)) * d;
}

__global__ void colsum_part_kernel(const float *__restrict__ x, uint64_t n, uint32_t d,
                                   uint32_t splits, double *__restrict__ part) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sp = blockIdx.y;
  if (c >= d) return;
  const uint64_t r0 = n * sp / splits, r1 = n * (sp + 1) / splits;
  double s = 0.0;
  for (uint64_t r = r0; r < r1; ++r) s += static_cast<double>(x[r * d + c]);
  part[static_cast<size_t>(sp) * d + c] = s;
}

// Row chunks of the PCA statistics (the rows of one 32 MiB upload piece):
// chunk k = k0 + blockIdx.y / per holds rows [k cr, min(n, (k + 1) cr)), and
// split sp = blockIdx.y % per of it a fixed share of those rows; partial
// slot (k per + sp).  One launch over every chunk (X in device memory) and
// one launch per chunk as its upload lands give the same sums.
__device__ __forceinline__ void chunk_split_rows(uint64_t n, uint64_t cr, uint32_t k0, uint32_t per, uint64_t &r0,
                                                 uint64_t &r1, uint32_t &slot) {
  const uint32_t k = k0 + blockIdx.y / per, sp = blockIdx.y % per;
  const uint64_t c0 = static_cast<uint64_t>(k) * cr;
  const uint64_t rows = c0 < n ? (n - c0 < cr ? n - c0 : cr) : 0;
  r0 = c0 + rows * sp / per;
  r1 = c0 + rows * (sp + 1) / per;
  slot = k * per + sp;
}

__global__ void colsum_chunk_kernel(const float *__restrict__ x, uint64_t n, uint32_t d, uint64_t cr, uint32_t k0,
                                    uint32_t per, double *__restrict__ part) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d) return;
  uint64_t r0, r1;
  uint32_t slot;
  chunk_split_rows(n, cr, k0, per, r0, r1, slot);
  double s = 0.0;
  for (uint64_t r = r0; r < r1; ++r) s += static_cast<double>(x[r * d + c]);
  part[static_cast<size_t>(slot) * d + c] = s;
}

// 16 columns per block; 16 thread groups stride the splits, then a fixed-order
// LDS sum over the groups (deterministic).
__global__ __launch_bounds__(256) void colmean_fin_kernel(const double *__restrict__ part, uint32_t d,
                                                          uint32_t splits, uint64_t n,
                                                          double *__restrict__ mean) {
  __shared__ double red[16][17];
  const int cl = threadIdx.x & 15, g = threadIdx.x >> 4;
  const uint32_t c = blockIdx.x * 16 + cl;
  double s = 0.0;
  if (c < d)
    for (uint32_t sp = g; sp < splits; sp += 16) s += part[static_cast<size_t>(sp) * d + c];
  red[g][cl] = s;
  __syncthreads();
  if (g == 0 && c < d) {
    double tot = 0.0;
    for (int k = 0; k < 16; ++k) tot += red[k][cl];
    mean[c] = tot / static_cast<double>(n);
  }
}

hipError_t launch_colmean_f64(const float *x, uint64_t n, uint32_t d, double *mean,
                              double *scratch, hipStream_t s) {
  const uint32_t splits = colsum_splits(n);
  dim3 g((d + 63) / 64, splits);
  colsum_part_kernel<<<g, 64, 0, s>>>(x, n, d, splits, scratch);
  colmean_fin_kernel<<<(d + 15) / 16, 256, 0, s>>>(scratch, d, splits, n, mean);
  return hipGetLastError();
}

typedef double doublex4 __attribute__((ext_vector_type(4)));

// Partial covariance of one row split for one 64x64 tile (upper triangle of
// tiles only) on the FP64 matrix cores: each of the 4 waves owns a 32x32
// quarter as 2x2 v_mfma_f64_16x16x4 tiles (A = the tile's columns a, B = its
// columns b, K = rows).  Stages of 32 centred rows are staged in LDS as
// doubles; the next stage's floats are loaded into registers while the
// current one is multiplied.  Operand lane l holds element (l % 16, l / 16)
// of A (16 x 4) and (l / 16, l % 16) of B (4 x 16); the result lane l holds
// rows l / 16 + 4 v, v = 0..3, of column l % 16 (unlike the FP32 16x16
// products, whose lane l holds rows 4 (l / 16) + v).
__global__ __launch_bounds__(256) void cov_part_kernel(const float *__restrict__ x, uint64_t n,
                                                       uint32_t d, const double *__restrict__ mean,
                                                       uint32_t splits, double *__restrict__ part,
                                                       uint64_t cr, uint32_t k0) {
  // blockIdx.x enumerates the upper tiles (ti <= tj) row by row
  const uint32_t tiles = (d + kCovTile - 1) / kCovTile;
  uint32_t ti = 0, u = blockIdx.x;
  while (u >= tiles - ti) {
    u -= tiles - ti;
    ++ti;
  }
  const uint32_t tj = ti + u;
  constexpr int RB = 32;
  __shared__ double sA[RB][kCovTile];
  __shared__ double sB[RB][kCovTile];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const uint32_t a0 = ti * kCovTile, b0 = tj * kCovTile;
  // rows of this block's split of its chunk, and its partial slot
  uint64_t r0, r1;
  uint32_t sp;
  chunk_split_rows(n, cr, k0, splits, r0, r1, sp);
  // staging: thread t moves column cc of rows rq + 4 f (f < 8) of a stage
  const int cc = t & 63, rq = t >> 6;
  const uint32_t ca = a0 + cc, cb = b0 + cc;
  const bool va = ca < d, vb = cb < d;
  const double ma = va ? mean[ca] : 0.0, mb = vb ? mean[cb] : 0.0;
  float pa[8], pb[8];
  auto fetch = [&](uint64_t rb) {
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      const uint64_t r = rb + rq + 4 * f;
      pa[f] = (r < r1 && va) ? x[r * d + ca] : 0.0f;
      pb[f] = (r < r1 && vb) ? x[r * d + cb] : 0.0f;
    }
  };
  auto stash = [&](uint64_t rb) {
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      const uint64_t r = rb + rq + 4 * f;
      sA[rq + 4 * f][cc] = (r < r1 && va) ? static_cast<double>(pa[f]) - ma : 0.0;
      sB[rq + 4 * f][cc] = (r < r1 && vb) ? static_cast<double>(pb[f]) - mb : 0.0;
    }
  };
  doublex4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
  const uint64_t nst = (r1 - r0 + RB - 1) / RB;
  if (nst) fetch(r0);
  const int kq = lane >> 4, m = lane & 15;
  for (uint64_t st = 0; st < nst; ++st) {
    const uint64_t rb = r0 + st * RB;
    __syncthreads();  // the previous stage has been read
    stash(rb);
    __syncthreads();
    if (st + 1 < nst) fetch(rb + RB);
#pragma unroll
    for (int kk = 0; kk < RB / 4; ++kk) {
      const int kr = kk * 4 + kq;
      double av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = sA[kr][32 * wr + 16 * i + m];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = sB[kr][32 * wc + 16 * j + m];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
  }
  double *P = part + static_cast<size_t>(sp) * d * d;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t a = a0 + 32 * wr + 16 * i + kq + 4 * v, b = b0 + 32 * wc + 16 * j + m;
        if (a < d && b < d) P[static_cast<size_t>(a) * d + b] = acc[i][j][v];
      }
}

// Chunked statistics (see chunk_split_rows).  The covariance is taken about
// a shift (the first chunk's column mean) and corrected at the end:
//   sum (x - mu)(x - mu)^T = sum (x - h)(x - h)^T - n (mu - h)(mu - h)^T,
// which keeps the FP64 sums free of cancellation for data far from the
// origin, and lets each chunk be summed before the others are in memory.
CovChunks cov_chunks(uint64_t n, uint32_t d) {
  CovChunks c;
  c.rows = kCovChunkBytes / (static_cast<uint64_t>(d) * sizeof(float));
  if (c.rows == 0) c.rows = 1;
  c.chunks = static_cast<uint32_t>((n + c.rows - 1) / c.rows);
  if (c.chunks == 0) c.chunks = 1;
  // splits per chunk: enough (upper tile, split) blocks for a round of the
  // resident slots in each chunk's launch, within kCovChunkScratch bytes of
  // partials over all chunks
  const uint64_t tiles = (d + kCovTile - 1) / kCovTile, upper = tiles * (tiles + 1) / 2;
  const uint64_t want = (kCovSlots + upper - 1) / upper;
  const uint64_t cap = kCovChunkScratch / (static_cast<uint64_t>(c.chunks) * d * d * sizeof(double));
  uint64_t per = want < cap ? want : cap;
  if (per < 1) per = 1;
  if (per > 4096) per = 4096;
  c.per = static_cast<uint32_t>(per);
  c.col_per = colsum_splits(c.rows < n ? c.rows : n);
  return c;
}
uint64_t cov_chunk_scratch_doubles(const CovChunks &c, uint32_t d) {
  return static_cast<uint64_t>(c.chunks) * c.per * d * d;
}
uint64_t colsum_chunk_scratch_doubles(const CovChunks &c, uint32_t d) {
  return static_cast<uint64_t>(c.chunks) * c.col_per * d;
}

// 16 lanes per element: lane j sums slots j, j + 16, ... in order, then the
// 16 sums are added in lane order (a fixed order for any slot count).
__global__ __launch_bounds__(256) void cov_fin_shift_kernel(const double *__restrict__ part, uint32_t d,
                                                            uint32_t splits, const double *__restrict__ mean,
                                                            const double *__restrict__ shift, uint64_t n,
                                                            double *__restrict__ cov) {
  __shared__ double red[16][17];
  const uint32_t j = threadIdx.x & 15, g = threadIdx.x >> 4;
  const uint64_t e = blockIdx.x * 16ull + g;
  const uint64_t dd = static_cast<uint64_t>(d) * d;
  const uint32_t a = static_cast<uint32_t>(e / d), b = static_cast<uint32_t>(e % d);
  // lower tiles: written by their mirror
  const bool live = e < dd && b / kCovTile >= a / kCovTile;
  double s = 0.0;
  if (live)
    for (uint32_t sp = j; sp < splits; sp += 16) s += part[static_cast<size_t>(sp) * dd + e];
  red[g][j] = s;
  __syncthreads();
  if (!live || j != 0) return;
  double t = 0.0;
  for (int k = 0; k < 16; ++k) t += red[g][k];
  const double da = mean[a] - shift[a], db = mean[b] - shift[b];
  t = t - static_cast<double>(n) * (da * db);
  cov[e] = t;
  if (b / kCovTile > a / kCovTile) cov[static_cast<uint64_t>(b) * d + a] = t;
}

// Chunks k0 .. k0 + nk - 1: column-sum partials, then (k0 == 0) the shift
// from the first chunk's sums, then covariance partials about the shift.
hipError_t launch_pca_stats_chunks(const float *x, uint64_t n, uint32_t d, const CovChunks &c, uint32_t k0,
                                   uint32_t nk, double *colpart, double *shift, double *covpart, hipStream_t s) {
  if (nk == 0) return hipSuccess;
  colsum_chunk_kernel<<<dim3((d + 63) / 64, nk * c.col_per), 64, 0, s>>>(x, n, d, c.rows, k0, c.col_per, colpart);
  if (k0 == 0) {
    const uint64_t rows0 = c.rows < n ? c.rows : n;
    colmean_fin_kernel<<<(d + 15) / 16, 256, 0, s>>>(colpart, d, c.col_per, rows0, shift);
  }
  const uint32_t tiles = (d + kCovTile - 1) / kCovTile;
  cov_part_kernel<<<dim3(tiles * (tiles + 1) / 2, nk * c.per), 256, 0, s>>>(x, n, d, shift, c.per, covpart,
                                                                                   c.rows, k0);
  return hipGetLastError();
}

// After every chunk: the mean from all column-sum partials and the
// corrected covariance.
hipError_t launch_pca_stats_finish(uint64_t n, uint32_t d, const CovChunks &c, const double *colpart,
                                   const double *shift, const double *covpart, double *mean, double *cov,
                                   hipStream_t s) {
  colmean_fin_kernel<<<(d + 15) / 16, 256, 0, s>>>(colpart, d, c.chunks * c.col_per, n, mean);
  const uint64_t total = static_cast<uint64_t>(d) * d;
  cov_fin_shift_kernel<<<static_cast<uint32_t>((total + 15) / 16), 256, 0, s>>>(covpart, d, c.chunks * c.per, mean,
                                                                                 shift, n, cov);
  return hipGetLastError();
}

// out[r][c] = sum_f (x[r][f] - mean[f]) * V[f][c]; 64x64 output tile per block.
__global__ __launch_bounds__(256) void project_kernel(const float *__restrict__ x, uint64_t n,
                                                      uint32_t d, const float *__restrict__ mean,
                                                      const float *__restrict__ V, uint32_t q,
                                                      float *__restrict__ out) {
  constexpr int TB = 64, KB = 16;
  __shared__ float sX[KB][TB + 1];
  __shared__ float sV[KB][TB];
  const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
  const uint64_t r0 = static_cast<uint64_t>(blockIdx.x) * TB;
  const uint32_t c0 = blockIdx.y * TB;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
  for (uint32_t f0 = 0; f0 < d; f0 += KB) {
    __syncthreads();
    for (int li = t; li < TB * KB; li += 256) {
      const int rr = li / KB, ff = li % KB;
      const uint64_t r = r0 + rr;
      const uint32_t f = f0 + ff;
      sX[ff][rr] = (r < n && f < d) ? x[r * d + f] - mean[f] : 0.0f;
    }
    for (int li = t; li < TB * KB; li += 256) {
      const int ff = li / TB, cc = li % TB;
      const uint32_t f = f0 + ff, c = c0 + cc;
      sV[ff][cc] = (f < d && c < q) ? V[static_cast<size_t>(f) * q + c] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int ff = 0; ff < KB; ++ff) {
      float xv[4], vv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) xv[i] = sX[ff][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) vv[j] = sV[ff][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(xv[i], vv[j], acc[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t r = r0 + ty * 4 + i;
      const uint32_t c = c0 + tx * 4 + j;
      if (r < n && c < q) out[r * q + c] = acc[i][j];
    }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// The same product on the FP32 matrix cores for q <= 16 CN: a block owns 64
// rows and every column; wave w owns rows 16 w .. 16 w + 15 as CN
// v_mfma_f32_16x16x4f32 tiles.  Operand lane l holds element (l % 16, l / 16)
// of A (16 x 4) and (l / 16, l % 16) of B (4 x 16); the result lane l holds
// rows 4 (l / 16) + v, v = 0..3, of column l % 16.  A 16-feature block is
// consumed in 4 steps; step s takes feature f + 4 (l / 16) + s, so each lane's
// A values for a block are one 16-byte load of its row and each B value set
// one 16-byte LDS read.  V and the mean are first packed, stage by stage, into
// exactly the LDS image a stage needs (project_pack_kernel), so staging is a
// straight copy of kProjStride float4 per KC = 32 features, loaded into
// registers one stage ahead like the row values.
constexpr int kProjKC = 32;
template <int CN>
constexpr int proj_stride() {
  return (kProjKC / 16) * CN * 64 + kProjKC / 4;  // V slots, then the mean
}

uint64_t project_workspace_floats(uint32_t d, uint32_t q) {
  const uint32_t cn = (q + 15) / 16;
  if (cn == 0 || cn > 8) return 0;
  const uint64_t stages = (d + kProjKC - 1) / kProjKC;
  return stages * ((kProjKC / 16) * cn * 64 + kProjKC / 4) * 4;
}

// Packs stage st's slot e: slots below KBLK CN 64 hold V[f + s][c] for s =
// 0..3 with (block kb, column tile j, lane l) = e's digits, f = 32 st + 16 kb
// + 4 (l / 16), c = 16 j + l % 16; the last KC / 4 slots hold the mean.
// Features >= d and columns >= q are zero.
__global__ void project_pack_kernel(const float *__restrict__ V, const float *__restrict__ mean,
                                    uint32_t d, uint32_t q, uint32_t cn, uint64_t total,
                                    float4 *__restrict__ packed) {
  const uint64_t e = blockIdx.x * 256ull + threadIdx.x;
  if (e >= total) return;
  const uint32_t vslots = (kProjKC / 16) * cn * 64, stride = vslots + kProjKC / 4;
  const uint32_t st = static_cast<uint32_t>(e / stride), slot = static_cast<uint32_t>(e % stride);
  float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (slot < vslots) {
    const uint32_t kb = slot / (cn * 64), rem = slot % (cn * 64);
    const uint32_t c = 16 * (rem / 64) + (rem & 15);
    const uint32_t fb = st * kProjKC + kb * 16 + 4 * ((rem & 63) >> 4);
    for (int s2 = 0; s2 < 4; ++s2)
      if (c < q && fb + s2 < d) v[s2] = V[static_cast<size_t>(fb + s2) * q + c];
  } else {
    const uint32_t fm = st * kProjKC + 4 * (slot - vslots);
    for (int s2 = 0; s2 < 4; ++s2)
      if (fm + s2 < d) v[s2] = mean[fm + s2];
  }
  packed[e] = make_float4(v[0], v[1], v[2], v[3]);
}

template <int CN, bool V4>
__global__ __launch_bounds__(256) void project_mfma_kernel(const float *__restrict__ x, uint64_t n,
                                                           uint32_t d,
                                                           const float4 *__restrict__ packed,
                                                           uint32_t q, float *__restrict__ out) {
  constexpr int KBLK = kProjKC / 16, SLOTS = KBLK * CN * 64, STRIDE = proj_stride<CN>();
  constexpr int NS = (STRIDE + 255) / 256;
  __shared__ float4 sV[STRIDE];
  const float4 *sM = sV + SLOTS;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = lane & 15, g = lane >> 4;
  const uint64_t rbase = static_cast<uint64_t>(blockIdx.x) * 64 + wave * 16;
  const uint64_t row = rbase + m;
  const bool rv = row < n;
  const float *xr = x + (rv ? row : 0) * d;
  float4 pv[NS], px[KBLK];
  auto fetch = [&](uint32_t st) __attribute__((always_inline)) {
    const float4 *src = packed + static_cast<size_t>(st) * STRIDE;
#pragma unroll
    for (int i = 0; i < NS; ++i)  // only the last copy can fall outside the stage
      pv[i] = (256 * (i + 1) <= STRIDE || t + 256 * i < STRIDE)
                  ? src[t + 256 * i]
                  : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t f0 = st * kProjKC;
#pragma unroll
    for (int kb = 0; kb < KBLK; ++kb) {
      const uint32_t fb = f0 + kb * 16 + 4 * g;
      if (V4) {
        px[kb] = (rv && fb < d) ? *reinterpret_cast<const float4 *>(xr + fb)
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      } else {
        float e[4];
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) e[s2] = (rv && fb + s2 < d) ? xr[fb + s2] : 0.0f;
        px[kb] = make_float4(e[0], e[1], e[2], e[3]);
      }
    }
  };
  floatx4 acc[CN];
#pragma unroll
  for (int j = 0; j < CN; ++j) acc[j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  float4 xa[KBLK];
  const uint32_t stages = (d + kProjKC - 1) / kProjKC;
  fetch(0);
  for (uint32_t st = 0; st < stages; ++st) {
    __syncthreads();  // the previous stage has been read
#pragma unroll
    for (int i = 0; i < NS; ++i)
      if (256 * (i + 1) <= STRIDE || t + 256 * i < STRIDE) sV[t + 256 * i] = pv[i];
#pragma unroll
    for (int kb = 0; kb < KBLK; ++kb) xa[kb] = px[kb];
    __syncthreads();
    if (st + 1 < stages) fetch(st + 1);
#pragma unroll
    for (int kb = 0; kb < KBLK; ++kb) {
      if (st * kProjKC + kb * 16 >= d) break;  // uniform: the tail stage's empty block
      float4 b[CN];
#pragma unroll
      for (int j = 0; j < CN; ++j) b[j] = sV[(kb * CN + j) * 64 + lane];
      const float4 mv = sM[kb * 4 + g];
      const float a0 = xa[kb].x - mv.x, a1 = xa[kb].y - mv.y;
      const float a2 = xa[kb].z - mv.z, a3 = xa[kb].w - mv.w;
      // step outer, column tile inner: consecutive products use different
      // accumulators
#pragma unroll
      for (int j = 0; j < CN; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b[j].x, acc[j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < CN; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b[j].y, acc[j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < CN; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b[j].z, acc[j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < CN; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a3, b[j].w, acc[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < CN; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint64_t r = rbase + 4 * g + v;
      const uint32_t c = 16 * j + m;
      if (r < n && c < q) out[r * q + c] = acc[j][v];
    }
}

template <int CN>
static void launch_project_mfma(const float *x, uint64_t n, uint32_t d, const float4 *packed,
                                uint32_t q, float *out, hipStream_t s) {
  const dim3 g(static_cast<uint32_t>((n + 63) / 64));
  const bool v4 = d % 4 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0;
  if (v4)
    project_mfma_kernel<CN, true><<<g, 256, 0, s>>>(x, n, d, packed, q, out);
  else
    project_mfma_kernel<CN, false><<<g, 256, 0, s>>>(x, n, d, packed, q, out);
}

hipError_t launch_project(const float *x, uint64_t n, uint32_t d, const float *mean,
                          const float *V, uint32_t q, float *workspace, float *out,
                          hipStream_t s) {
  if (n == 0 || q == 0) return hipSuccess;
  const uint32_t cn = (q + 15) / 16;
  if (cn > 8) {
    dim3 g(static_cast<uint32_t>((n + 63) / 64), (q + 63) / 64);
    project_kernel<<<g, 256, 0, s>>>(x, n, d, mean, V, q, out);
    return hipGetLastError();
  }
  float4 *packed = reinterpret_cast<float4 *>(workspace);
  const uint64_t total = project_workspace_floats(d, q) / 4;
  project_pack_kernel<<<static_cast<uint32_t>((total + 255) / 256), 256, 0, s>>>(V, mean, d, q, cn,
                                                                                 total, packed);
  switch (cn) {
    case 1: launch_project_mfma<1>(x, n, d, packed, q, out, s); break;
    case 2: launch_project_mfma<2>(x, n, d, packed, q, out, s); break;
    case 3: launch_project_mfma<3>(x, n, d, packed, q, out, s); break;
    case 4: launch_project_mfma<4>(x, n, d, packed, q, out, s); break;
    case 5: launch_project_mfma<5>(x, n, d, packed, q, out, s); break;
    case 6: launch_project_mfma<6>(x, n, d, packed, q, out, s); break;
    case 7: launch_project_mfma<7>(x, n, d, packed, q, out, s); break;
    default: launch_project_mfma<8>(x, n, d, packed, q, out, s); break;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// small utilities: Adam alone (adam.rs:51-59), zero fill, FP64 GEMM
// ---------------------------------------------------------------------------
__global__ void adam_kernel(float *__restrict__ y, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, uint64_t count,
                            float lr_t, float omb1, float omb2) {
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < count; e += gridDim.x * 256ull) {
    const float gv = g[e];
    float mm = m[e], vv = v[e];
    mm = mm + omb1 * (gv - mm);
    const float gg = gv * gv;
    vv = vv + omb2 * (gg - vv);
    const float num = lr_t * mm;
    const float den = sqrtf(vv) + 1e-7f;
    y[e] = y[e] - num / den;
    m[e] = mm;
    v[e] = vv;
  }
}

hipError_t launch_adam(float *y, const float *g, float *m, float *v, uint64_t count, float lr_t,
                       float omb1, float omb2, hipStream_t s) {
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  adam_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(y, g, m, v, count, lr_t, omb1, omb2);
  return hipGetLastError();
}

__global__ void fill_zero_kernel(uint32_t *__restrict__ p, uint64_t words) {
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < words; e += gridDim.x * 256ull) p[e] = 0u;
}

// dst (n x cols) = src[:, :cols] * scale (src has row stride ld)
__global__ void take_cols_kernel(const float *__restrict__ src, uint64_t n, uint32_t ld, uint32_t cols,
                                 float scale, float *__restrict__ dst) {
  const uint64_t total = n * cols;
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < total; e += gridDim.x * 256ull) {
    const uint64_t r = e / cols;
    dst[e] = src[r * ld + (e - r * cols)] * scale;
  }
}

hipError_t launch_take_cols(const float *src, uint64_t n, uint32_t ld, uint32_t cols, float scale,
                            float *dst, hipStream_t s) {
  uint64_t blocks = (n * cols + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  take_cols_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(src, n, ld, cols, scale, dst);
  return hipGetLastError();
}

// Plain streaming copy (16-byte vectors, grid stride): the reference workload
// for HBM bandwidth and for calibrating the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void hbm_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                       uint64_t n16) {
  const uint64_t stride = gridDim.x * 256ull;
  uint64_t e = blockIdx.x * 256ull + threadIdx.x;
  for (; e + 3 * stride < n16; e += 4 * stride) {  // four loads in flight per thread
    const uint4 v0 = src[e], v1 = src[e + stride], v2 = src[e + 2 * stride], v3 = src[e + 3 * stride];
    dst[e] = v0;
    dst[e + stride] = v1;
    dst[e + 2 * stride] = v2;
    dst[e + 3 * stride] = v3;
  }
  for (; e < n16; e += stride) dst[e] = src[e];
}

hipError_t launch_hbm_copy(void *dst, const void *src, uint64_t bytes, hipStream_t s) {
  if (bytes % 16 || reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
    return hipErrorInvalidValue;
  const uint64_t n16 = bytes / 16;
  uint64_t blocks = (n16 + 1023) / 1024;
  if (blocks > 16384) blocks = 16384;
  if (blocks == 0) return hipSuccess;
  hbm_copy_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(static_cast<const uint4 *>(src),
                                                                static_cast<uint4 *>(dst), n16);
  return hipGetLastError();
}

hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t s) {
  const uint64_t words = bytes / 4;
  uint64_t blocks = (words + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  fill_zero_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(static_cast<uint32_t *>(p), words);
  return hipGetLastError();
}

}  // namespace pm
