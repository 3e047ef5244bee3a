// pairs.hip -- mid-near / far pair sampling and the symmetrised CSR on the GPU.
//
// Sampling follows sampling.rs:37-142 and 296-313 with rand 0.8.5's SmallRng
// (Xoshiro256++ seeded through rand_core 0.6's PCG32 seed_from_u64) and
// UniformInt<u32>::sample_single (widening multiply, zone rejection), one
// thread per point: point i's streams depend only on (seed, i), so every
// point is independent.  MN distances use distance.rs:50-75's order.
//
// The CSR holds every non-self pair in both endpoint rows.  Rows are filled
// with integer atomics and then each row is sorted by entry value, which
// makes the layout (and so every row's summation order) independent of the
// atomic order.
#include "internal.h"

#include <algorithm>

#include <cstdint>

namespace pm {

struct DevRng {
  uint64_t s[4];
};

__device__ __forceinline__ uint64_t rotl64d(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__device__ inline DevRng rng_seed_from_u64(uint64_t state) {
  uint32_t w[8];
  for (int c = 0; c < 8; ++c) {
    state = state * 6364136223846793005ull + 11634580027462260723ull;
    const uint64_t st = state;
    const uint32_t xs = static_cast<uint32_t>(((st >> 18) ^ st) >> 27);
    const uint32_t rot = static_cast<uint32_t>(st >> 59);
    w[c] = (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  DevRng r;
  bool zero = true;
  for (int c = 0; c < 8; ++c) zero = zero && w[c] == 0;
  if (zero) {
    uint64_t st = 0;
    for (int q = 0; q < 4; ++q) {
      st += 0x9e3779b97f4a7c15ull;
      uint64_t z = st;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      r.s[q] = z ^ (z >> 31);
    }
    return r;
  }
  for (int q = 0; q < 4; ++q) r.s[q] = static_cast<uint64_t>(w[2 * q]) | (static_cast<uint64_t>(w[2 * q + 1]) << 32);
  return r;
}

__device__ __forceinline__ uint64_t rng_next_u64(DevRng &r) {
  const uint64_t result = rotl64d(r.s[0] + r.s[3], 23) + r.s[0];
  const uint64_t t = r.s[1] << 17;
  r.s[2] ^= r.s[0];
  r.s[3] ^= r.s[1];
  r.s[1] ^= r.s[2];
  r.s[0] ^= r.s[3];
  r.s[2] ^= t;
  r.s[3] = rotl64d(r.s[3], 45);
  return result;
}

__device__ __forceinline__ uint32_t rng_gen_range(DevRng &r, uint32_t range) {
  // [0, range), range >= 1
  const uint32_t zone = (range << __clz(range)) - 1u;
  for (;;) {
    const uint32_t v = static_cast<uint32_t>(rng_next_u64(r) >> 32);
    const uint64_t m = static_cast<uint64_t>(v) * range;
    if (static_cast<uint32_t>(m) <= zone) return static_cast<uint32_t>(m >> 32);
  }
}

// sampling.rs:37-61: n_samples distinct draws from [0, maximum) avoiding self
// and the reject list (reject[q * stride], q < n_reject); returns the count.
__device__ uint32_t dev_sample_fp(uint32_t n_samples, uint32_t maximum, const uint32_t *reject,
                                  uint32_t n_reject, uint32_t stride, uint32_t self_ind, DevRng &rng,
                                  uint32_t *out, uint32_t out_stride) {
  bool self_absent = true;
  for (uint32_t q = 0; q < n_reject; ++q)
    if (reject[q * stride] == self_ind) self_absent = false;
  uint64_t avail = maximum;
  avail = avail > n_reject ? avail - n_reject : 0;
  const uint32_t sa = self_absent ? 1u : 0u;
  avail = avail > sa ? avail - sa : 0;
  if (n_samples > avail) n_samples = static_cast<uint32_t>(avail);
  uint32_t len = 0;
  while (len < n_samples) {
    const uint32_t j = rng_gen_range(rng, maximum);
    if (j == self_ind) continue;
    bool dup = false;
    for (uint32_t q = 0; q < len && !dup; ++q) dup = out[q * out_stride] == j;
    for (uint32_t q = 0; q < n_reject && !dup; ++q) dup = reject[q * stride] == j;
    if (!dup) out[(len++) * out_stride] = j;
  }
  return len;
}

__device__ __forceinline__ uint32_t tkey_p(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Seeds: seeded runs use seed + i (sampling.rs:97, 133); unseeded runs draw
// one base seed per call and decorrelate points by an odd multiplier.
__device__ __forceinline__ uint64_t point_seed(uint64_t base, int seeded, uint64_t i) {
  return seeded ? base + i : (base ^ (0x9e3779b97f4a7c15ull * (i + 1)));
}

// sampling.rs:77-103: one thread per point; rows (i, j) of pair_fp, unfilled
// rows stay (0, 0) (callers zero the array).
__global__ void sample_fp_kernel(uint32_t n, const uint32_t *__restrict__ pair_nn, uint32_t knn,
                                 uint32_t n_fp, uint64_t seed, int seeded,
                                 uint32_t *__restrict__ pair_fp) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DevRng rng = rng_seed_from_u64(point_seed(seed, seeded, i));
  uint32_t *rows = pair_fp + static_cast<size_t>(2) * i * n_fp;
  const uint32_t got = dev_sample_fp(n_fp, n, pair_nn + static_cast<size_t>(2) * i * knn + 1, knn, 2, i,
                                     rng, rows + 1, 2);
  for (uint32_t q = 0; q < got; ++q) rows[2 * q] = i;
}

// The MN sampler's random streams, precomputed: point i's stream (seeded as
// in sampling.rs:133) is only ever used through rng_gen_range, which takes
// the upper 32 bits of each output, so the first kMnStream(n_mn) of those are
// written to vals[i][.] and the state after them to states[i] (a point that
// needs more continues from there).  One thread per point.
__host__ __device__ constexpr uint32_t mn_stream_len(uint32_t n_mn) {
  return (6 * n_mn + 16 + 7) / 8 * 8;  // six draws a step, a margin for rejections
}

__global__ void mn_stream_kernel(uint32_t n, uint32_t nv, uint64_t seed, int seeded,
                                 uint32_t *__restrict__ vals, uint64_t *__restrict__ states) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DevRng rng = rng_seed_from_u64(point_seed(seed, seeded, i));
  uint4 *dst = reinterpret_cast<uint4 *>(vals + static_cast<size_t>(i) * nv);
  for (uint32_t k = 0; k < nv / 4; ++k) {
    uint4 w;
    w.x = static_cast<uint32_t>(rng_next_u64(rng) >> 32);
    w.y = static_cast<uint32_t>(rng_next_u64(rng) >> 32);
    w.z = static_cast<uint32_t>(rng_next_u64(rng) >> 32);
    w.w = static_cast<uint32_t>(rng_next_u64(rng) >> 32);
    dst[k] = w;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) states[static_cast<size_t>(i) * 4 + q] = rng.s[q];
}

// sampling.rs:118-142 + 296-313: for each j, six distinct draws avoiding self
// and the earlier picks, then the second closest by distance (ties keep draw
// order; unfilled slots are (0.0, 0)).  Eight lanes per point.
//
// Draws: the group reads its stream eight values at a time (lane l holds
// value base + l) and takes them in order by shuffle, so the eight lanes see
// the sequence sampling.rs:37-61 consumes.  A candidate is tested against
// the draws of this step (lane q holds draw q) and the earlier picks (lane l
// holds picks l, l + 8, l + 16, l + 24) in parallel and the group's lanes
// are combined by ballot.  n_mn > 32 keeps the picks in the pair rows and
// tests them eight at a time.
//
// Distances: the rows come in the packed kNN layout (feature 8c + l at
// segment l, then the d mod 8 tail): lane l sums its segment in order of c
// with float4 loads of all six drawn rows in flight together, the partials
// are added in lane order 0..7, then the tail, exactly as distance.rs:50-75.
constexpr uint32_t kMnRegPicks = 32;

__global__ __launch_bounds__(256) void sample_mn_kernel(const float *__restrict__ xp, uint32_t n, uint32_t Cp,
                                                        uint32_t R, uint32_t Dp, uint32_t n_mn,
                                                        const uint32_t *__restrict__ vals,
                                                        uint64_t *__restrict__ states,
                                                        uint32_t *__restrict__ pair_mn) {
  const uint32_t pt = threadIdx.x >> 3;
  const uint32_t i = blockIdx.x * 32 + pt;
  const int lane = threadIdx.x & 63, l = lane & 7, g0 = lane & ~7;
  if (i >= n) return;  // the eight lanes of a point leave together
  const uint32_t nv = mn_stream_len(n_mn);
  const uint32_t *sv = vals + static_cast<size_t>(i) * nv;
  uint32_t buf = sv[l], base = 0, pos = 0;
  auto next_value = [&]() -> uint32_t {
    if (pos >= nv) {  // past the precomputed values (rare): continue the stream kept in states[i]
      uint64_t *st = states + static_cast<size_t>(i) * 4;
      DevRng rng;
#pragma unroll
      for (int q = 0; q < 4; ++q) rng.s[q] = st[q];
      const uint32_t v = static_cast<uint32_t>(rng_next_u64(rng) >> 32);
#pragma unroll
      for (int q = 0; q < 4; ++q) st[q] = rng.s[q];  // the group's lanes store the same state
      ++pos;
      return v;
    }
    if (pos - base >= 8) {
      base += 8;
      buf = sv[base + l];  // nv is a multiple of 8
    }
    const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(buf), g0 | static_cast<int>(pos - base), 64));
    ++pos;
    return v;
  };
  const uint32_t zone = (n << __clz(n)) - 1u;  // rng_gen_range(n)
  auto group_any = [&](bool p) -> bool { return ((__ballot(p) >> g0) & 0xffull) != 0; };
  uint32_t *rows = pair_mn + static_cast<size_t>(2) * i * n_mn;
  const bool in_reg = n_mn <= kMnRegPicks;
  uint32_t pk0 = 0, pk1 = 0, pk2 = 0, pk3 = 0;  // picks l, l + 8, l + 16, l + 24
  // is `c` among the first `cnt` picks?
  auto picked = [&](uint32_t c, uint32_t cnt) -> bool {
    if (in_reg) {
      const uint32_t ul = static_cast<uint32_t>(l);
      return group_any((ul < cnt && pk0 == c) || (ul + 8 < cnt && pk1 == c) || (ul + 16 < cnt && pk2 == c) ||
                       (ul + 24 < cnt && pk3 == c));
    }
    bool hit = false;
    for (uint32_t q0 = 0; q0 < cnt; q0 += 8) {
      const uint32_t q = q0 + static_cast<uint32_t>(l);
      hit = hit || (q < cnt && rows[2 * q + 1] == c);
    }
    return group_any(hit);
  };
  const float *xi = xp + static_cast<size_t>(i) * Dp;
  const float4 *own = reinterpret_cast<const float4 *>(xi + static_cast<size_t>(l) * Cp);
  const uint32_t c4n = Cp / 4, tail = 8u * Cp;
  for (uint32_t j = 0; j < n_mn; ++j) {
    // sampling.rs:37-61 with n_samples = 6, the reject list = the j picks
    const bool self_absent = !picked(i, j);
    uint64_t avail = n;
    avail = avail > j ? avail - j : 0;
    const uint32_t sa = self_absent ? 1u : 0u;
    avail = avail > sa ? avail - sa : 0;
    const uint32_t want = avail < 6 ? static_cast<uint32_t>(avail) : 6u;
    uint32_t got = 0, mine = 0;  // lane q < got holds draw q
    while (got < want) {
      uint32_t c;
      for (;;) {
        const uint64_t m = static_cast<uint64_t>(next_value()) * n;
        if (static_cast<uint32_t>(m) <= zone) {
          c = static_cast<uint32_t>(m >> 32);
          break;
        }
      }
      if (c == i) continue;
      const bool dup = group_any(static_cast<uint32_t>(l) < got && mine == c) || picked(c, j);
      if (!dup) {
        if (static_cast<uint32_t>(l) == got) mine = c;
        ++got;
      }
    }
    uint32_t sampled[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) sampled[q] = static_cast<uint32_t>(__shfl(static_cast<int>(mine), g0 | q, 64));
    const float4 *seg[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {  // unfilled slots read the point's own row (unused)
      const uint32_t r = static_cast<uint32_t>(q) < got ? sampled[q] : i;
      seg[q] = reinterpret_cast<const float4 *>(xp + static_cast<size_t>(r) * Dp + static_cast<size_t>(l) * Cp);
    }
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t c4 = 0; c4 < c4n; ++c4) {
      const float4 av = own[c4];
      float4 bv[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) bv[q] = seg[q][c4];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const float as[4] = {av.x, av.y, av.z, av.w}, bs[4] = {bv[q].x, bv[q].y, bv[q].z, bv[q].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float df = as[e] - bs[e];
          const float sq = df * df;
          acc[q] = acc[q] + sq;
        }
      }
    }
    // the tail: lane e < R holds the squared differences of feature 8 Cp + e
    float tq[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (static_cast<uint32_t>(l) < R) {
      const float xe = xi[tail + l];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const float df = xe - xp[static_cast<size_t>(q < static_cast<int>(got) ? sampled[q] : i) * Dp + tail + l];
        tq[q] = df * df;
      }
    }
    float dist[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t idx[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      float sd = 0.0f;
#pragma unroll
      for (int k = 0; k < 8; ++k) sd = sd + __shfl(acc[q], g0 | k, 64);
      for (uint32_t e = 0; e < R; ++e) sd = sd + __shfl(tq[q], g0 | static_cast<int>(e), 64);
      if (static_cast<uint32_t>(q) < got) {
        dist[q] = sqrtf(sd);
        idx[q] = sampled[q];
      }
    }
    for (int p = 1; p < 6; ++p) {  // insertion sort by total order: stable
      const float td = dist[p];
      const uint32_t ti = idx[p];
      int c = p - 1;
      while (c >= 0 && tkey_p(dist[c]) > tkey_p(td)) {
        dist[c + 1] = dist[c];
        idx[c + 1] = idx[c];
        --c;
      }
      dist[c + 1] = td;
      idx[c + 1] = ti;
    }
    if (l == 0) {
      rows[2 * j] = i;
      rows[2 * j + 1] = idx[1];
    }
    if (in_reg && static_cast<uint32_t>(l) == (j & 7)) {
      const uint32_t slot = j >> 3;
      pk0 = slot == 0 ? idx[1] : pk0;
      pk1 = slot == 1 ? idx[1] : pk1;
      pk2 = slot == 2 ? idx[1] : pk2;
      pk3 = slot == 3 ? idx[1] : pk3;
    }
    if (!in_reg) __threadfence_block();  // the row written above is read back by picked()
  }
}

uint64_t mn_stream_words(uint32_t n, uint32_t n_mn) {
  return n_mn ? static_cast<uint64_t>(n) * (mn_stream_len(n_mn) + 8) : 0;
}

hipError_t launch_sample_pairs(const float *xp, const KnnLayout &L, uint32_t n, const uint32_t *pair_nn,
                               uint32_t knn, uint32_t n_mn, uint32_t n_fp, uint64_t seed, int seeded,
                               uint32_t *pair_mn, uint32_t *pair_fp, uint32_t *stream_ws, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (n + 63) / 64;
  if (n_fp) sample_fp_kernel<<<blocks, 64, 0, s>>>(n, pair_nn, knn, n_fp, seed, seeded, pair_fp);
  if (n_mn) {
    const uint32_t nv = mn_stream_len(n_mn);
    uint64_t *states = reinterpret_cast<uint64_t *>(stream_ws + static_cast<size_t>(n) * nv);
    mn_stream_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, nv, seed, seeded, stream_ws, states);
    sample_mn_kernel<<<(n + 31) / 32, 256, 0, s>>>(xp, n, L.Cp, L.R, L.Dp, n_mn, stream_ws, states, pair_mn);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// symmetrised CSR
// ---------------------------------------------------------------------------
struct PairArrays {
  const uint32_t *p[3];
  uint64_t rows[3];
};

__device__ __forceinline__ bool pair_at(const PairArrays &pa, uint64_t g, uint32_t &i, uint32_t &j,
                                        uint32_t &type) {
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    if (g < pa.rows[t]) {
      i = pa.p[t][2 * g];
      j = pa.p[t][2 * g + 1];
      type = static_cast<uint32_t>(t);
      return true;
    }
    g -= pa.rows[t];
  }
  return false;
}

// Pair arrays are row-major by their first index (i), so the lanes of a
// wave mostly hold runs of equal i.  The i side is counted (and given its
// slots) once per run by the run's last lane; the j side, random, by one
// atomic per pair.  run_info: head = first lane of the lane's run, tail =
// last lane (valid lanes only; key 0xffffffff marks an invalid lane).
__device__ __forceinline__ void lane_run(uint32_t key, int lane, int &head, int &tail) {
  const uint32_t up = __shfl_up(key, 1, 64), dn = __shfl_down(key, 1, 64);
  const bool is_head = lane == 0 || up != key;
  const bool is_tail = lane == 63 || dn != key;
  const uint64_t H = __ballot(is_head), T = __ballot(is_tail);
  const uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  head = 63 - __clzll(H & le);
  tail = __ffsll(static_cast<long long>(T & ~((1ull << lane) - 1ull))) - 1;
}

// degree count and index validation (bad pair -> flag, pair index kept)
__global__ __launch_bounds__(256) void csr_degree_kernel(PairArrays pa, uint64_t total, uint32_t n,
                                                         uint32_t *__restrict__ deg, uint32_t *__restrict__ bad) {
  const int lane = threadIdx.x & 63;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t g0 = blockIdx.x * 256ull + (threadIdx.x & ~63u); g0 < total; g0 += stride) {
    const uint64_t g = g0 + lane;
    uint32_t i = 0, j = 0, t = 0;
    bool valid = g < total;
    if (valid) {
      pair_at(pa, g, i, j, t);
      if (i >= n || j >= n) {
        atomicOr(bad, 1u);
        valid = false;
      } else if (i == j) {
        valid = false;  // gradient.rs:116-118
      }
    }
    int head, tail;
    lane_run(valid ? i : 0xffffffffu, lane, head, tail);
    if (valid && tail == lane) atomicAdd(&deg[i], static_cast<uint32_t>(lane - head + 1));
    if (valid) atomicAdd(&deg[j], 1u);
  }
}

__global__ __launch_bounds__(256) void csr_fill_kernel(PairArrays pa, uint64_t total, uint32_t n,
                                                       uint32_t *__restrict__ cursor, uint32_t *__restrict__ ent) {
  const int lane = threadIdx.x & 63;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t g0 = blockIdx.x * 256ull + (threadIdx.x & ~63u); g0 < total; g0 += stride) {
    const uint64_t g = g0 + lane;
    uint32_t i = 0, j = 0, t = 0;
    bool valid = g < total;
    if (valid) {
      pair_at(pa, g, i, j, t);
      valid = i < n && j < n && i != j;
    }
    int head, tail;
    lane_run(valid ? i : 0xffffffffu, lane, head, tail);
    uint32_t base = 0;
    if (valid && tail == lane) base = atomicAdd(&cursor[i], static_cast<uint32_t>(lane - head + 1));
    base = __shfl(base, tail < 0 ? lane : tail, 64);
    const uint32_t tb = t << 30;
    if (valid) {
      ent[base + static_cast<uint32_t>(lane - head)] = j | tb | kOwnerBit;
      ent[atomicAdd(&cursor[j], 1u)] = i | tb;
    }
  }
}

// exclusive scan of n counts (three phases, 1024 elements per block)
__global__ __launch_bounds__(256) void scan_block_kernel(const uint32_t *__restrict__ in, uint32_t n,
                                                         uint32_t *__restrict__ out,
                                                         uint32_t *__restrict__ block_sums) {
  __shared__ uint32_t sm[1024];
  const uint32_t base = blockIdx.x * 1024;
  for (int q = 0; q < 4; ++q) {
    const uint32_t e = base + threadIdx.x * 4 + q;
    sm[threadIdx.x * 4 + q] = e < n ? in[e] : 0u;
  }
  __syncthreads();
  uint32_t local[4], run = 0;
  for (int q = 0; q < 4; ++q) {
    local[q] = run;
    run += sm[threadIdx.x * 4 + q];
  }
  __shared__ uint32_t ts[256];
  ts[threadIdx.x] = run;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {  // Hillis-Steele inclusive scan of thread totals
    const uint32_t v = threadIdx.x >= static_cast<unsigned>(off) ? ts[threadIdx.x - off] : 0u;
    __syncthreads();
    ts[threadIdx.x] += v;
    __syncthreads();
  }
  const uint32_t prefix = threadIdx.x ? ts[threadIdx.x - 1] : 0u;
  for (int q = 0; q < 4; ++q) {
    const uint32_t e = base + threadIdx.x * 4 + q;
    if (e < n) out[e] = prefix + local[q];
  }
  if (threadIdx.x == 255) block_sums[blockIdx.x] = ts[255];
}

__global__ void scan_sums_kernel(uint32_t *__restrict__ sums, uint32_t nb, uint32_t *__restrict__ total) {
  if (threadIdx.x != 0) return;
  uint32_t run = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    const uint32_t v = sums[b];
    sums[b] = run;
    run += v;
  }
  *total = run;
}

__global__ void scan_add_kernel(uint32_t *__restrict__ out, uint32_t n, const uint32_t *__restrict__ sums,
                                const uint32_t *__restrict__ total, uint32_t *__restrict__ cursor) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e < n) {
    out[e] += sums[e / 1024];
    cursor[e] = out[e];
  }
  if (e == n) out[n] = *total;
}

template <int R>
__device__ __forceinline__ void wave_sort_u32(uint32_t (&v)[R], int lane) {
  constexpr int N = 64 * R;
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 64) {
        const int rs = stride >> 6;
#pragma unroll
        for (int r = 0; r < R; ++r)
          if ((r & rs) == 0) {
            const bool asc = (((r * 64 + lane) & size) == 0);
            const uint32_t a = v[r], b = v[r | rs];
            const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
            v[r] = asc ? lo : hi;
            v[r | rs] = asc ? hi : lo;
          }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t o = __shfl_xor(v[r], stride, 64);
          const bool asc = (((r * 64 + lane) & size) == 0);
          const bool lower = (lane & stride) == 0;
          const uint32_t lo = v[r] < o ? v[r] : o, hi = v[r] < o ? o : v[r];
          v[r] = (lower == asc) ? lo : hi;
        }
      }
    }
  }
}

// one wavefront per row of length <= 64 * R (longer rows are left for the
// next size class; rows above the largest class are appended to long_list
// for csr_sort_long_kernel)
template <int R>
__global__ __launch_bounds__(256) void csr_sort_rows_kernel(const uint32_t *__restrict__ row_ptr, uint32_t n,
                                                            uint32_t min_len, uint32_t *__restrict__ ent,
                                                            uint32_t *__restrict__ long_count,
                                                            uint32_t *__restrict__ long_list, uint32_t long_cap) {
  const int lane = threadIdx.x & 63;
  const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const uint32_t b = row_ptr[row], len = row_ptr[row + 1] - b;
  if (len <= min_len || len < 2) return;
  if (len > 64u * R) {
    if (R == 32 && lane == 0) {
      const uint32_t slot = atomicAdd(long_count, 1u);
      if (slot < long_cap) long_list[slot] = row;
    }
    return;
  }
  uint32_t v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = r * 64 + lane;
    v[r] = e < len ? ent[b + e] : 0xffffffffu;
  }
  wave_sort_u32<R>(v, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = r * 64 + lane;
    if (e < len) ent[b + e] = v[r];
  }
}

// Ascending bitonic sort of a[0, len) by one workgroup, len padded to a
// power of two N with virtual +infinity entries.  Each stage of size S first
// compares i with its mirror in the S-block (S-1-offset), then runs the
// half-cleaners of stride S/4 .. 1; every comparison puts the smaller value
// at the lower index, so a partner at or beyond len (+infinity) never moves
// and is skipped.  a is in LDS or global memory; the barriers order both.
template <class P>
__device__ void block_bitonic_sort(P a, uint32_t len) {
  uint32_t N = 1;
  while (N < len) N <<= 1;
  const uint32_t pairs = N >> 1;
  for (uint32_t size = 2; size <= N; size <<= 1) {
    const uint32_t hs = __builtin_ctz(size) - 1;  // log2(size / 2)
    for (uint32_t p = threadIdx.x; p < pairs; p += blockDim.x) {
      const uint32_t blk = p >> hs, off = p & ((1u << hs) - 1u);
      const uint32_t i = (blk << (hs + 1)) + off, j = (blk << (hs + 1)) + size - 1u - off;
      if (j < len) {
        const uint32_t x = a[i], y = a[j];
        if (x > y) {
          a[i] = y;
          a[j] = x;
        }
      }
    }
    __syncthreads();
    for (uint32_t st = size >> 2; st > 0; st >>= 1) {
      const uint32_t ls = __builtin_ctz(st);
      for (uint32_t p = threadIdx.x; p < pairs; p += blockDim.x) {
        const uint32_t blk = p >> ls, off = p & (st - 1u);
        const uint32_t i = (blk << (ls + 1)) + off, j = i + st;
        if (j < len) {
          const uint32_t x = a[i], y = a[j];
          if (x > y) {
            a[i] = y;
            a[j] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Rows longer than the wave sorts take: one workgroup per listed row (the
// list count is read on the device, the grid strides over it).  A row of up
// to kLongRowLds entries is sorted in LDS, a longer one in place.
constexpr uint32_t kLongRowLds = 32768;
__global__ __launch_bounds__(1024) void csr_sort_long_kernel(const uint32_t *__restrict__ row_ptr,
                                                             const uint32_t *__restrict__ long_count,
                                                             const uint32_t *__restrict__ long_list, uint32_t long_cap,
                                                             uint32_t *__restrict__ ent) {
  extern __shared__ uint32_t srow[];
  const uint32_t cnt = min(*long_count, long_cap);
  for (uint32_t q = blockIdx.x; q < cnt; q += gridDim.x) {
    const uint32_t row = long_list[q];
    const uint32_t b = row_ptr[row], len = row_ptr[row + 1] - b;
    if (len <= kLongRowLds) {
      for (uint32_t e = threadIdx.x; e < len; e += blockDim.x) srow[e] = ent[b + e];
      __syncthreads();
      block_bitonic_sort(srow, len);
      for (uint32_t e = threadIdx.x; e < len; e += blockDim.x) ent[b + e] = srow[e];
    } else {
      block_bitonic_sort(ent + b, len);
    }
    __syncthreads();
  }
}

static uint32_t grid_for(uint64_t total) {
  uint64_t b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  return static_cast<uint32_t>(b ? b : 1);
}

hipError_t launch_csr_count(const uint32_t *const p[3], const uint64_t rows[3], uint32_t n,
                            uint32_t *deg, uint32_t *bad, hipStream_t s) {
  PairArrays pa;
  uint64_t total = 0;
  for (int t = 0; t < 3; ++t) {
    pa.p[t] = p[t];
    pa.rows[t] = rows[t];
    total += rows[t];
  }
  if (total) csr_degree_kernel<<<grid_for(total), 256, 0, s>>>(pa, total, n, deg, bad);
  return hipGetLastError();
}

hipError_t launch_csr_scan(const uint32_t *deg, uint32_t n, uint32_t *row_ptr, uint32_t *cursor,
                           uint32_t *scratch, hipStream_t s) {
  const uint32_t nb = (n + 1023) / 1024;
  scan_block_kernel<<<nb ? nb : 1, 256, 0, s>>>(deg, n, row_ptr, scratch);
  scan_sums_kernel<<<1, 64, 0, s>>>(scratch, nb, scratch + nb);
  scan_add_kernel<<<(n + 1 + 255) / 256, 256, 0, s>>>(row_ptr, n, scratch, scratch + nb, cursor);
  return hipGetLastError();
}

hipError_t launch_csr_fill_sort(const uint32_t *const p[3], const uint64_t rows[3], uint32_t n,
                                const uint32_t *row_ptr, uint32_t *cursor, uint32_t *ent,
                                uint32_t *long_count, uint32_t *long_list, uint32_t long_cap, hipStream_t s) {
  PairArrays pa;
  uint64_t total = 0;
  for (int t = 0; t < 3; ++t) {
    pa.p[t] = p[t];
    pa.rows[t] = rows[t];
    total += rows[t];
  }
  if (total) csr_fill_kernel<<<grid_for(total), 256, 0, s>>>(pa, total, n, cursor, ent);
  const uint32_t blocks = (n + 3) / 4;
  csr_sort_rows_kernel<1><<<blocks, 256, 0, s>>>(row_ptr, n, 0, ent, long_count, long_list, long_cap);
  csr_sort_rows_kernel<4><<<blocks, 256, 0, s>>>(row_ptr, n, 64, ent, long_count, long_list, long_cap);
  csr_sort_rows_kernel<32><<<blocks, 256, 0, s>>>(row_ptr, n, 256, ent, long_count, long_list, long_cap);
  const uint32_t g = std::min<uint32_t>(256u, std::max<uint32_t>(1u, long_cap));
  csr_sort_long_kernel<<<g, 1024, kLongRowLds * sizeof(uint32_t), s>>>(row_ptr, long_count, long_list, long_cap, ent);
  return hipGetLastError();
}

}  // namespace pm
