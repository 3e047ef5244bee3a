// sampling.cpp -- host pair sampling, the reference arithmetic the host needs,
// and the symmetrised CSR build.
//
// Pair sampling stays on the host with the reference's RNG so pair lists follow
// the reference's streams: rand 0.8.5 SmallRng (Xoshiro256++ on 64-bit
// targets) seeded by rand_core 0.6's default seed_from_u64 (PCG32 expansion),
// and UniformInt<u32>::sample_single (widening multiply, zone rejection).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <random>

#include "host_common.h"
#include "internal.h"

namespace pm {

int host_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  if (h > 16) h = 16;  // one GPU's share of host cores
  return static_cast<int>(h);
}

void parallel_for(uint64_t total, const std::function<void(uint64_t, uint64_t)> &fn, int threads) {
  if (threads <= 0) threads = host_threads();
  if (total < 2 || threads <= 1) {
    fn(0, total);
    return;
  }
  if (static_cast<uint64_t>(threads) > total) threads = static_cast<int>(total);
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (int t = 0; t < threads; ++t) {
    const uint64_t b = total * t / threads, e = total * (t + 1) / threads;
    pool.emplace_back([&fn, b, e] { fn(b, e); });
  }
  for (auto &th : pool) th.join();
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void find_weights(float w_mn_init, uint64_t itr, uint64_t p1, uint64_t p2, float out[3]) {
  if (itr < p1) {  // weights.rs:57-63
    const float progress = static_cast<float>(itr) / static_cast<float>(p1);
    const float a = (1.0f - progress) * w_mn_init;
    const float b = progress * 3.0f;
    out[0] = 2.0f;
    out[1] = a + b;
    out[2] = 1.0f;
  } else if (itr < p1 + p2) {  // weights.rs:64-71
    out[0] = 3.0f;
    out[1] = 3.0f;
    out[2] = 1.0f;
  } else {  // weights.rs:72-79
    out[0] = 1.0f;
    out[1] = 0.0f;
    out[2] = 1.0f;
  }
}

static float powi_f32(float a, int32_t b) {
  const bool recip = b < 0;
  float r = 1.0f;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

float adam_lr_t(float lr, float beta1, float beta2, uint64_t itr) {
  const int32_t t = static_cast<int32_t>(itr + 1);
  const float num = std::sqrt(1.0f - powi_f32(beta2, t));
  const float lr_num = lr * num;
  return lr_num / (1.0f - powi_f32(beta1, t));
}

float euclid_ref(const float *a, const float *b, uint64_t d) {
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const uint64_t chunks = d / 8;
  for (uint64_t c = 0; c < chunks; ++c)
    for (int l = 0; l < 8; ++l) {
      const float df = a[8 * c + l] - b[8 * c + l];
      const float sq = df * df;
      acc[l] = acc[l] + sq;
    }
  float s = 0.0f;
  for (int l = 0; l < 8; ++l) s = s + acc[l];
  for (uint64_t e = chunks * 8; e < d; ++e) {
    const float df = a[e] - b[e];
    const float sq = df * df;
    s = s + sq;
  }
  return std::sqrt(s);
}

void decide_num_pairs(uint64_t n, bool has_override, uint64_t override_k, float mn_ratio,
                      float fp_ratio, uint64_t out[3]) {
  uint64_t k;
  if (has_override) {
    k = override_k;
  } else if (n <= 10000) {
    k = 10;
  } else {
    const float lg = std::log10(static_cast<float>(n));
    const float v = 10.0f + 15.0f * (lg - 4.0f);
    const float r = std::round(v);
    k = r <= 0.0f ? 0 : static_cast<uint64_t>(r);
  }
  const float mn = std::round(static_cast<float>(k) * mn_ratio);
  const float fp = std::round(static_cast<float>(k) * fp_ratio);
  out[0] = k;
  out[1] = mn > 0.0f ? static_cast<uint64_t>(mn) : 0;  // `as usize` saturates
  out[2] = fp > 0.0f ? static_cast<uint64_t>(fp) : 0;
  if (k < 1) fail(PACMAP_ERR_INVALID_NEIGHBOR_COUNT, "The number of nearest neighbors can't be less than 1");
  if (out[2] < 1) fail(PACMAP_ERR_INVALID_FAR_POINT_COUNT, "The number of far points can't be less than 1");
}

// ---------------------------------------------------------------------------
// rand 0.8.5 SmallRng
// ---------------------------------------------------------------------------
static uint32_t pcg32_step(uint64_t &state) {
  const uint64_t kMul = 6364136223846793005ull, kInc = 11634580027462260723ull;
  state = state * kMul + kInc;
  const uint64_t st = state;
  const uint32_t xorshifted = static_cast<uint32_t>(((st >> 18) ^ st) >> 27);
  const uint32_t rot = static_cast<uint32_t>(st >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

SmallRng SmallRng::seed_from_u64(uint64_t state) {
  uint8_t seed[32];
  for (int c = 0; c < 8; ++c) {
    const uint32_t x = pcg32_step(state);
    for (int b = 0; b < 4; ++b) seed[4 * c + b] = static_cast<uint8_t>(x >> (8 * b));
  }
  SmallRng r;
  bool zero = true;
  for (int b = 0; b < 32; ++b) zero = zero && seed[b] == 0;
  if (zero) {  // Xoshiro256PlusPlus::from_seed maps an all-zero seed to SplitMix64(0)
    uint64_t st = 0;
    for (int w = 0; w < 4; ++w) {
      st += 0x9e3779b97f4a7c15ull;
      uint64_t z = st;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      r.s[w] = z ^ (z >> 31);
    }
    return r;
  }
  for (int w = 0; w < 4; ++w) {
    uint64_t v = 0;
    for (int b = 7; b >= 0; --b) v = (v << 8) | seed[8 * w + b];
    r.s[w] = v;
  }
  return r;
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t SmallRng::next_u64() {
  const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return result;
}

uint32_t SmallRng::gen_range(uint32_t low, uint32_t high) {
  const uint32_t range = high - low;
  if (range == 0) return next_u32();
  const uint32_t zone = (range << __builtin_clz(range)) - 1u;
  for (;;) {
    const uint32_t v = next_u32();
    const uint64_t m = static_cast<uint64_t>(v) * range;
    if (static_cast<uint32_t>(m) <= zone) return low + static_cast<uint32_t>(m >> 32);
  }
}

double SmallRng::next_f64() { return static_cast<double>(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

// sampling.rs:37-61
static uint64_t sample_fp(uint64_t n_samples, uint32_t maximum, const uint32_t *reject,
                          uint64_t n_reject, uint64_t stride, uint32_t self_ind, SmallRng &rng,
                          uint32_t *out) {
  bool self_absent = true;
  for (uint64_t q = 0; q < n_reject; ++q)
    if (reject[q * stride] == self_ind) self_absent = false;
  uint64_t avail = maximum;
  avail = avail > n_reject ? avail - n_reject : 0;
  avail = avail > (self_absent ? 1u : 0u) ? avail - (self_absent ? 1u : 0u) : 0;
  n_samples = std::min(n_samples, avail);
  uint64_t len = 0;
  while (len < n_samples) {
    const uint32_t j = rng.gen_range(0, maximum);
    if (j == self_ind) continue;
    bool dup = false;
    for (uint64_t q = 0; q < len && !dup; ++q) dup = out[q] == j;
    for (uint64_t q = 0; q < n_reject && !dup; ++q) dup = reject[q * stride] == j;
    if (!dup) out[len++] = j;
  }
  return len;
}

static uint64_t entropy_seed() {
  std::random_device rd;
  return (static_cast<uint64_t>(rd()) << 32) ^ rd();
}

void sample_fp_pairs(uint64_t n, const uint32_t *pair_neighbors, uint64_t n_neighbors,
                     uint64_t n_fp, const uint64_t *seed, uint32_t *pair_fp) {
  std::memset(pair_fp, 0, sizeof(uint32_t) * 2 * n * n_fp);
  const uint64_t base = seed ? *seed : entropy_seed();
  parallel_for(n, [&](uint64_t b, uint64_t e) {
    std::vector<uint32_t> buf(n_fp ? n_fp : 1);
    for (uint64_t i = b; i < e; ++i) {
      // seeded: SmallRng::seed_from_u64(seed + i) (sampling.rs:97); unseeded
      // (thread_rng, sampling.rs:176) draws an independent stream per point
      SmallRng rng = SmallRng::seed_from_u64(seed ? base + i : base ^ (0x9e3779b97f4a7c15ull * (i + 1)));
      const uint64_t got =
          sample_fp(n_fp, static_cast<uint32_t>(n), pair_neighbors + 2 * i * n_neighbors + 1,
                    n_neighbors, 2, static_cast<uint32_t>(i), rng, buf.data());
      for (uint64_t q = 0; q < got; ++q) {
        pair_fp[2 * (i * n_fp + q)] = static_cast<uint32_t>(i);
        pair_fp[2 * (i * n_fp + q) + 1] = buf[q];
      }
    }
  });
}

static inline uint32_t total_key(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

void sample_mn_pairs(const float *x, uint64_t n, uint64_t d, uint64_t n_mn, const uint64_t *seed,
                     uint32_t *pair_mn) {
  std::memset(pair_mn, 0, sizeof(uint32_t) * 2 * n * n_mn);
  const uint64_t base = seed ? *seed : entropy_seed();
  parallel_for(n, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      SmallRng rng = SmallRng::seed_from_u64(seed ? base + i : base ^ (0x9e3779b97f4a7c15ull * (i + 1)));
      uint32_t *rows = pair_mn + 2 * i * n_mn;
      for (uint64_t j = 0; j < n_mn; ++j) {
        uint32_t sampled[6];
        const uint64_t got = sample_fp(6, static_cast<uint32_t>(n), rows + 1, j, 2,
                                       static_cast<uint32_t>(i), rng, sampled);
        float dist[6] = {0, 0, 0, 0, 0, 0};
        uint32_t idx[6] = {0, 0, 0, 0, 0, 0};
        for (uint64_t q = 0; q < got; ++q) {
          dist[q] = euclid_ref(x + i * d, x + static_cast<uint64_t>(sampled[q]) * d, d);
          idx[q] = sampled[q];
        }
        // sampling.rs:308: order by total_cmp; equal distances keep slot order
        for (int a = 1; a < 6; ++a) {
          const float td = dist[a];
          const uint32_t ti = idx[a];
          int c = a - 1;
          while (c >= 0 && total_key(dist[c]) > total_key(td)) {
            dist[c + 1] = dist[c];
            idx[c + 1] = idx[c];
            --c;
          }
          dist[c + 1] = td;
          idx[c + 1] = ti;
        }
        rows[2 * j] = static_cast<uint32_t>(i);
        rows[2 * j + 1] = idx[1];
      }
    }
  });
}

// ---------------------------------------------------------------------------
// symmetrised CSR: every non-self pair (i, j) of type t appears in row i as
// (j, t, owner) and in row j as (i, t).  Row order follows the global pair
// order (NN, then MN, then FP, each in array order), which fixes the
// summation order of every row.
// ---------------------------------------------------------------------------
void build_csr(uint64_t n, const PairList lists[3], Csr &out) {
  if (n >= kMaxPoints) fail(PACMAP_ERR_UNSUPPORTED, "too many points for the CSR entry format");
  uint64_t total_pairs = 0;
  for (int t = 0; t < 3; ++t) total_pairs += lists[t].rows;
  const int T = host_threads();
  const uint64_t nt = std::max<uint64_t>(1, std::min<uint64_t>(T, total_pairs / 65536 + 1));
  auto pair_at = [&](uint64_t p, int &type) -> const uint32_t * {
    for (int t = 0; t < 3; ++t) {
      if (p < lists[t].rows) {
        type = t;
        return lists[t].p + 2 * p;
      }
      p -= lists[t].rows;
    }
    type = -1;
    return nullptr;
  };
  std::vector<std::vector<uint32_t>> cnt(nt, std::vector<uint32_t>(n, 0));
  std::atomic<bool> bad{false};
  std::atomic<uint64_t> bad_pair{0};
  parallel_for(nt, [&](uint64_t tb, uint64_t te) {
    for (uint64_t t = tb; t < te; ++t) {
      const uint64_t p0 = total_pairs * t / nt, p1 = total_pairs * (t + 1) / nt;
      auto &c = cnt[t];
      for (uint64_t p = p0; p < p1; ++p) {
        int type;
        const uint32_t *pr = pair_at(p, type);
        const uint32_t i = pr[0], j = pr[1];
        if (i >= n || j >= n) {
          bad = true;
          bad_pair = p;
          return;
        }
        if (i == j) continue;  // gradient.rs:116-118
        ++c[i];
        ++c[j];
      }
    }
  }, static_cast<int>(nt));
  if (bad) fail(PACMAP_ERR_INVALID_PAIR_INDEX, "pair index out of range at pair " + std::to_string(bad_pair.load()));
  out.row_ptr.assign(n + 1, 0);
  uint64_t acc = 0;
  for (uint64_t r = 0; r < n; ++r) {
    out.row_ptr[r] = static_cast<uint32_t>(acc);
    uint64_t deg = 0;
    for (uint64_t t = 0; t < nt; ++t) deg += cnt[t][r];
    acc += deg;
    if (acc >= 0xffffffffull) fail(PACMAP_ERR_UNSUPPORTED, "too many pairs for 32-bit CSR offsets");
  }
  out.row_ptr[n] = static_cast<uint32_t>(acc);
  // per-thread start offsets inside each row (turn counts into cursors)
  parallel_for(n, [&](uint64_t rb, uint64_t re) {
    for (uint64_t r = rb; r < re; ++r) {
      uint32_t cur = out.row_ptr[r];
      for (uint64_t t = 0; t < nt; ++t) {
        const uint32_t c = cnt[t][r];
        cnt[t][r] = cur;
        cur += c;
      }
    }
  });
  out.ent.assign(acc, 0);
  parallel_for(nt, [&](uint64_t tb, uint64_t te) {
    for (uint64_t t = tb; t < te; ++t) {
      const uint64_t p0 = total_pairs * t / nt, p1 = total_pairs * (t + 1) / nt;
      auto &c = cnt[t];
      for (uint64_t p = p0; p < p1; ++p) {
        int type;
        const uint32_t *pr = pair_at(p, type);
        const uint32_t i = pr[0], j = pr[1];
        if (i == j) continue;
        const uint32_t tb30 = static_cast<uint32_t>(type) << 30;
        out.ent[c[i]++] = j | tb30 | kOwnerBit;
        out.ent[c[j]++] = i | tb30;
      }
    }
  }, static_cast<int>(nt));
}

}  // namespace pm
