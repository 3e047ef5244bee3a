// sampling.cpp -- the reference arithmetic the host needs: the pair-count
// policy, the weight schedule, Adam's learning rate, and rand 0.8.5's SmallRng
// (Xoshiro256++ on 64-bit targets, seeded by rand_core 0.6's default
// seed_from_u64: PCG32 expansion) with UniformInt<u32>::sample_single and
// rand_distr's StandardNormal.  The pair samplers built on these streams run
// on the GPU (pairs.hip).
#include <algorithm>
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

// A value of `mant52` (52 mantissa bits) scaled into [2^e, 2^(e+1)).
static double f64_with_exponent(uint64_t mant52, int e) {
  const uint64_t b = mant52 | (static_cast<uint64_t>(1023 + e) << 52);
  double v;
  std::memcpy(&v, &b, sizeof(v));
  return v;
}

// Ziggurat layers of N(0, 1) as rand_distr 0.4.3 tabulates them: Doornik's
// ZIGNOR construction with 256 layers of area V under exp(-x^2 / 2), the
// base layer at R including the tail.
namespace {
constexpr double kZigR = 3.6541528853610088, kZigV = 0.00492867323399;
struct ZigTables {
  double x[257], f[257];
  ZigTables() {
    auto pdf = [](double v) { return std::exp(-v * v / 2.0); };
    x[0] = kZigV / pdf(kZigR);
    x[1] = kZigR;
    for (int i = 2; i < 256; ++i) x[i] = std::sqrt(-2.0 * std::log(kZigV / x[i - 1] + pdf(x[i - 1])));
    x[256] = 0.0;
    for (int i = 0; i < 257; ++i) f[i] = pdf(x[i]);
  }
};
}  // namespace

// rand_distr 0.4.3 StandardNormal (f64): one u64 gives the layer (low 8
// bits) and a uniform in [-1, 1) (high 52 bits); the wedge test draws a
// Standard f64, the base layer's tail draws Open01 pairs.
double SmallRng::standard_normal() {
  static const ZigTables z;
  for (;;) {
    const uint64_t bits = next_u64();
    const unsigned i = static_cast<unsigned>(bits & 0xff);
    const double u = f64_with_exponent(bits >> 12, 1) - 3.0;
    const double x = u * z.x[i];
    if (std::fabs(x) < z.x[i + 1]) return x;
    if (i == 0) {
      double xt = 1.0, yt = 0.0;
      while (-2.0 * yt < xt * xt) {
        const double a = f64_with_exponent(next_u64() >> 12, 0) - (1.0 - 1.1102230246251565e-16);
        const double b = f64_with_exponent(next_u64() >> 12, 0) - (1.0 - 1.1102230246251565e-16);
        xt = std::log(a) / kZigR;
        yt = std::log(b);
      }
      return u < 0.0 ? xt - kZigR : kZigR - xt;
    }
    if (z.f[i + 1] + (z.f[i] - z.f[i + 1]) * next_f64() < std::exp(-x * x / 2.0)) return x;
  }
}

}  // namespace pm
