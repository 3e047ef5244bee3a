// grad_paged.hip -- paged form of the target-blocked gradient (gradient.rs:98-151)
// for embeddings whose target range no longer fits in LDS.
//
// The tb layout splits the other points into kTargetBlocks index ranges and
// runs range b on XCD b, so each XCD gathers from one eighth of Y.  At a
// million points a range is 1 MB: it stays in the XCD's L2, but every entry
// is then one random L2 request, and the pass is bound by how many of those
// a CU keeps in flight.  Here a workgroup (range b, chunk c of rc rows)
// keeps its rows' Y and their gradient sums in LDS and walks range b in
// pages of pp points, each copied into LDS with coalesced loads while the
// previous page is computed, so every gather is an LDS read.
//
// Entries of a (range, chunk, page) block are stored ELL-style: the chunk's
// rows are sorted by their entry count in the page (descending, ties by
// row), cut into slices of 64 (one row per lane), and each slice stores its
// entries step-major (step k holds entry k of its 64 rows, padded to the
// slice's longest row).  The slices of a page are dealt to the 16 waves to
// even out their steps; a wave's slices of one page are contiguous ("a
// run"), and its runs of all pages follow each other, so a wave streams its
// words and loads pages ahead while it computes the current one.
// Each word carries the pair type, the owner bit, the row (local to the
// chunk) and the target (local to the page); the last step of a slice is
// flagged, and there the lanes add their row's page sum into the row's LDS
// sum.  A row appears once per page, so no two lanes update one sum
// together; pages are separated by barriers.  Padding words point at a far
// point staged after the page (their term is a signed zero, no loss); lanes
// of a slice without a row use a dummy row after the chunk.
//
// Summation order of a row's range-b partial: pages in order, within a page
// the row's entries in CSR order.  The range partials are then combined in
// range order by the tb combine pass.  Deterministic: no atomics on sums.
#include "grad_device.h"

#include <cstdint>

namespace pm {

constexpr uint32_t kPageThreads = 1024;                 // gradient pass workgroup
constexpr uint32_t kPageWaves = kPageThreads / 64;      // runs per block
constexpr uint32_t kPageBuildThreads = 512;             // layout passes
constexpr uint32_t kPageMaxRows = 6400;                 // rows per chunk (13-bit row field, dummy = rc; layout LDS)
constexpr uint32_t kPageMaxPoints = 16380;              // points per page: 14-bit target field, far = pp
constexpr uint32_t kPageBuckets = 16;                   // count buckets of the layout sort
constexpr uint32_t kPageTgtBits = 14;
constexpr uint32_t kPageTgtMask = (1u << kPageTgtBits) - 1u;
constexpr uint32_t kPageRowShift = 14;
constexpr uint32_t kPageRowBits = 13;
constexpr uint32_t kPageLast = 1u << 27;                // last step of a slice
constexpr uint32_t kPagePrefetch = 2;                   // float4 per thread of a page copy
constexpr uint32_t kPageStaticLds = 128;                // the gradient pass's static LDS (wave sums)

// LDS floats of the rows' Y and sums (rc + 1 rows each: the dummy row),
// rounded to 16 bytes so that the page buffers start aligned.
__host__ __device__ __forceinline__ uint32_t page_row_floats(uint32_t rc, uint32_t dims) {
  return (2u * (rc + 1u) * dims + 3u) & ~3u;
}
__host__ __device__ __forceinline__ uint32_t page_buf_floats(uint32_t pp, uint32_t dims) { return (pp + 4u) * dims; }

// One plan for a cap on the rows per chunk: chunks a multiple of 32 (so that
// the 8 ranges x chunks workgroups are a multiple of the 256 CUs), pages as
// large as the LDS left by the rows allows.
static bool page_plan_rows(uint32_t n, uint32_t dims, uint32_t rcap, uint32_t max_page_points, uint32_t min_pp,
                           PagePlan &p) {
  p = PagePlan{};
  p.span = tb_span(n);
  rcap &= ~3u;
  if (rcap < 64) return false;
  const uint64_t per = 32ull * rcap;
  const uint64_t k = (n + per - 1) / per;
  const uint64_t want = 32ull * k;
  uint64_t rc = (n + want - 1) / want;
  rc = (rc + 3) & ~3ull;
  p.rc = static_cast<uint32_t>(rc);
  p.chunks = static_cast<uint32_t>((n + rc - 1) / rc);
  const uint32_t row_floats = page_row_floats(p.rc, dims);
  const int64_t avail = static_cast<int64_t>((kLdsPerCu - kPageStaticLds) / 4) - row_floats;
  if (avail <= 0) return false;
  int64_t ppmax = avail / (2 * dims) - 4;
  ppmax = ppmax < kPageMaxPoints ? ppmax : kPageMaxPoints;
  const int64_t pf_cap = static_cast<int64_t>(kPagePrefetch * 4 * kPageThreads / dims);  // one page copy per thread
  ppmax = ppmax < pf_cap ? ppmax : pf_cap;
  if (max_page_points && ppmax > max_page_points) ppmax = max_page_points;
  ppmax &= ~int64_t{3};
  if (ppmax < 64 || (ppmax < min_pp && p.span > ppmax)) return false;
  p.pages = static_cast<uint32_t>((p.span + ppmax - 1) / ppmax);
  if (p.pages > 64) return false;  // the gradient pass holds one page's run offsets per lane
  p.pp = ((p.span + p.pages - 1) / p.pages + 3u) & ~3u;
  p.lds_bytes = (row_floats + 2u * page_buf_floats(p.pp, dims)) * 4u;
  return p.lds_bytes + kPageStaticLds <= kLdsPerCu;
}

// Rows per chunk: the largest cap that leaves pages of at least 1,024
// points (measured at 10^6 points, 2 dims: 6,252 rows and pages of 3,908
// points beat 3,908 rows and pages of 6,252 points, and 8,188 rows with
// pages of 1,636 points); max_rows pins the cap.
bool page_plan(uint32_t n, uint32_t dims, uint32_t max_rows, uint32_t max_page_points, PagePlan &p) {
  p = PagePlan{};
  if (n < kTargetBlocks || dims == 0 || dims > 8) return false;
  if (max_rows)
    return page_plan_rows(n, dims, max_rows < kPageMaxRows ? max_rows : kPageMaxRows, max_page_points, 0, p);
  for (uint32_t rcap : {6400u, 4096u, 2048u, 1024u})
    if (page_plan_rows(n, dims, rcap, max_page_points, 1024, p)) return true;
  return page_plan_rows(n, dims, 512, max_page_points, 0, p);
}
uint64_t page_runs(const PagePlan &p) {
  return static_cast<uint64_t>(kTargetBlocks) * p.chunks * p.pages * kPageWaves;
}
uint32_t page_grad_blocks(const PagePlan &p) { return kTargetBlocks * p.chunks; }

// Points [lo, hi) of page pg of range b.
__device__ __forceinline__ void page_bounds(const PagePlan &p, uint32_t n, uint32_t b, uint32_t pg, uint32_t &lo,
                                            uint32_t &hi) {
  const uint32_t rlo = b * p.span;
  const uint32_t rhi = min(rlo + p.span, n);
  lo = min(rlo + pg * p.pp, rhi);
  hi = min(lo + p.pp, rhi);
}

__device__ __forceinline__ uint32_t page_bucket(uint32_t cnt) {
  return cnt >= kPageBuckets - 1 ? 0u : kPageBuckets - 1 - cnt;
}

// ---------------------------------------------------------------------------
// Layout passes: one workgroup per (range b, chunk c), walking the pages of
// range b in order.  A row's range-b segment is sorted by entry word, so it
// is six runs by the word's top three bits (pair type, owner), each
// ascending by target; a page's entries of a row are the next entries of
// each run below the page's end, and six cursors per row move through the
// segment once.  For every page both passes count each row's entries and
// sort the rows the same way (a stable counting sort by count bucket,
// largest counts first; counts of 15 or more share the first bucket), so
// the fill pass writes exactly the runs the count pass sized, each row's
// entries in CSR order.
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ __launch_bounds__(kPageBuildThreads) void page_layout_kernel(const uint32_t *__restrict__ row_ptr,
                                                                        const uint32_t *__restrict__ tb_ent,
                                                                        const uint16_t *__restrict__ off, uint32_t n,
                                                                        PagePlan p, uint32_t *__restrict__ rsz,
                                                                        const uint32_t *__restrict__ roff,
                                                                        uint32_t *__restrict__ went) {
  constexpr uint32_t kSl = (kPageMaxRows + 63) / 64;
  constexpr uint32_t kBw = kPageBuildThreads / 64;  // waves of the sort
  constexpr int kRuns = 6;                          // (type, owner) runs of a segment
  __shared__ uint32_t sBeg[kPageMaxRows];           // segment start (absolute)
  __shared__ uint16_t sLen[kPageMaxRows];           // segment length
  __shared__ uint16_t sCur[kRuns][kPageMaxRows];    // cursors into the runs (relative)
  __shared__ uint16_t sCnt[kPageMaxRows];           // entries in the current page
  __shared__ uint16_t sRowAt[kPageMaxRows];
  __shared__ uint32_t sW[kSl];                      // slice widths (steps)
  __shared__ uint32_t sBefore[kSl];                 // steps before the slice in its wave's run
  __shared__ uint8_t sWave[kSl];
  __shared__ uint32_t sLoad[kPageWaves];
  __shared__ uint32_t sBk[kBw][kPageBuckets];       // per-wave bucket counts, then starts
  const uint32_t c = blockIdx.x % p.chunks, b = blockIdx.x / p.chunks;
  const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const uint32_t r0 = c * p.rc, rows = min(p.rc, n - r0);
  // run r of a row starts after the entries whose key (word >> 29) is below r
  for (uint32_t i = t; i < rows; i += kPageBuildThreads) {
    const uint32_t row = r0 + i;
    const uint32_t rb = row_ptr[row];
    const uint32_t s0 = tb_seg_beg(off, n, b, row), len = tb_seg_end(row_ptr, off, n, b, row) - s0;
    uint32_t st[kRuns];
#pragma unroll
    for (int r = 0; r < kRuns; ++r) st[r] = 0;
    for (uint32_t k = 0; k < len; ++k) {
      const uint32_t key = tb_ent[rb + s0 + k] >> 29;
#pragma unroll
      for (int r = 0; r < kRuns; ++r) st[r] += key < static_cast<uint32_t>(r) ? 1u : 0u;
    }
    sBeg[i] = rb + s0;
    sLen[i] = static_cast<uint16_t>(len);
#pragma unroll
    for (int r = 0; r < kRuns; ++r) sCur[r][i] = static_cast<uint16_t>(st[r]);
  }
  __syncthreads();
  // the sort's rows: wave wv owns rows [wv q, wv q + q) in order
  const uint32_t q_per = (rows + kBw - 1) / kBw;
  const uint32_t sq0 = min(wv * q_per, rows), sq1 = min(sq0 + q_per, rows);
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t nsl = (rows + 63) / 64;
  const size_t run_base = (static_cast<size_t>(b) * p.chunks + c) * kPageWaves * p.pages;
  for (uint32_t pg = 0; pg < p.pages; ++pg) {
    uint32_t plo, phi;
    page_bounds(p, n, b, pg, plo, phi);
    // entries of each row below the page's end (the count pass moves the
    // cursors here; the fill pass moves them as it writes).  A row meets a
    // page about 0.27 times on average, so the word at each run's cursor is
    // usually already past the page: the six runs' words are loaded
    // together and a run is followed further only when its word is inside.
    for (uint32_t i = t; i < rows; i += kPageBuildThreads) {
      const uint32_t beg = sBeg[i], len = sLen[i];
      uint32_t k[kRuns], w[kRuns];
#pragma unroll
      for (int r = 0; r < kRuns; ++r) {
        k[r] = sCur[r][i];
        w[r] = k[r] < len ? tb_ent[beg + k[r]] : 0xffffffffu;
      }
      uint32_t cnt = 0;
#pragma unroll
      for (int r = 0; r < kRuns; ++r) {
        const uint32_t k0 = k[r];
        while (k[r] < len && (w[r] >> 29) == static_cast<uint32_t>(r) && (w[r] & kIdxMask) < phi) {
          ++k[r];
          w[r] = k[r] < len ? tb_ent[beg + k[r]] : 0xffffffffu;
        }
        cnt += k[r] - k0;
        if (!FILL) sCur[r][i] = static_cast<uint16_t>(k[r]);
      }
      sCnt[i] = static_cast<uint16_t>(cnt);
    }
    __syncthreads();
    // stable counting sort by bucket: per-wave bucket counts, starts, places
    {
      uint32_t cur[kPageBuckets];
#pragma unroll
      for (uint32_t v = 0; v < kPageBuckets; ++v) cur[v] = 0;
      for (uint32_t q0 = sq0; q0 < sq1; q0 += 64) {
        const uint32_t i = q0 + lane;
        const uint32_t bk = i < sq1 ? page_bucket(sCnt[i]) : kPageBuckets;
#pragma unroll
        for (uint32_t v = 0; v < kPageBuckets; ++v) cur[v] += __popcll(__ballot(bk == v));
      }
      if (lane < kPageBuckets) {
        uint32_t x = 0;
#pragma unroll
        for (uint32_t v = 0; v < kPageBuckets; ++v)
          if (v == lane) x = cur[v];
        sBk[wv][lane] = x;
      }
      __syncthreads();
      // exclusive scan of the counts in (bucket, wave) order by wave 0: lane
      // l holds entries 2l and 2l + 1 of the kPageBuckets x kBw table
      if (wv == 0) {
        static_assert(kPageBuckets * kBw == 128, "two scan entries per lane");
        const uint32_t e0 = 2 * lane, e1 = e0 + 1;
        const uint32_t x0 = sBk[e0 % kBw][e0 / kBw], x1 = sBk[e1 % kBw][e1 / kBw];
        uint32_t inc = x0 + x1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(inc, o, 64);
          if (static_cast<int>(lane) >= o) inc += y;
        }
        const uint32_t ex = inc - x0 - x1;
        sBk[e0 % kBw][e0 / kBw] = ex;
        sBk[e1 % kBw][e1 / kBw] = ex + x0;
      }
      __syncthreads();
#pragma unroll
      for (uint32_t v = 0; v < kPageBuckets; ++v) cur[v] = sBk[wv][v];
      for (uint32_t q0 = sq0; q0 < sq1; q0 += 64) {
        const uint32_t i = q0 + lane;
        const uint32_t bk = i < sq1 ? page_bucket(sCnt[i]) : kPageBuckets;
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t v = 0; v < kPageBuckets; ++v) {
          const uint64_t m = __ballot(bk == v);
          if (bk == v) pos = cur[v] + __popcll(m & below);
          cur[v] += __popcll(m);
        }
        if (i < sq1) sRowAt[pos] = static_cast<uint16_t>(i);
      }
    }
    __syncthreads();
    // slice widths: the longest row of each slice of 64 places, one wave
    // per slice
    for (uint32_t sl = wv; sl < nsl; sl += kBw) {
      const uint32_t q = sl * 64 + lane;
      uint32_t w = q < rows ? static_cast<uint32_t>(sCnt[sRowAt[q]]) : 0u;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) w = max(w, static_cast<uint32_t>(__shfl_xor(static_cast<int>(w), o, 64)));
      if (lane == 0) sW[sl] = w;
    }
    __syncthreads();
    // slices to waves, each to the least loaded wave so far (ties: the lower
    // wave), so that every wave of the gradient pass has about the same
    // steps in this page; a wave's slices are stored in slice order.  Wave
    // 0 keeps the loads in lanes 0..15 and finds the minimum by shuffles.
    if (wv == 0) {
      uint32_t load = 0;
      for (uint32_t sl = 0; sl < nsl; ++sl) {
        uint32_t key = lane < kPageWaves ? (load << 4) | lane : 0xffffffffu;
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) key = min(key, static_cast<uint32_t>(__shfl_xor(static_cast<int>(key), o, 64)));
        const uint32_t best = key & 15u;
        const uint32_t wsl = sW[sl];
        if (lane == 0) {
          sWave[sl] = static_cast<uint8_t>(best);
          sBefore[sl] = key >> 4;
        }
        if (lane == best) load += wsl;
      }
      if (lane < kPageWaves) sLoad[lane] = load;
    }
    __syncthreads();
    // run of wave w: ((b chunks + c) 16 + w) pages + pg (a wave's runs of
    // all pages are contiguous)
    const size_t run0 = run_base + pg;
    if (!FILL) {
      if (t < kPageWaves) rsz[run0 + static_cast<size_t>(t) * p.pages] = 64u * sLoad[t];
    } else {
      for (uint32_t q = t; q < nsl * 64; q += kPageBuildThreads) {
        const uint32_t sl = q >> 6, j = q & 63, ws = sW[sl];
        if (ws == 0) continue;
        uint32_t *dst = went + roff[run0 + static_cast<size_t>(sWave[sl]) * p.pages] + 64u * sBefore[sl] + j;
        const uint32_t i = q < rows ? sRowAt[q] : p.rc;
        const uint32_t cnt = q < rows ? sCnt[i] : 0u;
        const uint32_t rf = (cnt ? i : p.rc) << kPageRowShift;
        uint32_t k = 0;
        if (cnt) {
          // as in the count: the six runs' cursor words loaded together
          const uint32_t beg = sBeg[i], len = sLen[i];
          uint32_t e[kRuns], w[kRuns];
#pragma unroll
          for (int r = 0; r < kRuns; ++r) {
            e[r] = sCur[r][i];
            w[r] = e[r] < len ? tb_ent[beg + e[r]] : 0xffffffffu;
          }
#pragma unroll
          for (int r = 0; r < kRuns; ++r) {
            while (e[r] < len && (w[r] >> 29) == static_cast<uint32_t>(r) && (w[r] & kIdxMask) < phi) {
              const uint32_t o = w[r] & kIdxMask;
              dst[64u * k] = (w[r] & ~kIdxMask) | rf | (o - plo) | (k + 1 == ws ? kPageLast : 0u);
              ++k;
              ++e[r];
              w[r] = e[r] < len ? tb_ent[beg + e[r]] : 0xffffffffu;
            }
            sCur[r][i] = static_cast<uint16_t>(e[r]);
          }
        }
        for (; k < ws; ++k) dst[64u * k] = (2u << 30) | rf | p.pp | (k + 1 == ws ? kPageLast : 0u);
      }
    }
    __syncthreads();
  }
}

static uint32_t page_layout_blocks(const PagePlan &p) { return kTargetBlocks * p.chunks; }

hipError_t launch_page_count(const uint32_t *row_ptr, const uint32_t *tb_ent, const uint16_t *off, uint32_t n,
                             const PagePlan &p, uint32_t *rsz, hipStream_t s) {
  if (p.rc > kPageMaxRows || p.pp > kPageMaxPoints) return hipErrorInvalidValue;
  page_layout_kernel<false><<<page_layout_blocks(p), kPageBuildThreads, 0, s>>>(row_ptr, tb_ent, off, n, p, rsz,
                                                                               nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_page_fill(const uint32_t *row_ptr, const uint32_t *tb_ent, const uint16_t *off, uint32_t n,
                            const PagePlan &p, const uint32_t *roff, uint32_t *went, hipStream_t s) {
  if (p.rc > kPageMaxRows || p.pp > kPageMaxPoints) return hipErrorInvalidValue;
  page_layout_kernel<true><<<page_layout_blocks(p), kPageBuildThreads, 0, s>>>(row_ptr, tb_ent, off, n, p, nullptr,
                                                                              roff, went);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gradient pass: one workgroup per (range b = blockIdx.x % 8, chunk c), so
// range b runs on XCD b (workgroups are dealt to the 8 XCDs round-robin)
// and its pages are read from that XCD's L2.
//
// A page period is short (a few steps per wave), so the loads run ahead:
// the words of page pg + 3 and the Y copy of page pg + 4 are issued while
// page pg is computed (four word and four copy register sets in turn, the
// page loop unrolled by 4 so that every set has a static name).  Each
// wave's run offsets of all pages are read once, one page per lane.  A run
// longer than kPageRunSteps steps loads the rest when it is computed, and
// that wait drains every load in flight: the host keeps the gather form of
// the kernel for layouts where such runs are not rare.
// ---------------------------------------------------------------------------
constexpr uint32_t kPageMaxPages = 64;  // run offsets held one page per lane

template <int DIMS, bool LOSS>
__global__ __launch_bounds__(kPageThreads) void grad_page_kernel(const GradArgs a, const PagePlan p,
                                                                 const uint32_t *__restrict__ roff,
                                                                 const uint32_t *__restrict__ went,
                                                                 float *__restrict__ part) {
  extern __shared__ __align__(16) float lds[];
  __shared__ float sLoss[kPageWaves];
  const uint32_t n = a.n;
  const uint32_t b = blockIdx.x % kTargetBlocks, c = blockIdx.x / kTargetBlocks;
  const uint32_t r0 = c * p.rc, rows = min(p.rc, n - r0);
  // (rc + 1) rows (the chunk's, then a dummy row) of [y, range-b sum]:
  // a slice's first step reads both, its last step writes the sum
  float *sRA = lds;
  float *sPg0 = lds + page_row_floats(p.rc, DIMS);     // page buffers, each with the far point at pp
  float *sPg1 = sPg0 + page_buf_floats(p.pp, DIMS);
  const uint32_t t = threadIdx.x;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lane = t & 63;
  const uint32_t pages = p.pages;

  // page copies: kPagePrefetch float4 per thread plus one float of the < 4
  // left over; every load is issued (clamped to the first point), so each
  // copy is the same number of loads
  typedef float v4f __attribute__((ext_vector_type(4)));
  struct Copy {
    v4f v[kPagePrefetch];
    float r;
    uint32_t n4, rem;
  };
  auto load_page = [&](uint32_t pg, Copy &q) __attribute__((always_inline)) {
    uint32_t lo, hi;
    page_bounds(p, n, b, min(pg, pages - 1), lo, hi);
    const uint32_t cnt = pg < pages ? (hi - lo) * DIMS : 0u;
    q.n4 = cnt / 4;
    q.rem = cnt % 4;
    const float *src = a.y_in + static_cast<size_t>(lo) * DIMS;  // lo is a multiple of 4 points
    const v4f *s4 = reinterpret_cast<const v4f *>(src);
#pragma unroll
    for (uint32_t k = 0; k < kPagePrefetch; ++k) {
      const uint32_t i = t + k * kPageThreads;
      q.v[k] = s4[i < q.n4 ? i : 0u];
    }
    q.r = src[t < q.rem ? q.n4 * 4 + t : 0u];
  };
  auto store_page = [&](const Copy &q, float *dst) __attribute__((always_inline)) {
    v4f *d4 = reinterpret_cast<v4f *>(dst);
#pragma unroll
    for (uint32_t k = 0; k < kPagePrefetch; ++k) {
      const uint32_t i = t + k * kPageThreads;
      if (i < q.n4) d4[i] = q.v[k];
    }
    if (t < q.rem) dst[q.n4 * 4 + t] = q.r;
  };

  // this wave's runs: lane l holds the offsets of page l
  const size_t stream = (static_cast<size_t>(b) * p.chunks + c) * kPageWaves * pages +
                        static_cast<size_t>(wave) * pages;
  uint32_t run_beg = 0, run_end = 0;
  if (static_cast<uint32_t>(lane) < pages) {
    run_beg = roff[stream + lane];
    run_end = roff[stream + lane + 1];
  }
  constexpr uint32_t kU = kPageRunSteps;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(went), 0, static_cast<int>(p.words * 4u), 0x00020000);
  struct Run {
    uint32_t w[kU];
  };
  auto run_beg_of = [&](uint32_t pg) __attribute__((always_inline)) { return __builtin_amdgcn_readlane(run_beg, static_cast<int>(min(pg, pages - 1))); };
  auto run_steps_of = [&](uint32_t pg) __attribute__((always_inline)) {
    const int l = static_cast<int>(min(pg, pages - 1));
    return pg < pages ? (__builtin_amdgcn_readlane(run_end, l) - __builtin_amdgcn_readlane(run_beg, l)) >> 6 : 0u;
  };
  // steps past the run point past the word array: such a load returns 0
  // without a memory access, and every fetch is the same number of loads
  auto fetch = [&](uint32_t pg, Run &r) __attribute__((always_inline)) {
    const uint32_t beg = run_beg_of(pg), steps = run_steps_of(pg);
#pragma unroll
    for (uint32_t k = 0; k < kU; ++k)
      r.w[k] = __builtin_amdgcn_raw_buffer_load_b32(
          rs, lane * 4, static_cast<int>(k < steps ? (beg + 64u * k) * 4u : p.words * 4u), 0);
  };

  float loss = 0.0f;
  float g[DIMS];
#pragma unroll
  for (int d = 0; d < DIMS; ++d) g[d] = 0.0f;
  const PairConsts pc = pair_consts(a.w);
  const float kn = pc.k_nn, km = pc.k_mn, kf = pc.k_fp, wn = pc.w_nn, wm = pc.w_mn, wf = pc.w_fp;
  // one step: `gather` reads the target's Y (and at a slice's first step the
  // row's [y, sum]); the reads of a group of steps are issued together, then
  // `apply` adds the terms in step order.  The row's y and sum stay in
  // registers for the steps of its slice; at the slice's last step the lanes
  // write their rows' sums back.  Slices start and end together on all lanes
  // (the flag is wave-uniform).
  struct Operands {
    float yo[DIMS];
    float ra[2 * DIMS];
  };
  float yr[DIMS], ar[DIMS];
#pragma unroll
  for (int d = 0; d < DIMS; ++d) yr[d] = ar[d] = 0.0f;
  auto gather = [&](uint32_t w, bool first, const float *pg, Operands &o) __attribute__((always_inline)) {
    const uint32_t tg = w & kPageTgtMask;
#pragma unroll
    for (int d = 0; d < DIMS; ++d) o.yo[d] = pg[tg * DIMS + d];
    if (first) {
      const float *ra = sRA + __builtin_amdgcn_ubfe(w, kPageRowShift, kPageRowBits) * (2 * DIMS);
#pragma unroll
      for (int q = 0; q < 2 * DIMS; ++q) o.ra[q] = ra[q];
    }
  };
  auto is_last = [&](uint32_t w) __attribute__((always_inline)) {
    return (__builtin_amdgcn_readfirstlane(w) & kPageLast) != 0u;
  };
  auto apply = [&](uint32_t w, bool first, const Operands &o) __attribute__((always_inline)) {
    if (first) {
#pragma unroll
      for (int d = 0; d < DIMS; ++d) {
        yr[d] = o.ra[d];
        ar[d] = o.ra[DIMS + d];
      }
    }
    pair_term<DIMS, LOSS>(w, yr, o.yo, kn, km, kf, wn, wm, wf, g, loss);
    if (is_last(w)) {
      float *ra = sRA + __builtin_amdgcn_ubfe(w, kPageRowShift, kPageRowBits) * (2 * DIMS);
#pragma unroll
      for (int d = 0; d < DIMS; ++d) {
        ra[DIMS + d] = ar[d] + g[d];
        g[d] = 0.0f;
      }
    }
  };
  constexpr uint32_t kG = 3;  // steps whose reads are in flight together
  auto compute = [&](uint32_t pgi, const Run &r, const float *pg) __attribute__((always_inline)) {
    const uint32_t steps = run_steps_of(pgi);
    bool first = true;  // a run starts and ends on slice boundaries
#pragma unroll
    for (uint32_t k0 = 0; k0 < kU; k0 += kG) {
      if (k0 < steps) {
        Operands o[kG];
        bool fst[kG];
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k)
          if (k0 + k < steps) {
            fst[k] = first;
            gather(r.w[k0 + k], first, pg, o[k]);
            first = is_last(r.w[k0 + k]);
          }
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k)
          if (k0 + k < steps) apply(r.w[k0 + k], fst[k], o[k]);
      }
    }
    if (steps > kU) {
      const uint32_t beg = run_beg_of(pgi);
      for (uint32_t k = kU; k < steps; ++k) {
        const uint32_t w =
            __builtin_amdgcn_raw_buffer_load_b32(rs, lane * 4, static_cast<int>((beg + 64u * k) * 4u), 0);
        Operands o;
        const bool f = first;
        gather(w, f, pg, o);
        first = is_last(w);
        apply(w, f, o);
      }
    }
  };

  Run w0, w1, w2, w3;
  Copy q0, q1, q2, q3;
  fetch(0, w0);
  fetch(1, w1);
  fetch(2, w2);
  load_page(0, q0);
  load_page(1, q1);
  load_page(2, q2);
  load_page(3, q3);
  {
    const float *src = a.y_in + static_cast<size_t>(r0) * DIMS;
    for (uint32_t i = t; i < (p.rc + 1) * DIMS; i += kPageThreads) {
      const uint32_t row = i / DIMS, d = i % DIMS;
      sRA[row * (2 * DIMS) + d] = i < rows * DIMS ? src[i] : 0.0f;
      sRA[row * (2 * DIMS) + DIMS + d] = 0.0f;
    }
  }
  if (t < DIMS) {
    sPg0[p.pp * DIMS + t] = kSellFar;
    sPg1[p.pp * DIMS + t] = kSellFar;
  }
  store_page(q0, sPg0);
  __syncthreads();
  // page pg: its Y is in buffer pg % 2 and its words in set pg % 4.  Page
  // pg + 1 waits in copy set (pg + 1) % 4 and is stored into the other
  // buffer first; page pg + 4 is loaded into copy set pg % 4 (free since
  // page pg was stored) and page pg + 3's words into the set page pg - 1
  // used.  Copies and words run three pages ahead.
  auto page = [&](uint32_t pgi, const Run &wu, Run &wp, const Copy &qs, Copy &ql, float *bcur, float *bnxt)
      __attribute__((always_inline)) {
    if (pgi + 1 < pages) store_page(qs, bnxt);
    load_page(pgi + 4, ql);
    fetch(pgi + 3, wp);
    compute(pgi, wu, bcur);
    __syncthreads();
  };
  for (uint32_t pg = 0; pg < pages; pg += 4) {
    page(pg, w0, w3, q1, q0, sPg0, sPg1);
    if (pg + 1 >= pages) break;
    page(pg + 1, w1, w0, q2, q1, sPg1, sPg0);
    if (pg + 2 >= pages) break;
    page(pg + 2, w2, w1, q3, q2, sPg0, sPg1);
    if (pg + 3 >= pages) break;
    page(pg + 3, w3, w2, q0, q3, sPg1, sPg0);
  }
  float *dst = part + (static_cast<size_t>(b) * n + r0) * DIMS;
  for (uint32_t i = t; i < rows * DIMS; i += kPageThreads) dst[i] = sRA[(i / DIMS) * (2 * DIMS) + DIMS + i % DIMS];
  if (LOSS) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss = loss + __shfl_xor(loss, o, 64);
    if (lane == 0) sLoss[wave] = loss;
    __syncthreads();
    if (t == 0) {
      float s = 0.0f;
      for (uint32_t w = 0; w < kPageWaves; ++w) s = s + sLoss[w];
      a.loss_part[blockIdx.x] = s;
    }
  }
}

template <int DIMS>
static hipError_t launch_page_dims(const GradArgs &a, const PagePlan &p, const uint32_t *roff, const uint32_t *went,
                                   float *part, bool adam, hipStream_t s) {
  const uint32_t g1 = page_grad_blocks(p);
  if (g1 == 0) return hipSuccess;
  if (a.loss_part)
    grad_page_kernel<DIMS, true><<<g1, kPageThreads, p.lds_bytes, s>>>(a, p, roff, went, part);
  else
    grad_page_kernel<DIMS, false><<<g1, kPageThreads, p.lds_bytes, s>>>(a, p, roff, went, part);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_tb_combine(a, part, DIMS, adam, s);
}

hipError_t launch_grad_adam_paged(const GradArgs &a, const PagePlan &p, const uint32_t *roff, const uint32_t *went,
                                  float *part, uint32_t dims, bool adam, hipStream_t s) {
  if (p.rc > kPageMaxRows || p.pp > kPageMaxPoints || p.words >= kPageMaxWords) return hipErrorInvalidValue;
  switch (dims) {
    case 1: return launch_page_dims<1>(a, p, roff, went, part, adam, s);
    case 2: return launch_page_dims<2>(a, p, roff, went, part, adam, s);
    case 3: return launch_page_dims<3>(a, p, roff, went, part, adam, s);
    case 4: return launch_page_dims<4>(a, p, roff, went, part, adam, s);
    case 5: return launch_page_dims<5>(a, p, roff, went, part, adam, s);
    case 6: return launch_page_dims<6>(a, p, roff, went, part, adam, s);
    case 7: return launch_page_dims<7>(a, p, roff, went, part, adam, s);
    case 8: return launch_page_dims<8>(a, p, roff, went, part, adam, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pm
