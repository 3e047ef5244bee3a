// host.cpp -- C ABI of libpacmap_mi355x: fit_transform orchestration, kNN and
// the per-iteration seams.  Every entry point validates on the host first,
// then runs its kernels on a stream of its own and blocks until done.
#include <algorithm>
#include <cmath>
#include <memory>
#include <new>
#include <cstdlib>
#include <random>

#include "host_common.h"
#include "internal.h"

namespace pm {
namespace {

thread_local pacmap_timings g_tm;
// rows of the last kNN on this thread that the certificate sent to the exact
// rescan: copied into pinned memory on the stream (no host wait) and read
// after the call returns, once the copy's event has completed
struct RescannedRows {
  uint32_t *rows = nullptr;  // pinned host memory
  uint64_t cap = 0, count = 0;
  hipEvent_t done = nullptr;
  ~RescannedRows() {
    if (process_exiting()) return;
    if (rows) (void)hipHostFree(rows);
    if (done) (void)hipEventDestroy(done);
  }
  void record(const uint32_t *dev_rows, uint64_t n, hipStream_t s) {
    if (n > cap) {
      if (done) PM_HIP(hipEventSynchronize(done));  // the last copy into rows is done
      if (rows) PM_HIP(hipHostFree(rows));
      rows = nullptr;
      cap = 0;
      note_runtime_use();
      PM_HIP(hipHostMalloc(reinterpret_cast<void **>(&rows), n * sizeof(uint32_t), 0));
      cap = n;
    }
    if (!done) PM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    PM_HIP(hipMemcpyAsync(rows, dev_rows, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PM_HIP(hipEventRecord(done, s));
    count = n;
  }
  std::vector<uint32_t> read() const {
    if (count == 0) return {};
    PM_HIP(hipEventSynchronize(done));
    return std::vector<uint32_t>(rows, rows + count);
  }
};
thread_local RescannedRows g_rescanned;

// ---------------------------------------------------------------------------
// exact kNN on a device matrix (knn.rs:137-194)
// ---------------------------------------------------------------------------
// Relative error coefficient of the MFMA lower bound lb = a_q + a_j - 2 dot
// (see knn_mfma.hip): bf16 input rounding (split: hi + lo leaves ~3 u^2,
// single: 2 u, u = 2^-8), FP32 accumulation of the MFMA (counted as one
// truncating rounding per product), the rounding of a, of the sums, and of
// the centring; doubled for safety.
double knn_kappa(uint32_t kpad, bool split) {
  const double u = 1.0 / 256.0, e23 = 1.0 / 8388608.0, e24 = e23 / 2;
  const double dotc = split ? 3.1 * u * u + (3.0 * kpad + 1.0) * e23 * 1.02
                            : (2.0 * u + u * u) + (kpad + 1.0) * e23 * 1.02;
  return 2.0 * (dotc + 16.0 * e24 + 4.0 * e24 * 2.0);
}

// The reference distance of an unlisted candidate is at least
// sqrt(tau * cert_scale): (1 - gammaA) bounds the 8-lane FP32 summation
// error, (1 - 2^-24)^2 the final sqrt rounding, 2^-20 is margin.
double knn_cert_scale(uint32_t d) {
  const double e24 = 1.0 / 16777216.0;
  const double gammaA = (d / 8.0 + 16.0) * e24 * 1.1;
  return (1.0 - gammaA) * (1.0 - e24) * (1.0 - e24) * (1.0 - 1.0 / 1048576.0);
}

// Upper bounds of the rerank: d^2 <= (aq + ac) ub_scale1 - 2 dot covers
// the product and norm errors that kappa covers for the lower bound (and
// the rounding of a to f32); ub_scale2 the reference's FP32 summation
// (1 + gammaA) and sqrt rounding (1 + 2^-24)^2, with 2^-20 margin.
double knn_ub_scale1(double kappa) { return (1.0 + kappa) / ((1.0 - kappa) * (1.0 - 1.0 / 8388608.0)); }
double knn_ub_scale2(uint32_t d) {
  const double e24 = 1.0 / 16777216.0;
  const double gammaA = (d / 8.0 + 16.0) * e24 * 1.1;
  return (1.0 + gammaA) * (1.0 + e24) * (1.0 + e24) * (1.0 + 1.0 / 1048576.0);
}

bool knn_precision_valid(int precision) {
  const int base = precision & 0xff;
  return (precision & ~(0xff | PACMAP_KNN_APPROX)) == 0 && base >= PACMAP_KNN_FP32 && base <= PACMAP_KNN_DIRECT;
}

// Candidate precision of the MFMA path: 0 single bf16, 1 split bf16, 2 FP8.
int knn_mode(int precision) {
  const int base = precision & 0xff;
  return base == PACMAP_KNN_BF16 ? 0 : base == PACMAP_KNN_FP8 ? 2 : 1;
}
uint32_t knn_chunks(uint64_t d, int mode) { return static_cast<uint32_t>(mode == 2 ? (d + 127) / 128 : (d + 31) / 32); }

bool knn_mfma_eligible(uint64_t n, uint64_t d, uint32_t k, int precision) {
  if ((precision & 0xff) == PACMAP_KNN_DIRECT) return false;
  const uint64_t kch = knn_chunks(d, knn_mode(precision));
  const uint64_t ntiles = (n + 15) / 16;
  const uint32_t ss = std::max<uint32_t>(1, (3 * k + 15) / 16);
  return n >= 1024 && d >= 1 && kch <= 8 && k <= 128 && ntiles >= 8ull * ss;
}

// MFMA candidates + exact rerank.  certify: every row is either proven exact
// (certificate) or rescanned exactly; otherwise (approximate mode) the
// listed candidates' exact top-k is returned as is.  Returns false when the
// input is not finite (the caller then uses the direct kernel).
bool knn_mfma_path(const float *x_dev, uint64_t n, uint64_t d, uint32_t k, const float *xp,
                   const KnnLayout &L, int mode, bool certify, DevBuf<uint32_t> &idx, DevBuf<float> &dist,
                   hipStream_t s) {
  const bool split = mode == 1;
  const uint32_t kch = knn_chunks(d, mode);
  const uint64_t ntiles = (n + 15) / 16;
  const uint64_t rows_padded = ntiles * 16;
  DevBuf<double> mean(d), scratch(colmean_scratch_doubles(n, static_cast<uint32_t>(d)));
  PM_HIP(launch_colmean_f64(x_dev, n, static_cast<uint32_t>(d), mean.get(), scratch.get(), s));
  const size_t frag_words = ntiles * kch * 64 * 4;  // 16 B per lane fragment
  DevBuf<uint32_t> fhi(frag_words), flo(mode != 0 ? frag_words : 4);
  const uint32_t ss = std::max<uint32_t>(1, (3 * k + 15) / 16);
  const uint32_t capg = (3.0 * k * 2.5 <= 512.0) ? 512 : 1024;
  // split the candidate range until there are enough (group, split) blocks
  // to keep every CU busy with a short tail
  const uint32_t groups = knn_cand_groups(static_cast<uint32_t>(ntiles));
  // symmetric pass (each tile pair of the upper triangle once, appends to
  // both rows) unless PACMAP_KNN_SYM=0
  bool sym = true;
  if (const char *e = std::getenv("PACMAP_KNN_SYM")) sym = std::atoi(e) != 0;
  // candidate windows of at most 2 MiB of fragments (half an XCD's L2;
  // PACMAP_KNN_WINDOW_KB overrides)
  KnnSymPlan splan;
  if (sym) {
    uint32_t window_bytes = 2u << 20;
    if (const char *e = std::getenv("PACMAP_KNN_WINDOW_KB"))
      window_bytes = static_cast<uint32_t>(std::max<long>(64, std::min<long>(1 << 20, std::atol(e)))) << 10;
    const uint32_t tile_bytes = kch * (mode != 0 ? 2u : 1u) * 1024u;
    // about 16k items (some 20 rounds of the resident workgroups: measured
    // 2.78 ms at 70,000 points against 2.99 ms with 5,443 items)
    splan = knn_sym_plan(static_cast<uint32_t>(ntiles), std::max<uint32_t>(16384, 4 * groups),
                         std::max<uint32_t>(16, window_bytes / tile_bytes));
    sym = splan.items != 0;
  }
  // The candidate passes see the rows sorted by the Morton code of their
  // first six coordinates, so that a tile's rows lie close together and
  // the symmetric pass can skip tile pairs whose boxes lie apart
  // (PACMAP_KNN_PRUNE=0: the rows' own order and no skipping).
  bool prune = sym;
  if (const char *e = std::getenv("PACMAP_KNN_PRUNE")) prune = prune && std::atoi(e) != 0;
  DevBuf<uint32_t> perm(n);
  {
    DevBuf<uint32_t> omm(16), okey(n), okey2(n), oidx(n);
    const size_t tb = knn_order_temp_bytes(n);
    DevBuf<uint8_t> otemp(std::max<size_t>(tb, 16));
    PM_HIP(launch_knn_order(x_dev, n, prune ? static_cast<uint32_t>(d) : 0u, omm.get(), okey.get(), okey2.get(),
                            oidx.get(), perm.get(), otemp.get(), tb, s));
  }
  DevBuf<float> arow(rows_padded), arow_orig(n), rinv(mode == 2 ? rows_padded : 1);
  DevBuf<uint32_t> flags(2);
  PM_HIP(launch_fill_zero(flags.get(), flags.bytes(), s));
  const double kappa = certify ? knn_kappa(kch * 32, split) : 0.0;
  PM_HIP(launch_knn_rownorm(x_dev, n, static_cast<uint32_t>(d), mean.get(), static_cast<float>(1.0 - kappa),
                            rows_padded, perm.get(), arow.get(), arow_orig.get(), flags.get(),
                            mode == 2 ? rinv.get() : nullptr, s));
  if (mode == 2)
    PM_HIP(launch_knn_frag8(x_dev, n, static_cast<uint32_t>(d), mean.get(), kch, ntiles, rinv.get(), perm.get(),
                            fhi.get(), flo.get(), s));
  else
    PM_HIP(launch_knn_frag(x_dev, n, static_cast<uint32_t>(d), mean.get(), kch, ntiles, perm.get(), fhi.get(),
                           split ? flo.get() : nullptr, s));
  // the tau sample: ceil(n / ss) rows spread over the input, in tiles of
  // their own (in the candidate order a tile's rows are neighbours, and a
  // sample of whole tiles would then set tau from a row's nearest tile)
  const uint64_t srows = (n + ss - 1) / ss, stiles = (srows + 15) / 16;
  DevBuf<uint32_t> srow(srows), sfhi(stiles * kch * 64 * 4), sflo(mode != 0 ? stiles * kch * 64 * 4 : 4);
  DevBuf<float> sarow(stiles * 16), srinv(mode == 2 ? stiles * 16 : 1);
  PM_HIP(launch_knn_sample_rows(n, srows, knn_sample_stride(n), srow.get(), s));
  PM_HIP(launch_knn_rownorm(x_dev, srows, static_cast<uint32_t>(d), mean.get(), static_cast<float>(1.0 - kappa),
                            stiles * 16, srow.get(), sarow.get(), arow_orig.get(), flags.get(),
                            mode == 2 ? srinv.get() : nullptr, s));
  if (mode == 2)
    PM_HIP(launch_knn_frag8(x_dev, srows, static_cast<uint32_t>(d), mean.get(), kch, stiles, srinv.get(), srow.get(),
                            sfhi.get(), sflo.get(), s));
  else
    PM_HIP(launch_knn_frag(x_dev, srows, static_cast<uint32_t>(d), mean.get(), kch, stiles, srow.get(), sfhi.get(),
                           split ? sflo.get() : nullptr, s));
  DevBuf<float> box(prune ? ntiles * 16 : 1);
  if (prune) PM_HIP(launch_knn_tilebox(x_dev, n, static_cast<uint32_t>(d), perm.get(), ntiles, box.get(), s));
  // the finiteness flag is read with the rerank's results: non-finite
  // input only makes the candidate passes list garbage (every access stays
  // within its list), and the caller then runs the direct kernel
  uint32_t hflags[2] = {0, 0};
  const uint32_t nsplit =
      sym ? 1u : std::min<uint32_t>(16, std::max<uint32_t>(1, (4096 + groups - 1) / groups));
  DevBuf<uint32_t> gcnt(n * nsplit), gcand(n * nsplit * capg * (split ? 2 : 1)), fb_list(n);
  DevBuf<float> gtau(n), gtau_orig(n), tmax(ntiles);
  // 256 (listed, reranked) pairs by row mod 256, then the skipped tile pairs
  DevBuf<unsigned long long> kstats(513);
  PM_HIP(launch_fill_zero(kstats.get(), kstats.bytes(), s));
  const void *flo_p = mode != 0 ? flo.get() : fhi.get();
  const float *rinv_p = mode == 2 ? rinv.get() : nullptr;
  if (sym) PM_HIP(launch_fill_zero(gcnt.get(), gcnt.bytes(), s));
  Timer tsamp, tk, tr;
  tsamp.start(s);
  KnnOrder ord;
  ord.perm = perm.get();
  ord.sample.fhi = reinterpret_cast<const uint4 *>(sfhi.get());
  ord.sample.flo = reinterpret_cast<const uint4 *>(mode != 0 ? sflo.get() : sfhi.get());
  ord.sample.arow = sarow.get();
  ord.sample.rinv = mode == 2 ? srinv.get() : nullptr;
  ord.sample.tiles = static_cast<uint32_t>(stiles);
  PM_HIP(launch_knn_cand(0, fhi.get(), flo_p, arow.get(), rinv_p, static_cast<uint32_t>(n), kch,
                         static_cast<uint32_t>(ntiles), ss, capg, nsplit, gcnt.get(), gcand.get(),
                         gtau.get(), mode, s, KnnSymPlan{}, ord));
  // tau by row for the rerank; the largest tau per tile and the tile pairs
  // each query group can skip
  PM_HIP(launch_knn_tiletau(gtau.get(), perm.get(), n, ntiles, tmax.get(), gtau_orig.get(), s));
  const uint32_t pwords = static_cast<uint32_t>((ntiles + 31) / 32);
  DevBuf<uint32_t> pbits(prune ? static_cast<size_t>(groups) * pwords : 1);
  if (prune) {
    DevBuf<float> gbox(static_cast<size_t>(groups) * 17);
    PM_HIP(launch_knn_prune(box.get(), tmax.get(), ntiles, static_cast<uint32_t>(d), pwords, gbox.get(), pbits.get(),
                            kstats.get() + 512, s));
    ord.prune = pbits.get();
    ord.pwords = pwords;
  }
  tsamp.stop(s);
  tk.start(s);
  PM_HIP(launch_knn_cand(sym ? 2 : 1, fhi.get(), flo_p, arow.get(), rinv_p, static_cast<uint32_t>(n), kch,
                         static_cast<uint32_t>(ntiles), ss, capg, nsplit, gcnt.get(), gcand.get(),
                         gtau.get(), mode, s, splan, ord));
  tk.stop(s);
  tr.start(s);
  // upper-bound pruning (certified split lists): see knn_rerank_kernel
  const double ub_s1 = split && certify ? knn_ub_scale1(kappa) : 0.0;
  PM_HIP(launch_knn_rerank(xp, L, static_cast<uint32_t>(n), k, capg, nsplit, gcnt.get(), gcand.get(),
                           gtau_orig.get(), certify ? knn_cert_scale(static_cast<uint32_t>(d)) : -1.0, arow_orig.get(),
                           split, ub_s1, knn_ub_scale2(static_cast<uint32_t>(d)), idx.get(), dist.get(),
                           flags.get() + 1, fb_list.get(), kstats.get(), s));
  tr.stop(s);
  download(hflags, flags.get(), 2, s);
  std::vector<unsigned long long> hstats(513);
  download(hstats.data(), kstats.get(), 513, s);
  PM_HIP(hipStreamSynchronize(s));
  if (hflags[0]) return false;
  tk.add_to(g_tm.knn_kernel_ms);
  g_tm.knn_kernel_launches += 1;
  tsamp.add_to(g_tm.knn_sample_ms);
  tr.add_to(g_tm.knn_rerank_ms);
  g_tm.knn_candidates = capg;
  for (int i = 0; i < 256; ++i) {
    g_tm.knn_listed += hstats[2 * i];
    g_tm.knn_reranked += hstats[2 * i + 1];
  }
  g_tm.knn_path = mode == 1 ? 1 : mode == 0 ? 2 : 3;
  if (sym) {
    // (query group, candidate tile) pairs past the groups' own tiles, and
    // those the symmetric pass skipped
    const uint64_t gt = splan.gt;  // tiles per query group
    for (uint64_t g = 0; g < groups; ++g)
      g_tm.knn_tile_pairs += (g + 1) * gt < ntiles ? ntiles - (g + 1) * gt : 0;
    g_tm.knn_tile_pairs_skipped = hstats[512];
  }
  g_tm.knn_certified = certify ? 1 : 0;
  const uint32_t nfb = hflags[1];
  g_tm.knn_rescanned_rows += nfb;
  if (nfb) {
    // split each listed row's candidates over S blocks (about 512 workgroups
    // in all, at least 256 candidates each), then merge
    const uint64_t qblocks = (nfb + 15) / 16;
    const uint32_t S = static_cast<uint32_t>(
        std::max<uint64_t>(1, std::min<uint64_t>({(512 + qblocks - 1) / qblocks, 128, n / 256})));
    DevBuf<uint64_t> part(static_cast<size_t>(nfb) * S * k);
    Timer tf;
    tf.start(s);
    PM_HIP(launch_knn_rescan(xp, n, L, k, fb_list.get(), nfb, part.get(), S, idx.get(), dist.get(), s));
    tf.stop(s);
    tf.add_to(g_tm.knn_fallback_ms);
    g_rescanned.record(fb_list.get(), nfb, s);
  }
  return true;
}

void knn_device(const float *x_dev, uint64_t n, uint64_t d, uint32_t k, DevBuf<uint32_t> &idx,
                DevBuf<float> &dist, hipStream_t s, int precision = PACMAP_KNN_FP32) {
  if (k > kKnnMaxK)
    fail(PACMAP_ERR_UNSUPPORTED, "k = " + std::to_string(k) + " exceeds the kNN kernel limit of " +
                                     std::to_string(kKnnMaxK));
  if (n >= (1ull << 31)) fail(PACMAP_ERR_UNSUPPORTED, "too many points for the kNN kernel");
  if (d >= (1ull << 28)) fail(PACMAP_ERR_UNSUPPORTED, "too many features for the kNN kernel");
  idx.alloc(n * k);
  dist.alloc(n * k);
  g_rescanned.count = 0;
  if (k == 0) return;
  const KnnLayout L = knn_layout(static_cast<uint32_t>(d));
  DevBuf<float> xp(n * L.Dp);
  Timer all;
  all.start(s);
  PM_HIP(launch_knn_pack(x_dev, xp.get(), n, L, s));
  bool done = false;
  if (knn_mfma_eligible(n, d, k, precision)) {
    const int mode = knn_mode(precision);
    // FP8 candidates cannot be certified (the E4M3 error bound is too wide)
    const bool certify = mode != 2 && !(precision & PACMAP_KNN_APPROX);
    done = knn_mfma_path(x_dev, n, d, k, xp.get(), L, mode, certify, idx, dist, s);
  }
  if (!done) {
    Timer ker;
    ker.start(s);
    PM_HIP(launch_knn_exact(xp.get(), n, L, k, idx.get(), dist.get(), s));
    ker.stop(s);
    ker.add_to(g_tm.knn_kernel_ms);
    g_tm.knn_kernel_launches += 1;
    g_tm.knn_candidates = n - 1;
    g_tm.knn_path = 0;
    g_tm.knn_certified = 1;
  }
  all.stop(s);
  all.add_to(g_tm.knn_ms);
}

// ---------------------------------------------------------------------------
// the optimiser (lib.rs:781-837) as one hipGraph of fused iterations
// ---------------------------------------------------------------------------
struct GraphGuard {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  ~GraphGuard() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
  }
};

// The optimiser graph of this thread's previous call, kept with a key made
// of every launch argument it captured.  A call whose launches would be
// identical (same sizes and schedule and, through the caching allocator, the
// same buffers) replays it instead of capturing and instantiating 450+
// nodes again.  Each call still writes all of the graph's inputs first.
struct GraphCache {
  std::vector<uint64_t> key;
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  void reset() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
    e = nullptr;
    g = nullptr;
    key.clear();
  }
  ~GraphCache() {
    if (!process_exiting()) reset();
  }
};
thread_local GraphCache g_graph;  // the thread's kept graph, destroyed when the thread exits

}  // namespace
void graph_cache_reset() { g_graph.reset(); }
namespace {

struct KeyWriter {
  std::vector<uint64_t> &k;
  void add(uint64_t v) { k.push_back(v); }
  void add(const void *p) { k.push_back(reinterpret_cast<uint64_t>(p)); }
  void add(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    k.push_back(u);
  }
  void add(const GradArgs &a) {
    add(a.row_ptr), add(a.ent), add(a.y_in), add(static_cast<const void *>(a.y_out)), add(a.m), add(a.v);
    add(a.grad_out), add(a.loss_part), add(a.snap), add(static_cast<uint64_t>(a.n));
    add(a.w[0]), add(a.w[1]), add(a.w[2]), add(a.lr_t), add(a.omb1), add(a.omb2), add(a.mean_row_len);
  }
};

// Symmetrised CSR in device memory (see internal.h for the entry format).
struct DevCsr {
  DevBuf<uint32_t> row_ptr, ent;
  uint64_t nnz = 0;
};

// Builds the CSR of three device pair arrays on the GPU, every row sorted by
// entry value (hub rows too, by a workgroup each).
void build_csr_dev(uint64_t n, const uint32_t *const dp[3], const uint64_t rows[3], DevCsr &out,
                   hipStream_t s) {
  if (n >= kMaxPoints) fail(PACMAP_ERR_UNSUPPORTED, "too many points for the CSR entry format");
  Timer tcsr;  // stream time from the first CSR launch to the last
  tcsr.start(s);
  DevBuf<uint32_t> deg(n), flags(2), scratch((n + 1023) / 1024 + 2), cursor(n);
  PM_HIP(launch_fill_zero(deg.get(), deg.bytes(), s));
  PM_HIP(launch_fill_zero(flags.get(), flags.bytes(), s));
  PM_HIP(launch_csr_count(dp, rows, static_cast<uint32_t>(n), deg.get(), flags.get(), s));
  out.row_ptr.alloc(n + 1, kSlotCsrRow);
  PM_HIP(launch_csr_scan(deg.get(), static_cast<uint32_t>(n), out.row_ptr.get(), cursor.get(), scratch.get(), s));
  uint32_t hf[2] = {0, 0}, nnz = 0;
  download(hf, flags.get(), 2, s);
  download(&nnz, out.row_ptr.get() + n, 1, s);
  PM_HIP(hipStreamSynchronize(s));
  if (hf[0]) fail(PACMAP_ERR_INVALID_PAIR_INDEX, "pair index out of range");
  out.nnz = nnz;
  out.ent.alloc(std::max<uint64_t>(1, nnz), kSlotCsrEnt);
  const uint32_t long_cap = nnz / 2049 + 1;  // rows above 2048 entries
  DevBuf<uint32_t> long_list(long_cap);
  PM_HIP(launch_csr_fill_sort(dp, rows, static_cast<uint32_t>(n), out.row_ptr.get(), cursor.get(),
                              out.ent.get(), flags.get() + 1, long_list.get(), long_cap, s));
  tcsr.stop(s);
  tcsr.add_to(g_tm.csr_ms);
  g_tm.csr_entries = nnz;
}

// Uploads host pair arrays and builds their CSR on the device.
void build_csr_from_host(uint64_t n, const PairList lists[3], DevCsr &out, hipStream_t s) {
  DevBuf<uint32_t> dpair[3];
  const uint32_t *dp[3];
  uint64_t rows[3];
  for (int t = 0; t < 3; ++t) {
    rows[t] = lists[t].rows;
    dpair[t].alloc(std::max<uint64_t>(2, 2 * rows[t]));
    upload(dpair[t], lists[t].p, 2 * rows[t], s);
    dp[t] = dpair[t].get();
  }
  build_csr_dev(n, dp, rows, out, s);
}

// The target-blocked gradient pays off once Y no longer fits one XCD's L2
// (n dims 4 B >= 4 MiB).  PACMAP_GRAD_TARGET_BLOCKS=0 / 1 forces it off / on.
bool use_target_blocks(uint64_t n, uint64_t dims) {
  if (const char *e = std::getenv("PACMAP_GRAD_TARGET_BLOCKS")) {
    if (e[0] == '0') return false;
    if (e[0] == '1') return n >= kTargetBlocks;
  }
  return n * dims * 4 >= (4ull << 20);
}

// The LDS-staged (sliced ELL) gradient, for embeddings small enough that a
// target range of Y plus a row chunk fits in LDS and large enough to fill
// the GPU.  PACMAP_GRAD_SELL=0 / 1 forces it off / on (when it fits); 4
// target ranges when they fit (one workgroup per CU), else 8.
bool use_sell(uint64_t n, uint64_t dims, SellPlan &plan) {
  if (n >= kMaxPoints) return false;
  // PACMAP_GRAD_SELL_RANGES=4 / 8 picks the number of target ranges
  uint32_t order[2] = {4, 8};
  if (const char *e = std::getenv("PACMAP_GRAD_SELL_RANGES")) {
    if (e[0] == '8') order[0] = 8, order[1] = 0;
    if (e[0] == '4') order[1] = 0;
  }
  bool ok = false;
  for (uint32_t r : order)
    if (r && !ok) ok = sell_plan(static_cast<uint32_t>(n), static_cast<uint32_t>(dims), r, plan);
  if (!ok) return false;
  if (const char *e = std::getenv("PACMAP_GRAD_SELL")) {
    if (e[0] == '0') return false;
    if (e[0] == '1') return true;
  }
  return n >= 16384;
}

// The paged form of the target-blocked gradient (Y pages staged through LDS),
// used in place of it when its plan fits.  PACMAP_GRAD_PAGED=0 turns it off
// and =1 keeps it where its layout has many long runs (see build_grad_layout);
// PACMAP_GRAD_PAGE_ROWS / PACMAP_GRAD_PAGE_POINTS cap the rows per chunk and
// the points per page (tuning and tests).
bool use_paged(uint64_t n, uint64_t dims, PagePlan &plan) {
  if (n >= kMaxPoints) return false;
  if (const char *e = std::getenv("PACMAP_GRAD_PAGED"))
    if (e[0] == '0') return false;
  uint32_t max_rows = 0, max_pp = 0;
  if (const char *e = std::getenv("PACMAP_GRAD_PAGE_ROWS")) max_rows = static_cast<uint32_t>(std::strtoul(e, nullptr, 10));
  if (const char *e = std::getenv("PACMAP_GRAD_PAGE_POINTS")) max_pp = static_cast<uint32_t>(std::strtoul(e, nullptr, 10));
  if (!page_plan(static_cast<uint32_t>(n), static_cast<uint32_t>(dims), max_rows, max_pp, plan)) return false;
  // every chunk copies all of Y through LDS once per iteration: past about
  // 6 GB of copies (n around 1.5 million at 2 dims) the gather form is kept
  return static_cast<uint64_t>(plan.chunks) * n * dims * 4 <= (6ull << 30);
}

// The gradient kernel for a CSR and its per-fit layout: the target-blocked
// kernel when Y outgrows an XCD's L2 (in its paged form when the plan fits),
// its LDS-staged (sliced ELL) form when a whole target range fits in LDS,
// else the one-pass kernel (see kernels.hip, grad_paged.hip).  The optimiser
// and the gradient seam both build it, so they run the same kernel.
struct GradLayout {
  bool tb = false, sell = false, paged = false;
  SellPlan sp;
  PagePlan pg;
  DevBuf<uint32_t> tb_ent, sell_perm, sell_off, sell_ent, page_off, page_ent;
  DevBuf<uint16_t> tb_off;
  DevBuf<float> tb_part;
  uint32_t loss_blocks(uint32_t n) const {
    if (paged) return page_grad_blocks(pg);
    return sell ? sell_grad_blocks(sp) : (tb ? grad_tb_blocks(n) : grad_blocks(n));
  }
  const uint32_t *entries(const DevCsr &csr) const { return tb && !sell && !paged ? tb_ent.get() : csr.ent.get(); }
  hipError_t launch(const GradArgs &a, uint32_t dims, bool adam, hipStream_t s) const {
    if (paged)
      return launch_grad_adam_paged(a, pg, page_off.get(), page_ent.get(), tb_part.get(), dims, adam, s);
    if (sell)
      return launch_grad_adam_sell(a, sp, sell_perm.get(), sell_off.get(), sell_ent.get(), tb_part.get(), dims,
                                   adam, s);
    if (tb) return launch_grad_adam_tb(a, tb_off.get(), tb_part.get(), dims, adam, s);
    return launch_grad_adam(a, dims, adam, s);
  }
};

void build_grad_layout(const DevCsr &csr, uint64_t n, uint64_t dims, bool want, GradLayout &L, hipStream_t s) {
  const uint64_t nd = n * dims;
  SellPlan &sp = L.sp;
  bool tb = want && use_target_blocks(n, dims) && csr.nnz > 0;
  bool sell = want && !tb && csr.nnz > 0 && use_sell(n, dims, sp);
  bool paged = tb && use_paged(n, dims, L.pg);
  if (tb || sell) {
    L.tb_ent.alloc(csr.nnz, kSlotTbEnt);
    L.tb_off.alloc(static_cast<size_t>(kTargetBlocks) * n, kSlotTbOff);
    DevBuf<uint32_t> bad(1);
    PM_HIP(launch_fill_zero(bad.get(), bad.bytes(), s));
    PM_HIP(launch_tb_partition(csr.row_ptr.get(), csr.ent.get(), static_cast<uint32_t>(n), L.tb_ent.get(),
                               L.tb_off.get(), bad.get(), s));
    uint32_t hb = 0;
    download(&hb, bad.get(), 1, s);
    PM_HIP(hipStreamSynchronize(s));
    if (hb) {
      tb = sell = paged = false;  // a row too long for the u16 range offsets: one-pass kernel
    } else {
      L.tb_part.alloc(static_cast<size_t>(kTargetBlocks) * nd, kSlotTbPart);
    }
  }
  if (paged) {
    PagePlan &pg = L.pg;
    const uint32_t n32 = static_cast<uint32_t>(n);
    const uint64_t runs = page_runs(pg);
    L.page_off.alloc(runs + 1, kSlotPageOff);
    DevBuf<uint32_t> rsz(runs), cur(runs), scr((runs + 1023) / 1024 + 1);
    PM_HIP(launch_page_count(csr.row_ptr.get(), L.tb_ent.get(), L.tb_off.get(), n32, pg, rsz.get(), s));
    // the scan is 32-bit and the gradient pass addresses words with 32-bit
    // byte offsets: the run sizes are summed in 64 bits first, and a layout
    // past kPageMaxWords keeps the tb kernel
    std::vector<uint32_t> hsz(runs);
    download(hsz.data(), rsz.get(), runs, s);
    PM_HIP(hipStreamSynchronize(s));
    uint64_t total = 0, long_runs = 0;
    for (uint32_t v : hsz) {
      total += v;
      long_runs += v > 64u * kPageRunSteps ? 1u : 0u;
    }
    // a run longer than kPageRunSteps steps stalls its workgroup for a memory
    // round trip (data whose neighbours sit close in index, hub rows): when
    // more than 1% of the runs are, the gather form is faster
    const char *force = std::getenv("PACMAP_GRAD_PAGED");
    const bool forced = force && force[0] == '1';
    if (total < kPageMaxWords && (forced || long_runs * 100 <= runs)) {
      PM_HIP(launch_csr_scan(rsz.get(), static_cast<uint32_t>(runs), L.page_off.get(), cur.get(), scr.get(), s));
      L.page_ent.alloc(total ? total : 1, kSlotPageEnt);
      pg.words = static_cast<uint32_t>(total ? total : 1);
      PM_HIP(launch_page_fill(csr.row_ptr.get(), L.tb_ent.get(), L.tb_off.get(), n32, pg, L.page_off.get(),
                              L.page_ent.get(), s));
      g_tm.grad_page_words = total;
      g_tm.grad_page_pages = pg.pages;
    } else {
      paged = false;
      L.page_off.release();
    }
  }
  if (sell) {
    const uint32_t n32 = static_cast<uint32_t>(n);
    const uint64_t ns_ = sell_slices(sp);
    L.sell_perm.alloc(static_cast<size_t>(sp.ranges) * n, kSlotSellPerm);
    L.sell_off.alloc(ns_ + 1, kSlotSellOff);
    DevBuf<uint32_t> ssz(ns_), cur(ns_), scr((ns_ + 1023) / 1024 + 1);
    PM_HIP(launch_sell_sort(csr.row_ptr.get(), L.tb_off.get(), n32, sp, L.sell_perm.get(), ssz.get(), s));
    PM_HIP(launch_csr_scan(ssz.get(), static_cast<uint32_t>(ns_), L.sell_off.get(), cur.get(), scr.get(), s));
    uint32_t total = 0;
    download(&total, L.sell_off.get() + ns_, 1, s);
    PM_HIP(hipStreamSynchronize(s));
    L.sell_ent.alloc(total ? total : 1, kSlotSellEnt);
    sp.words = total ? total : 1;
    PM_HIP(launch_sell_fill(csr.row_ptr.get(), L.tb_ent.get(), L.tb_off.get(), n32, sp, L.sell_perm.get(),
                            L.sell_off.get(), L.sell_ent.get(), s));
    g_tm.grad_sell_entries = total;
    g_tm.grad_sell_ranges = sp.ranges;
  }
  L.tb = tb;
  L.sell = sell;
  L.paged = paged;
  g_tm.grad_target_blocks = tb ? kTargetBlocks : 0;
  g_tm.grad_sell = sell ? 1 : 0;
  g_tm.grad_paged = paged ? 1 : 0;
}

// The initial Y is y (host), or y_dev (device, n x dims) when that is not
// null; y ends as the final Y (y.size() >= n dims).
void optimize_device(const DevCsr &csr, uint64_t n, uint64_t dims, std::vector<float> &y, const float *y_dev,
                     float lr, const uint64_t iters[3], const uint64_t *snap_idx, uint64_t ns,
                     float *snaps_out, float *loss_out, bool record_loss, hipStream_t s) {
  const uint64_t T = iters[0] + iters[1] + iters[2];
  const uint64_t nd = n * dims;
  // snapshot slots (lib.rs:830-836): slot k is written at the iteration whose
  // itr+1 equals snapshots[k], when k is the first position with that value;
  // slot 0 starts as the initial Y (lib.rs:789-791)
  std::vector<int64_t> writer(T, -1);
  std::vector<bool> written(ns, false);
  for (uint64_t sidx = 0; sidx < ns; ++sidx) {
    const uint64_t v = snap_idx[sidx];
    if (v >= 1 && v <= T && writer[v - 1] < 0) {
      writer[v - 1] = static_cast<int64_t>(sidx);
      written[sidx] = true;
    }
  }
  if (snaps_out && ns) {
    std::memset(snaps_out, 0, sizeof(float) * ns * nd);
    if (y_dev) {
      download(snaps_out, y_dev, nd, s);
      PM_HIP(hipStreamSynchronize(s));
    } else {
      std::memcpy(snaps_out, y.data(), sizeof(float) * nd);
    }
  }
  // the two Y buffers start 256-byte aligned (the gradient kernels stage
  // Y with 16-byte loads)
  const uint64_t ystride = (nd + 63) / 64 * 64;
  DevBuf<float> ybuf(2 * ystride, kSlotY), m(nd, kSlotM), v(nd, kSlotV);
  if (y_dev) {
    if (nd) PM_HIP(hipMemcpyAsync(ybuf.get(), y_dev, nd * sizeof(float), hipMemcpyDeviceToDevice, s));
  } else {
    upload(ybuf, y.data(), nd, s);
  }
  PM_HIP(launch_fill_zero(m.get(), m.bytes(), s));
  PM_HIP(launch_fill_zero(v.get(), v.bytes(), s));
  DevBuf<float> d_snap(ns && snaps_out ? ns * nd : 0, kSlotSnap);
  GradLayout lay;
  build_grad_layout(csr, n, dims, T > 0, lay, s);
  const bool tb = lay.tb, sell = lay.sell;
  SellPlan &sp = lay.sp;
  const uint32_t nb = lay.loss_blocks(static_cast<uint32_t>(n));
  DevBuf<float> part(record_loss ? T * nb : 0, kSlotLossPart), d_loss(record_loss ? T : 0);
  if (T == 0 && y_dev) download(y.data(), y_dev, nd, s);
  PM_HIP(hipStreamSynchronize(s));
  g_tm.iterations = T;
  if (T == 0) return;

  const float beta1 = 0.9f, beta2 = 0.999f, w_mn_init = 1000.0f;  // lib.rs:782-784
  const double tg = now_ms();
  std::vector<GradArgs> plan(T);
  for (uint64_t it = 0; it < T; ++it) {
    GradArgs &a = plan[it];
    a.row_ptr = csr.row_ptr.get();
    a.ent = lay.entries(csr);
    a.mean_row_len = static_cast<float>(static_cast<double>(csr.nnz) / static_cast<double>(n));
    a.y_in = ybuf.get() + (it % 2) * ystride;
    a.y_out = ybuf.get() + ((it + 1) % 2) * ystride;
    a.m = m.get();
    a.v = v.get();
    a.n = static_cast<uint32_t>(n);
    find_weights(w_mn_init, it, iters[0], iters[1], a.w);
    a.lr_t = adam_lr_t(lr, beta1, beta2, it);
    a.omb1 = 1.0f - beta1;
    a.omb2 = 1.0f - beta2;
    a.loss_part = record_loss ? part.get() + it * nb : nullptr;
    a.snap = (writer[it] >= 0 && d_snap.get()) ? d_snap.get() + writer[it] * nd : nullptr;
  }
  std::vector<uint64_t> key;
  {
    KeyWriter kw{key};
    kw.add(static_cast<uint64_t>(sell ? 2 : (tb ? 1 : 0))), kw.add(dims), kw.add(T);
    kw.add(static_cast<uint64_t>(sp.ranges)), kw.add(static_cast<uint64_t>(sp.group)),
        kw.add(static_cast<uint64_t>(sp.span)), kw.add(static_cast<uint64_t>(sp.rc)),
        kw.add(static_cast<uint64_t>(sp.chunks)), kw.add(static_cast<uint64_t>(sp.smax)),
        kw.add(static_cast<uint64_t>(sp.words));
    kw.add(lay.sell_perm.get()), kw.add(lay.sell_off.get()), kw.add(lay.sell_ent.get()), kw.add(lay.tb_off.get()),
        kw.add(lay.tb_part.get());
    for (const GradArgs &a : plan) kw.add(a);
  }
  if (!(g_graph.e && g_graph.key == key)) {
    g_graph.reset();
    PM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t cap_err = hipSuccess;
    for (uint64_t it = 0; it < T && cap_err == hipSuccess; ++it) {
      cap_err = lay.launch(plan[it], static_cast<uint32_t>(dims), true, s);
    }
    const hipError_t end_err = hipStreamEndCapture(s, &g_graph.g);
    if (cap_err != hipSuccess) fail_gpu(cap_err, "grad_adam launch during capture", __FILE__, __LINE__);
    if (end_err != hipSuccess) fail_gpu(end_err, "hipStreamEndCapture", __FILE__, __LINE__);
    PM_HIP(hipGraphInstantiate(&g_graph.e, g_graph.g, nullptr, nullptr, 0));
    g_graph.key = std::move(key);
    g_tm.graph_reused = 0;
  } else {
    g_tm.graph_reused = 1;
  }
  g_tm.graph_build_ms += now_ms() - tg;
  Timer opt;
  opt.start(s);
  PM_HIP(hipGraphLaunch(g_graph.e, s));
  opt.stop(s);
  if (record_loss) PM_HIP(launch_loss_reduce(part.get(), nb, static_cast<uint32_t>(T), d_loss.get(), s));
  download(y.data(), ybuf.get() + (T % 2) * ystride, nd, s);
  std::vector<float> loss(record_loss ? T : 0);
  if (record_loss) download(loss.data(), d_loss.get(), T, s);
  PM_HIP(hipStreamSynchronize(s));
  opt.add_to(g_tm.optimize_ms);
  if (loss_out) {
    if (record_loss)
      std::memcpy(loss_out, loss.data(), sizeof(float) * T);
    else
      std::fill(loss_out, loss_out + T, std::nanf(""));
  }
  if (snaps_out && ns) {
    for (uint64_t sidx = 0; sidx < ns; ++sidx)
      if (written[sidx]) download(snaps_out + sidx * nd, d_snap.get() + sidx * nd, nd, s);
    PM_HIP(hipStreamSynchronize(s));
  }
}

// ---------------------------------------------------------------------------
// PCA: top-q principal axes of the rows of x, all on the GPU: FP64
// covariance, then (d <= 112) a Jacobi eigensolve of it, or (d > 112) a block
// subspace iteration with CholeskyQR2 followed by a Jacobi eigensolve of the
// Rayleigh-Ritz matrix.  Axes are sorted by variance, each signed so its
// largest-magnitude loading is positive.  Outputs stay on the device.
// ---------------------------------------------------------------------------
struct PcaAxes {
  DevBuf<float> Vf;    // d x q (FP32 copy for the projection)
  DevBuf<float> mean;  // d
};

// FP64 column means and covariance sum of the rows of x, by the row chunks
// of cov_chunks (see launch_pca_stats_chunks).
struct PcaStats {
  DevBuf<double> mean, cov, shift, colpart, covpart;
  CovChunks c;
  void alloc(uint64_t n, uint32_t d) {
    c = cov_chunks(n, d);
    mean.alloc(d);
    shift.alloc(d);
    cov.alloc(static_cast<size_t>(d) * d);
    colpart.alloc(colsum_chunk_scratch_doubles(c, d));
    covpart.alloc(cov_chunk_scratch_doubles(c, d));
  }
  void finish(uint64_t n, uint32_t d, hipStream_t s) {
    PM_HIP(launch_pca_stats_finish(n, d, c, colpart.get(), shift.get(), covpart.get(), mean.get(), cov.get(), s));
    colpart.release();
    covpart.release();
  }
};

// X already in device memory: every chunk in one launch.
void pca_stats_dev(const float *x_dev, uint64_t n, uint32_t d, PcaStats &st, hipStream_t s) {
  Timer tc;
  tc.start(s);
  st.alloc(n, d);
  PM_HIP(launch_pca_stats_chunks(x_dev, n, d, st.c, 0, st.c.chunks, st.colpart.get(), st.shift.get(),
                                 st.covpart.get(), s));
  st.finish(n, d, s);
  tc.stop(s);
  tc.add_to(g_tm.pca_cov_ms);
}

// X in host memory: uploaded in the statistics' row chunks on the copy
// stream; each chunk's sums run on s as soon as it lands, while the next
// chunk is in flight.  The sums are the same bits as pca_stats_dev's.
void upload_pca_stats(const float *x_host, uint64_t n, uint32_t d, float *x_dev, PcaStats &st, hipStream_t s) {
  CopyStream cs;
  st.alloc(n, d);
  Timer tu, tc;
  tu.start(s);
  upload_rows_pieces(x_dev, x_host, n, static_cast<size_t>(d) * sizeof(float), st.c.rows, s, cs.get(),
                     [&](uint32_t k, hipEvent_t landed) {
                       PM_HIP(hipStreamWaitEvent(s, landed, 0));
                       if (k + 1 == st.c.chunks) tc.start(s);  // the sums left after the last piece landed
                       PM_HIP(launch_pca_stats_chunks(x_dev, n, d, st.c, k, 1, st.colpart.get(), st.shift.get(),
                                                      st.covpart.get(), s));
                     });
  tu.stop(cs.get());
  st.finish(n, d, s);
  tc.stop(s);
  tu.add_to(g_tm.upload_ms);
  tc.add_to(g_tm.pca_cov_ms);
}

// The subspace iteration's random start (uniform in [-0.5, 0.5), SmallRng
// seeded 0x5eed, row-major d x b), drawn and uploaded once per thread and
// shape, then kept on the device for the thread's lifetime.
struct PcaStartBlock {
  double *p = nullptr;
  uint32_t d = 0, b = 0;
  int dev = -1;
  ~PcaStartBlock() {
    if (p && !process_exiting()) (void)hipFree(p);
  }
};
const double *pca_start_block(uint32_t d, uint32_t b) {
  static thread_local PcaStartBlock c;
  int dev = 0;
  PM_HIP(hipGetDevice(&dev));
  if (c.p && c.d == d && c.b == b && c.dev == dev) return c.p;
  std::vector<double> q0(static_cast<size_t>(d) * b);
  SmallRng rng = SmallRng::seed_from_u64(0x5eedull);
  for (auto &e : q0) e = rng.next_f64() - 0.5;
  if (c.p) {
    PM_HIP(hipSetDevice(c.dev));
    PM_HIP(hipDeviceSynchronize());  // earlier calls' copies out of it are done
    PM_HIP(hipFree(c.p));
    PM_HIP(hipSetDevice(dev));
    c.p = nullptr;
  }
  note_runtime_use();
  PM_HIP(hipMalloc(reinterpret_cast<void **>(&c.p), q0.size() * sizeof(double)));
  PM_HIP(hipMemcpy(c.p, q0.data(), q0.size() * sizeof(double), hipMemcpyHostToDevice));
  c.d = d;
  c.b = b;
  c.dev = dev;
  return c.p;
}

void pca_axes_dev(PcaStats &stats, uint64_t n, uint32_t d, uint32_t q, PcaAxes &out, hipStream_t s) {
  constexpr uint32_t kIters = 6;
  const uint32_t jmax = jacobi_max_size();
  (void)n;
  double *const d_cov = stats.cov.get();
  double *const d_mean = stats.mean.get();
  Timer te;
  te.start(s);
  DevBuf<double> V(static_cast<size_t>(d) * q);
  DevBuf<uint32_t> status(1);
  PM_HIP(launch_fill_zero(status.get(), status.bytes(), s));
  DevBuf<double> jwork(jacobi_work_doubles(std::min(d, jmax)));
  if (d <= jmax) {
    PM_HIP(launch_jacobi(d_cov, d, q, V.get(), jwork.get(), s));
  } else {
    const uint32_t b = std::min<uint32_t>(std::min<uint32_t>(d, q + 12), jmax);
    if (q > b) fail(PACMAP_ERR_PCA, "too many principal components requested");
    const size_t qn = static_cast<size_t>(d) * b;
    DevBuf<double> Q(qn), Z(qn), G(static_cast<size_t>(b) * b), Rinv(static_cast<size_t>(b) * b),
        U(static_cast<size_t>(b) * q);
    PM_HIP(hipMemcpyAsync(Q.get(), pca_start_block(d, b), qn * sizeof(double), hipMemcpyDeviceToDevice, s));
    auto gemm = [&](uint32_t M, uint32_t N, uint32_t K, const double *A, uint32_t lda, bool tA, const double *B,
                    uint32_t ldb, double *C, uint32_t ldc) {
      PM_HIP(launch_gemm64(M, N, K, A, lda, tA, B, ldb, C, ldc, s));
    };
    // Q <- orthonormal basis of the columns of Z by CholeskyQR (Q = Z R^-1);
    // `twice` repeats it on Q (CholeskyQR2) for full orthogonality.  Inside
    // the iteration one pass suffices: its loss of orthogonality (~u cond^2)
    // does not accumulate, and the last pass is doubled.  Z is overwritten.
    auto orth = [&](bool twice) {
      gemm(b, b, d, Z.get(), b, true, Z.get(), b, G.get(), b);
      PM_HIP(launch_chol_inv(G.get(), b, Rinv.get(), status.get(), s));
      gemm(d, b, b, Z.get(), b, false, Rinv.get(), b, Q.get(), b);
      if (!twice) return;
      gemm(b, b, d, Q.get(), b, true, Q.get(), b, G.get(), b);
      PM_HIP(launch_chol_inv(G.get(), b, Rinv.get(), status.get(), s));
      gemm(d, b, b, Q.get(), b, false, Rinv.get(), b, Z.get(), b);
      PM_HIP(hipMemcpyAsync(Q.get(), Z.get(), Z.bytes(), hipMemcpyDeviceToDevice, s));
    };
    // the random start is not orthonormalised: the iteration depends only on
    // its span, and a d x b uniform random block is well conditioned
    // (singular values within ~(sqrt d +- sqrt b))
    // the covariance is exactly symmetric (cov_fin_kernel writes both halves
    // from one sum), so it is read as transposed: 128-byte rows per load
    for (uint32_t it = 0; it < kIters; ++it) {
      gemm(d, b, d, d_cov, d, true, Q.get(), b, Z.get(), b);
      orth(it + 1 == kIters);
    }
    // Rayleigh-Ritz: T = Q^T C Q, eigenvectors U, axes V = Q U
    gemm(d, b, d, d_cov, d, true, Q.get(), b, Z.get(), b);
    gemm(b, b, d, Q.get(), b, true, Z.get(), b, G.get(), b);
    PM_HIP(launch_jacobi(G.get(), b, q, U.get(), jwork.get(), s));
    gemm(d, q, b, Q.get(), b, false, U.get(), q, V.get(), q);
  }
  out.Vf.alloc(static_cast<size_t>(d) * q);
  out.mean.alloc(d);
  PM_HIP(launch_sign_fix(V.get(), d, q, out.Vf.get(), d_mean, out.mean.get(), s));
  te.stop(s);
  // checked when the entry point returns (the status stays 0 unless the
  // orthonormalisation failed even with a shift)
  defer_flag_check(status.get(), PACMAP_ERR_PCA, "subspace orthonormalisation failed", s);
  te.add_to(g_tm.pca_eig_ms);
}

void pca_axes_dev(const float *x_dev, uint64_t n, uint32_t d, uint32_t q, PcaAxes &out, hipStream_t s) {
  PcaStats st;
  pca_stats_dev(x_dev, n, d, st, s);
  pca_axes_dev(st, n, d, q, out, s);
}

// out (n x q) = (x - mean) V on the device
void pca_project_dev(const float *x_dev, uint64_t n, uint32_t d, uint32_t q, const PcaAxes &ax,
                     DevBuf<float> &out, hipStream_t s) {
  out.alloc(n * q);
  DevBuf<float> work(std::max<uint64_t>(project_workspace_floats(d, q), 4));
  PM_HIP(launch_project(x_dev, n, d, ax.mean.get(), ax.Vf.get(), q, work.get(), out.get(), s));
}

// ---------------------------------------------------------------------------
// Value initialisation (lib.rs:753-766) with ndarray 0.16's reduction orders
// ---------------------------------------------------------------------------
void value_init(const float *yin, uint64_t n, uint64_t dims, std::vector<float> &y) {
  std::vector<float> mean(dims, 0.0f), wm(dims, 0.0f), ssq(dims, 0.0f);
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t d = 0; d < dims; ++d) mean[d] = mean[d] + yin[i * dims + d];
  for (uint64_t d = 0; d < dims; ++d) mean[d] = mean[d] / static_cast<float>(n);
  for (uint64_t i = 0; i < n; ++i) {
    const float count = static_cast<float>(i + 1);
    for (uint64_t d = 0; d < dims; ++d) {
      const float x = yin[i * dims + d];
      const float delta = x - wm[d];
      wm[d] = wm[d] + delta / count;
      ssq[d] = std::fma(x - wm[d], delta, ssq[d]);
    }
  }
  for (uint64_t d = 0; d < dims; ++d) ssq[d] = std::sqrt(ssq[d] / static_cast<float>(n));
  y.resize(n * dims);
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t d = 0; d < dims; ++d) {
      const float c = yin[i * dims + d] - mean[d];
      const float sc = c * 0.0001f;
      y[i * dims + d] = sc / ssq[d];
    }
}

// Random(seed) initialisation (lib.rs:769-778): Normal<f32>(0, 1) draws in
// row-major order from SmallRng::seed_from_u64(seed), each the f64
// StandardNormal sample rounded to f32, times 1, plus 0.
void random_init(uint64_t n, uint64_t dims, bool has_seed, uint64_t seed, std::vector<float> &y) {
  if (!has_seed) {
    std::random_device rd;
    seed = (static_cast<uint64_t>(rd()) << 32) ^ rd();
  }
  SmallRng rng = SmallRng::seed_from_u64(seed);
  y.resize(n * dims);
  for (uint64_t e = 0; e < n * dims; ++e) y[e] = 0.0f + 1.0f * static_cast<float>(rng.standard_normal());
}

// ---------------------------------------------------------------------------
// fit_transform (lib.rs:294-367)
// ---------------------------------------------------------------------------
// Outputs of the pair-generation entry point: fit_impl stops after sampling
// and returns the three pair lists (rows x 2) and the preprocessed features.
struct PairsOut {
  uint32_t *nn = nullptr, *mn = nullptr, *fp = nullptr;
  float *x_pre = nullptr;  // n x dk, optional
  uint64_t counts[4] = {0, 0, 0, 0};  // NN, MN, FP rows; dk
  bool query = false;                 // only fill counts
};

void fit_impl(const float *x_host, const float *x_dev_in, uint64_t n, uint64_t d,
              const pacmap_config *cfg, float *y_out, float *snaps_out, float *loss_out,
              PairsOut *po = nullptr) {
  const double t0 = now_ms();
  if (!cfg) fail(PACMAP_ERR_INVALID_ARGUMENT, "config is NULL");
  if (n <= 1) fail(PACMAP_ERR_SAMPLE_SIZE, "Sample size must be larger than one");  // lib.rs:300-302
  if (!y_out && !po) fail(PACMAP_ERR_INVALID_ARGUMENT, "y_out is NULL");
  if (!x_host && !x_dev_in && d) fail(PACMAP_ERR_INVALID_ARGUMENT, "x is NULL");
  const uint64_t dims = cfg->embedding_dimensions;
  if (dims == 0) fail(PACMAP_ERR_INVALID_ARGUMENT, "embedding_dimensions must be at least 1");
  if (dims > 8) fail(PACMAP_ERR_UNSUPPORTED, "embedding_dimensions above 8 is not supported");
  if (n >= kMaxPoints) fail(PACMAP_ERR_UNSUPPORTED, "too many points");
  const int init = cfg->init_kind;
  if (init < PACMAP_INIT_PCA || init > PACMAP_INIT_RANDOM) fail(PACMAP_ERR_INVALID_ARGUMENT, "unknown initialization");
  const bool pca_path = d > 100 && init == PACMAP_INIT_PCA;  // lib.rs:456
  if (!pca_path && d == 0) fail(PACMAP_ERR_EMPTY_ARRAY_MIN_MAX, "Failed to compute min or max of X: the array is empty");
  if (init == PACMAP_INIT_VALUE &&
      (!cfg->init_value || cfg->init_value_rows != n || cfg->init_value_cols != dims))
    fail(PACMAP_ERR_INVALID_ARGUMENT, "Initialization::Value must be n x embedding_dimensions");
  if (init == PACMAP_INIT_PCA && !pca_path && std::min<uint64_t>(n, dims) < dims)
    fail(PACMAP_ERR_INVALID_ARGUMENT, "PCA initialisation needs n >= embedding_dimensions");
  uint64_t counts[3];
  decide_num_pairs(n, cfg->has_override_neighbors != 0, cfg->override_neighbors, cfg->mid_near_ratio,
                   cfg->far_pair_ratio, counts);  // lib.rs:333-338
  const uint64_t k = counts[0], n_mn = counts[1], n_fp = counts[2];
  const int pk = cfg->pair_kind;
  if (pk < PACMAP_PAIRS_GENERATE || pk > PACMAP_PAIRS_ALL_PROVIDED) fail(PACMAP_ERR_INVALID_ARGUMENT, "unknown pair configuration");
  if (!knn_precision_valid(cfg->knn_precision)) fail(PACMAP_ERR_INVALID_ARGUMENT, "unknown kNN precision");
  if (pk == PACMAP_PAIRS_NEIGHBORS_PROVIDED) {  // lib.rs:651-658
    if (cfg->pair_neighbors_rows != n * k || cfg->pair_neighbors_cols != 2) {
      PmError e(PACMAP_ERR_INVALID_NEAREST_NEIGHBOR_SHAPE, "Invalid shape for nearest neighbor pairs");
      e.expected[0] = n * k;
      e.expected[1] = 2;
      e.actual[0] = cfg->pair_neighbors_rows;
      e.actual[1] = cfg->pair_neighbors_cols;
      throw e;
    }
  }
  if (pk != PACMAP_PAIRS_GENERATE && cfg->pair_neighbors_rows && !cfg->pair_neighbors)
    fail(PACMAP_ERR_INVALID_ARGUMENT, "pair_neighbors is NULL");
  if (pk == PACMAP_PAIRS_ALL_PROVIDED) {
    if (cfg->pair_neighbors_cols != 2 && cfg->pair_neighbors_rows)
      fail(PACMAP_ERR_INVALID_ARGUMENT, "pair arrays must have 2 columns");
    if ((cfg->pair_mn_rows && !cfg->pair_mn) || (cfg->pair_fp_rows && !cfg->pair_fp))
      fail(PACMAP_ERR_INVALID_ARGUMENT, "pair array is NULL");
  }
  if (cfg->has_snapshots && cfg->num_snapshots && !cfg->snapshots)
    fail(PACMAP_ERR_INVALID_ARGUMENT, "snapshots is NULL");
  if (x_host) require_host(x_host, "x");
  if (x_dev_in) require_device(x_dev_in, "x_dev");
  if (y_out) require_host(y_out, "y_out");
  if (snaps_out) require_host(snaps_out, "snapshots_out");
  if (loss_out) require_host(loss_out, "loss_out");
  if (init == PACMAP_INIT_VALUE) require_host(cfg->init_value, "init_value");
  if (pk != PACMAP_PAIRS_GENERATE) {
    require_host(cfg->pair_neighbors, "pair_neighbors");
    require_host(cfg->pair_mn, "pair_mn");
    require_host(cfg->pair_fp, "pair_fp");
  }
  if (po) {
    if (pk != PACMAP_PAIRS_GENERATE) fail(PACMAP_ERR_INVALID_ARGUMENT, "pair generation needs PairConfiguration::Generate");
    po->counts[0] = n * std::min<uint64_t>(k, n - 1);
    po->counts[1] = n * n_mn;
    po->counts[2] = n * n_fp;
    po->counts[3] = pca_path ? std::min<uint64_t>(100, n) : d;
    if (po->query) return;
    if (!po->nn || (n_mn && !po->mn) || (n_fp && !po->fp)) fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL pair output");
    require_host(po->nn, "pair_nn_out");
    require_host(po->mn, "pair_mn_out");
    require_host(po->fp, "pair_fp_out");
    require_host(po->x_pre, "x_pre_out");
  }

  Stream stream;
  const hipStream_t s = stream.get();
  DevBuf<float> d_xin;
  const float *xd = x_dev_in;
  // on the PCA path a host X is uploaded in row chunks whose FP64 sums run
  // while the next chunk is in flight
  PcaStats pca_stats;
  bool have_stats = false;
  if (x_host) {
    d_xin.alloc(n * d);
    xd = d_xin.get();
    if (pca_path) {
      upload_pca_stats(x_host, n, static_cast<uint32_t>(d), d_xin.get(), pca_stats, s);
      have_stats = true;
    } else {
      Timer tu;
      tu.start(s);
      upload(d_xin, x_host, n * d, s);
      tu.stop(s);
      tu.add_to(g_tm.upload_ms);
    }
  }

  // --- preprocess_x (lib.rs:442-532) ---
  Timer tp;
  tp.start(s);
  DevBuf<float> d_xk;
  uint64_t dk = d;
  std::vector<float> y0;
  bool have_y = false;
  if (pca_path) {
    const uint32_t q = static_cast<uint32_t>(std::min<uint64_t>(100, n));  // lib.rs:457
    if (dims > q) fail(PACMAP_ERR_INVALID_ARGUMENT, "embedding_dimensions exceeds the PCA width");
    PcaAxes ax;
    if (!have_stats) pca_stats_dev(xd, n, static_cast<uint32_t>(d), pca_stats, s);
    pca_axes_dev(pca_stats, n, static_cast<uint32_t>(d), q, ax, s);
    pca_project_dev(xd, n, static_cast<uint32_t>(d), q, ax, d_xk, s);
    dk = q;
  } else {
    d_xk.alloc(n * d);
    DevBuf<uint32_t> keys(2);
    DevBuf<float> mean(d);
    PM_HIP(launch_minmax(xd, n * d, keys.get(), s));
    PM_HIP(launch_normalize(xd, d_xk.get(), n * d, keys.get(), s));
    PM_HIP(launch_colmean_seq(d_xk.get(), n, static_cast<uint32_t>(d), mean.get(), s));
    PM_HIP(launch_sub_mean(d_xk.get(), n, static_cast<uint32_t>(d), mean.get(), s));
    if (init == PACMAP_INIT_PCA) {  // lib.rs:513-519 then lib.rs:768
      const uint32_t nc = static_cast<uint32_t>(std::min<uint64_t>(n, dims));
      if (nc > d) fail(PACMAP_ERR_PCA, "PCA needs at least as many features as components");
      PcaAxes ax;
      pca_axes_dev(d_xk.get(), n, static_cast<uint32_t>(d), nc, ax, s);
      DevBuf<float> proj;
      pca_project_dev(d_xk.get(), n, static_cast<uint32_t>(d), nc, ax, proj, s);
      y0.resize(n * dims);
      download(y0.data(), proj.get(), n * dims, s);
      PM_HIP(hipStreamSynchronize(s));
      for (auto &e : y0) e = e * 0.01f;
      have_y = true;
    }
  }
  tp.stop(s);
  tp.add_to(g_tm.preprocess_ms);

  // --- sample_pairs (lib.rs:629-692) ---
  DevCsr csr;
  uint64_t seed_base = cfg->seed;
  if (!cfg->has_seed) {  // thread_rng in the reference: a fresh stream per call
    std::random_device rd;
    seed_base = (static_cast<uint64_t>(rd()) << 32) ^ rd();
  }
  // MN/FP sampling on the device from the NN pairs `d_nn` (knn per point)
  auto sample_and_build = [&](const uint32_t *d_nn, uint32_t knn) {
    Timer tsmp;
    tsmp.start(s);
    DevBuf<uint32_t> d_mn(std::max<uint64_t>(2, 2 * n * n_mn)), d_fp(std::max<uint64_t>(2, 2 * n * n_fp));
    PM_HIP(launch_fill_zero(d_fp.get(), d_fp.bytes(), s));
    const KnnLayout Ls = knn_layout(static_cast<uint32_t>(dk));
    DevBuf<float> d_xs(n_mn ? n * Ls.Dp : 0);
    if (n_mn) PM_HIP(launch_knn_pack(d_xk.get(), d_xs.get(), n, Ls, s));
    DevBuf<uint32_t> d_ws(mn_stream_words(static_cast<uint32_t>(n), static_cast<uint32_t>(n_mn)));
    PM_HIP(launch_sample_pairs(d_xs.get(), Ls, static_cast<uint32_t>(n), d_nn, knn, static_cast<uint32_t>(n_mn),
                               static_cast<uint32_t>(n_fp), seed_base, cfg->has_seed ? 1 : 0, d_mn.get(),
                               d_fp.get(), d_ws.get(), s));
    tsmp.stop(s);
    tsmp.add_to(g_tm.sampling_ms);
    if (po) {
      download(po->nn, d_nn, 2 * n * knn, s);
      if (po->mn) download(po->mn, d_mn.get(), 2 * n * n_mn, s);
      if (po->fp) download(po->fp, d_fp.get(), 2 * n * n_fp, s);
      if (po->x_pre) download(po->x_pre, d_xk.get(), n * dk, s);
      PM_HIP(hipStreamSynchronize(s));
      return;
    }
    const uint32_t *dp[3] = {d_nn, d_mn.get(), d_fp.get()};
    const uint64_t rows[3] = {n * knn, n * n_mn, n * n_fp};
    build_csr_dev(n, dp, rows, csr, s);
  };
  if (pk == PACMAP_PAIRS_GENERATE) {  // neighbors.rs:56-108
    const uint32_t kp = static_cast<uint32_t>(std::min<uint64_t>(k + 50, n - 1));  // neighbors.rs:67
    const uint32_t knn = static_cast<uint32_t>(std::min<uint64_t>(k, n - 1));      // neighbors.rs:68
    // n >= approx_threshold selects USearch in the reference; the exact
    // neighbours computed here are a valid answer for that path too
    DevBuf<uint32_t> idx;
    DevBuf<float> dist;
    knn_device(d_xk.get(), n, dk, kp, idx, dist, s, cfg->knn_precision);
    Timer tsg;
    tsg.start(s);
    DevBuf<float> sig(n), scaled(n * kp);
    DevBuf<uint32_t> d_nn(std::max<uint64_t>(2, n * knn * 2));
    PM_HIP(launch_sigma(dist.get(), n, kp, sig.get(), s));
    PM_HIP(launch_scale_nn(dist.get(), idx.get(), sig.get(), n, kp, knn, scaled.get(), d_nn.get(), s));
    tsg.stop(s);
    tsg.add_to(g_tm.sigma_ms);
    sample_and_build(d_nn.get(), knn);
    if (po) return;
  } else if (pk == PACMAP_PAIRS_NEIGHBORS_PROVIDED) {  // lib.rs:651-676
    for (uint64_t r = 0; r < 2 * n * k; ++r)
      if (cfg->pair_neighbors[r] >= n) fail(PACMAP_ERR_INVALID_PAIR_INDEX, "pair_neighbors index out of range");
    DevBuf<uint32_t> d_nn(std::max<uint64_t>(2, 2 * n * k));
    upload(d_nn, cfg->pair_neighbors, 2 * n * k, s);
    sample_and_build(d_nn.get(), static_cast<uint32_t>(k));
  } else {  // lib.rs:679-690
    PairList lists[3] = {{cfg->pair_neighbors, cfg->pair_neighbors_rows},
                         {cfg->pair_mn, cfg->pair_mn_rows},
                         {cfg->pair_fp, cfg->pair_fp_rows}};
    build_csr_from_host(n, lists, csr, s);
  }

  // --- YInit (lib.rs:752-779) ---
  DevBuf<float> y_init_dev;  // the initial Y when it is computed on the device
  if (pca_path) {  // YInit::Preprocessed: x[:, :dims] * 0.01 (lib.rs:767)
    y0.resize(n * dims);
    y_init_dev.alloc(n * dims);
    PM_HIP(launch_take_cols(d_xk.get(), n, static_cast<uint32_t>(dk), static_cast<uint32_t>(dims), 0.01f,
                            y_init_dev.get(), s));
  } else if (!have_y) {
    if (init == PACMAP_INIT_VALUE)
      value_init(cfg->init_value, n, dims, y0);
    else
      random_init(n, dims, cfg->init_random_has_seed != 0, cfg->init_random_seed, y0);
  }
  optimize_device(csr, n, dims, y0, y_init_dev.get(), cfg->learning_rate, cfg->num_iters,
                  cfg->has_snapshots ? cfg->snapshots : nullptr,
                  cfg->has_snapshots ? cfg->num_snapshots : 0, snaps_out, loss_out,
                  cfg->record_loss != 0 && loss_out != nullptr, s);
  std::memcpy(y_out, y0.data(), sizeof(float) * n * dims);
  g_tm.total_ms = now_ms() - t0;
}

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" {

void pacmap_config_default(pacmap_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->embedding_dimensions = 2;  // lib.rs:211-221
  cfg->init_kind = PACMAP_INIT_PCA;
  cfg->mid_near_ratio = 0.5f;
  cfg->far_pair_ratio = 2.0f;
  cfg->pair_kind = PACMAP_PAIRS_GENERATE;
  cfg->learning_rate = 1.0f;
  cfg->num_iters[0] = 100;
  cfg->num_iters[1] = 100;
  cfg->num_iters[2] = 250;
  cfg->approx_threshold = 8000;
  cfg->knn_precision = PACMAP_KNN_FP32;
  cfg->record_loss = 1;
}

int pacmap_fit_transform(const float *x, uint64_t n, uint64_t d, const pacmap_config *cfg,
                         float *y_out, float *snapshots_out, float *loss_out, pacmap_status *st) {
  g_tm = pacmap_timings{};
  return guarded(st, [&] { fit_impl(x, nullptr, n, d, cfg, y_out, snapshots_out, loss_out); });
}

int pacmap_fit_transform_dev(const float *x_dev, uint64_t n, uint64_t d, const pacmap_config *cfg,
                             float *y_out, float *snapshots_out, float *loss_out, pacmap_status *st) {
  g_tm = pacmap_timings{};
  return guarded(st, [&] { fit_impl(nullptr, x_dev, n, d, cfg, y_out, snapshots_out, loss_out); });
}

int pacmap_knn(const float *x, uint64_t n, uint64_t d, uint64_t k, int32_t precision, uint32_t *idx,
               float *dist, uint64_t *stats_out, pacmap_status *st) {
  g_tm = pacmap_timings{};
  return guarded(st, [&] {
    if (!knn_precision_valid(precision)) fail(PACMAP_ERR_INVALID_ARGUMENT, "unknown kNN precision");
    if (n == 0) return;  // knn.rs:141-143
    if (!x && d) fail(PACMAP_ERR_INVALID_ARGUMENT, "x is NULL");
    const uint64_t kk = std::min<uint64_t>(k, n - 1);  // knn.rs:146
    if (kk && (!idx || !dist)) fail(PACMAP_ERR_INVALID_ARGUMENT, "output is NULL");
    if (kk == 0) return;
    require_host(x, "x");
    require_host(idx, "idx");
    require_host(dist, "dist");
    Stream stream;
    const hipStream_t s = stream.get();
    DevBuf<float> dx(n * d);
    upload(dx, x, n * d, s);
    DevBuf<uint32_t> di;
    DevBuf<float> dd;
    knn_device(dx.get(), n, d, static_cast<uint32_t>(kk), di, dd, s, precision);
    download(idx, di.get(), n * kk, s);
    download(dist, dd.get(), n * kk, s);
    PM_HIP(hipStreamSynchronize(s));
    if (stats_out) {
      stats_out[0] = g_tm.knn_rescanned_rows;
      stats_out[1] = g_tm.knn_candidates;
    }
  });
}

int pacmap_knn_exact(const float *x, uint64_t n, uint64_t d, uint64_t k, uint32_t *idx, float *dist,
                     pacmap_status *st) {
  return pacmap_knn(x, n, d, k, PACMAP_KNN_FP32, idx, dist, nullptr, st);
}

int pacmap_grad(const float *y, uint64_t n, uint64_t dims, const uint32_t *pair_nn, uint64_t n_nn,
                const uint32_t *pair_mn, uint64_t n_mn, const uint32_t *pair_fp, uint64_t n_fp,
                float w_nn, float w_mn, float w_fp, float *grad_out, pacmap_status *st) {
  g_tm = pacmap_timings{};
  return guarded(st, [&] {
    if (dims == 0 || dims > 8) fail(PACMAP_ERR_UNSUPPORTED, "dims must be in 1..8");
    if (n == 0) return;
    if (!y || !grad_out) fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL array");
    if ((n_nn && !pair_nn) || (n_mn && !pair_mn) || (n_fp && !pair_fp))
      fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL pair array");
    for (const void *hp : {static_cast<const void *>(y), static_cast<const void *>(pair_nn),
                           static_cast<const void *>(pair_mn), static_cast<const void *>(pair_fp),
                           static_cast<const void *>(grad_out)})
      require_host(hp, "an array argument");
    PairList lists[3] = {{pair_nn, n_nn}, {pair_mn, n_mn}, {pair_fp, n_fp}};
    Stream stream;
    const hipStream_t s = stream.get();
    DevCsr csr;
    build_csr_from_host(n, lists, csr, s);
    // the kernel the optimiser runs at this size, in gradient mode; Y starts
    // 256-byte aligned as the optimiser's buffers do
    GradLayout lay;
    build_grad_layout(csr, n, dims, true, lay, s);
    DevBuf<float> dy(n * dims, kSlotY), dg(n * dims);
    upload(dy, y, n * dims, s);
    const uint32_t nb = lay.loss_blocks(static_cast<uint32_t>(n));
    DevBuf<float> part(nb), loss(1);
    GradArgs a;
    a.row_ptr = csr.row_ptr.get();
    a.ent = lay.entries(csr);
    a.mean_row_len = static_cast<float>(static_cast<double>(csr.nnz) / static_cast<double>(n));
    a.y_in = dy.get();
    a.grad_out = dg.get();
    a.loss_part = part.get();
    a.n = static_cast<uint32_t>(n);
    a.w[0] = w_nn;
    a.w[1] = w_mn;
    a.w[2] = w_fp;
    PM_HIP(lay.launch(a, static_cast<uint32_t>(dims), false, s));
    PM_HIP(launch_loss_reduce(part.get(), nb, 1, loss.get(), s));
    download(grad_out, dg.get(), n * dims, s);
    float lv = 0.f;
    download(&lv, loss.get(), 1, s);
    PM_HIP(hipStreamSynchronize(s));
    for (uint64_t c = 0; c < dims; ++c) grad_out[n * dims + c] = 0.0f;
    grad_out[n * dims] = lv;  // gradient.rs:72
  });
}

int pacmap_adam_step(float *y, const float *grad, float *m, float *v, uint64_t n, uint64_t dims,
                     float beta1, float beta2, float lr, uint64_t itr, pacmap_status *st) {
  return guarded(st, [&] {
    const uint64_t count = n * dims;
    if (count == 0) return;
    if (!y || !grad || !m || !v) fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL array");
    Stream stream;
    const hipStream_t s = stream.get();
    DevBuf<float> dy(count), dg(count), dm(count), dv(count);
    upload(dy, y, count, s);
    upload(dg, grad, count, s);
    upload(dm, m, count, s);
    upload(dv, v, count, s);
    PM_HIP(launch_adam(dy.get(), dg.get(), dm.get(), dv.get(), count, adam_lr_t(lr, beta1, beta2, itr),
                       1.0f - beta1, 1.0f - beta2, s));
    download(y, dy.get(), count, s);
    download(m, dm.get(), count, s);
    download(v, dv.get(), count, s);
    PM_HIP(hipStreamSynchronize(s));
  });
}

void pacmap_weights(float w_mn_init, uint64_t itr, uint64_t p1, uint64_t p2, float out[3]) {
  if (out) find_weights(w_mn_init, itr, p1, p2, out);
}

int pacmap_sigma_scale(const float *knn_dist, const uint32_t *knn_idx, uint64_t n, uint64_t kprime,
                       uint64_t n_neighbors, float *sigma_out, float *scaled_out, uint32_t *pairs_out,
                       pacmap_status *st) {
  return guarded(st, [&] {
    if (n == 0) return;
    if (kprime > 512) fail(PACMAP_ERR_UNSUPPORTED, "kprime above 512");
    if (n_neighbors > kprime) fail(PACMAP_ERR_INVALID_ARGUMENT, "n_neighbors exceeds kprime");
    if (kprime && (!knn_dist || !knn_idx)) fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL array");
    for (uint64_t e = 0; e < n * kprime; ++e)
      if (knn_idx[e] >= n) fail(PACMAP_ERR_INVALID_PAIR_INDEX, "neighbour index out of range");
    Stream stream;
    const hipStream_t s = stream.get();
    DevBuf<float> dd(n * kprime), dsig(n), dsc(n * kprime);
    DevBuf<uint32_t> di(n * kprime), dp(pairs_out ? n * n_neighbors * 2 : 0);
    upload(dd, knn_dist, n * kprime, s);
    upload(di, knn_idx, n * kprime, s);
    PM_HIP(launch_sigma(dd.get(), n, static_cast<uint32_t>(kprime), dsig.get(), s));
    PM_HIP(launch_scale_nn(dd.get(), di.get(), dsig.get(), n, static_cast<uint32_t>(kprime),
                           static_cast<uint32_t>(n_neighbors), dsc.get(), dp.get(), s));
    if (sigma_out) download(sigma_out, dsig.get(), n, s);
    if (scaled_out) download(scaled_out, dsc.get(), n * kprime, s);
    if (pairs_out) download(pairs_out, dp.get(), n * n_neighbors * 2, s);
    PM_HIP(hipStreamSynchronize(s));
  });
}

int pacmap_scale_dist(const float *knn_dist, const float *sig, const uint32_t *nbrs, uint64_t n,
                      uint64_t k, float *scaled_out, pacmap_status *st) {
  return guarded(st, [&] {
    if (n == 0 || k == 0) return;
    if (k > 512) fail(PACMAP_ERR_UNSUPPORTED, "k above 512");
    if (!knn_dist || !sig || !nbrs || !scaled_out) fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL array");
    for (uint64_t e = 0; e < n * k; ++e)
      if (nbrs[e] >= n) fail(PACMAP_ERR_INVALID_PAIR_INDEX, "neighbour index out of range");
    Stream stream;
    const hipStream_t s = stream.get();
    DevBuf<float> dd(n * k), dsig(n), dsc(n * k);
    DevBuf<uint32_t> di(n * k);
    upload(dd, knn_dist, n * k, s);
    upload(dsig, sig, n, s);
    upload(di, nbrs, n * k, s);
    PM_HIP(launch_scale_nn(dd.get(), di.get(), dsig.get(), n, static_cast<uint32_t>(k), 0, dsc.get(), nullptr, s));
    download(scaled_out, dsc.get(), n * k, s);
    PM_HIP(hipStreamSynchronize(s));
  });
}

int pacmap_generate_pairs(const float *x, uint64_t n, uint64_t d, const pacmap_config *cfg,
                          uint32_t *pair_nn_out, uint32_t *pair_mn_out, uint32_t *pair_fp_out,
                          float *x_pre_out, uint64_t counts_out[4], pacmap_status *st) {
  g_tm = pacmap_timings{};
  return guarded(st, [&] {
    PairsOut po;
    po.query = pair_nn_out == nullptr;
    po.nn = pair_nn_out;
    po.mn = pair_mn_out;
    po.fp = pair_fp_out;
    po.x_pre = x_pre_out;
    fit_impl(x, nullptr, n, d, cfg, nullptr, nullptr, nullptr, &po);
    if (counts_out)
      for (int i = 0; i < 4; ++i) counts_out[i] = po.counts[i];
  });
}

int pacmap_sample_pairs(const float *x, uint64_t n, uint64_t d, const uint32_t *pair_nn,
                        uint64_t n_neighbors, uint64_t n_mn, uint64_t n_fp, uint64_t seed,
                        uint32_t *pair_mn_out, uint32_t *pair_fp_out, pacmap_status *st) {
  return guarded(st, [&] {
    if (n == 0) return;
    if (n >= kMaxPoints) fail(PACMAP_ERR_UNSUPPORTED, "too many points");
    if ((n * d && !x) || (n * n_neighbors && !pair_nn) || (n * n_mn && !pair_mn_out) ||
        (n * n_fp && !pair_fp_out))
      fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL array");
    for (uint64_t r = 0; r < 2 * n * n_neighbors; ++r)
      if (pair_nn[r] >= n) fail(PACMAP_ERR_INVALID_PAIR_INDEX, "pair_nn index out of range");
    Stream stream;
    const hipStream_t s = stream.get();
    DevBuf<float> dx(std::max<uint64_t>(1, n * d));
    DevBuf<uint32_t> dnn(std::max<uint64_t>(2, 2 * n * n_neighbors)), dmn(std::max<uint64_t>(2, 2 * n * n_mn)),
        dfp(std::max<uint64_t>(2, 2 * n * n_fp));
    upload(dx, x, n * d, s);
    upload(dnn, pair_nn, 2 * n * n_neighbors, s);
    PM_HIP(launch_fill_zero(dfp.get(), dfp.bytes(), s));
    const KnnLayout L = knn_layout(static_cast<uint32_t>(d));
    DevBuf<float> dxs(n_mn ? n * L.Dp : 0);
    if (n_mn) PM_HIP(launch_knn_pack(dx.get(), dxs.get(), n, L, s));
    DevBuf<uint32_t> dws(mn_stream_words(static_cast<uint32_t>(n), static_cast<uint32_t>(n_mn)));
    PM_HIP(launch_sample_pairs(dxs.get(), L, static_cast<uint32_t>(n), dnn.get(),
                               static_cast<uint32_t>(n_neighbors), static_cast<uint32_t>(n_mn),
                               static_cast<uint32_t>(n_fp), seed, 1, dmn.get(), dfp.get(), dws.get(), s));
    download(pair_mn_out, dmn.get(), 2 * n * n_mn, s);
    download(pair_fp_out, dfp.get(), 2 * n * n_fp, s);
    PM_HIP(hipStreamSynchronize(s));
  });
}

int pacmap_optimize(float *y, uint64_t n, uint64_t dims, const uint32_t *pair_nn, uint64_t n_nn,
                    const uint32_t *pair_mn, uint64_t n_mn, const uint32_t *pair_fp, uint64_t n_fp,
                    float lr, const uint64_t num_iters[3], const uint64_t *snapshots,
                    uint64_t num_snapshots, float *snapshots_out, float *loss_out, pacmap_status *st) {
  g_tm = pacmap_timings{};
  return guarded(st, [&] {
    if (dims == 0 || dims > 8) fail(PACMAP_ERR_UNSUPPORTED, "dims must be in 1..8");
    if (!num_iters) fail(PACMAP_ERR_INVALID_ARGUMENT, "num_iters is NULL");
    if (n == 0) return;
    if (!y) fail(PACMAP_ERR_INVALID_ARGUMENT, "y is NULL");
    if ((n_nn && !pair_nn) || (n_mn && !pair_mn) || (n_fp && !pair_fp))
      fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL pair array");
    if (num_snapshots && !snapshots) fail(PACMAP_ERR_INVALID_ARGUMENT, "snapshots is NULL");
    for (const void *hp : {static_cast<const void *>(y), static_cast<const void *>(pair_nn),
                           static_cast<const void *>(pair_mn), static_cast<const void *>(pair_fp),
                           static_cast<const void *>(snapshots_out), static_cast<const void *>(loss_out)})
      require_host(hp, "an array argument");
    const double t0 = now_ms();
    PairList lists[3] = {{pair_nn, n_nn}, {pair_mn, n_mn}, {pair_fp, n_fp}};
    Stream stream;
    DevCsr csr;
    build_csr_from_host(n, lists, csr, stream.get());
    std::vector<float> yv(y, y + n * dims);
    optimize_device(csr, n, dims, yv, nullptr, lr, num_iters, snapshots, num_snapshots, snapshots_out, loss_out,
                    loss_out != nullptr, stream.get());
    std::memcpy(y, yv.data(), sizeof(float) * n * dims);
    g_tm.total_ms = now_ms() - t0;
  });
}

int pacmap_dev_alloc(uint64_t bytes, void **ptr_out, pacmap_status *st) {
  return guarded(st, [&] {
    if (!ptr_out) fail(PACMAP_ERR_INVALID_ARGUMENT, "ptr_out is NULL");
    *ptr_out = nullptr;
    PM_HIP(hipMalloc(ptr_out, bytes ? bytes : 1));
  });
}

int pacmap_dev_free(void *ptr) { return ptr ? (hipFree(ptr) == hipSuccess ? 0 : PACMAP_ERR_GPU) : 0; }

int pacmap_dev_upload(void *dst_dev, const void *src_host, uint64_t bytes, pacmap_status *st) {
  return guarded(st, [&] { PM_HIP(hipMemcpy(dst_dev, src_host, bytes, hipMemcpyHostToDevice)); });
}

int pacmap_dev_download(void *dst_host, const void *src_dev, uint64_t bytes, pacmap_status *st) {
  return guarded(st, [&] { PM_HIP(hipMemcpy(dst_host, src_dev, bytes, hipMemcpyDeviceToHost)); });
}

int pacmap_dev_copy(void *dst_dev, const void *src_dev, uint64_t bytes, uint32_t repeats, double *ms_out,
                    pacmap_status *st) {
  return guarded(st, [&] {
    if (!dst_dev || !src_dev) fail(PACMAP_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (bytes % 16) fail(PACMAP_ERR_INVALID_ARGUMENT, "bytes must be a multiple of 16");
    Stream stream;
    const hipStream_t s = stream.get();
    PM_HIP(launch_hbm_copy(dst_dev, src_dev, bytes, s));  // first touch of the pages, untimed
    Timer t;
    t.start(s);
    for (uint32_t r = 0; r < std::max<uint32_t>(1, repeats); ++r) PM_HIP(launch_hbm_copy(dst_dev, src_dev, bytes, s));
    t.stop(s);
    const double ms = t.ms();
    if (ms_out) *ms_out = ms / std::max<uint32_t>(1, repeats);
  });
}

void pacmap_last_timings(pacmap_timings *out) {
  if (out) *out = g_tm;
}

uint64_t pacmap_last_rescanned_rows(uint32_t *rows_out, uint64_t cap) {
  std::vector<uint32_t> v = g_rescanned.read();
  std::sort(v.begin(), v.end());
  if (rows_out)
    for (uint64_t i = 0; i < std::min<uint64_t>(cap, v.size()); ++i) rows_out[i] = v[i];
  return v.size();
}

int pacmap_device_info(char *name_out, uint64_t name_len, int32_t *cu_count_out) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return PACMAP_ERR_GPU;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return PACMAP_ERR_GPU;
  if (name_out && name_len) {
    std::strncpy(name_out, prop.gcnArchName, name_len - 1);
    name_out[name_len - 1] = 0;
  }
  if (cu_count_out) *cu_count_out = prop.multiProcessorCount;
  return 0;
}

const char *pacmap_version(void) { return "pacmap-mi355x 0.1.0 (gfx950)"; }

}  // extern "C"
