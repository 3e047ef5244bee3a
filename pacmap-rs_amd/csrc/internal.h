// internal.h -- launchers shared between the HIP kernels and the host library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace pm {

// Feature layout for the exact-order distance kernels.  distance.rs:50-75 sums
// (a-b)^2 into 8 lane accumulators over chunks of 8 features, adds the lanes
// 0..7 in order, then adds the tail features one by one.  The packed layout
// puts lane l's features (8c + l, c < C) contiguously in segment l of length
// Cp >= C (zero padded), then the tail, then zero padding up to Dp.  Zero
// padding is exact: (0 - 0)^2 = 0 and s + 0 = s.
struct KnnLayout {
  uint32_t d = 0;   // input features
  uint32_t C = 0;   // full 8-chunks (d / 8)
  uint32_t R = 0;   // tail features (d % 8)
  uint32_t Cp = 0;  // padded segment length (multiple of 4)
  uint32_t Dp = 0;  // padded row length (multiple of kKnnChunk)
};
constexpr uint32_t kKnnChunk = 16;  // features staged per LDS chunk
KnnLayout knn_layout(uint32_t d);

// Largest neighbour count the LDS top-k kernels hold per query.
constexpr uint32_t kKnnMaxK = 448;

hipError_t launch_knn_pack(const float *x, float *xp, uint64_t n, const KnnLayout &L,
                           hipStream_t s);
// Exact FP32 kNN in the reference's distance order; k = min(k, n-1) already.
// qlist (optional): only the nq listed query rows are computed.
hipError_t launch_knn_exact(const float *xp, uint64_t n, const KnnLayout &L, uint32_t k,
                            uint32_t *idx, float *dist, hipStream_t s,
                            const uint32_t *qlist = nullptr, uint32_t nq = 0);

// Pair sampling on the GPU (pairs.hip) from the points in the packed kNN
// layout (launch_knn_pack); pair arrays are rows x 2 u32 and pair_fp must be
// zeroed by the caller (unfilled far pairs stay (0, 0)).
// stream_ws holds mn_stream_words(n, n_mn) words (the MN sampler's
// precomputed random streams).
uint64_t mn_stream_words(uint32_t n, uint32_t n_mn);
hipError_t launch_sample_pairs(const float *xp, const KnnLayout &L, uint32_t n, const uint32_t *pair_nn,
                               uint32_t knn, uint32_t n_mn, uint32_t n_fp, uint64_t seed, int seeded,
                               uint32_t *pair_mn, uint32_t *pair_fp, uint32_t *stream_ws, hipStream_t s);
// Symmetrised CSR on the GPU: count degrees (flags out-of-range indices in
// bad[0]), scan to row offsets, fill, and sort each row by entry value.  Rows
// of up to 2048 entries are sorted by a wavefront each; longer rows are
// listed (long_count zeroed by the caller, long_list of long_cap >=
// entries / 2049 + 1 words) and sorted by a workgroup each.
hipError_t launch_csr_count(const uint32_t *const p[3], const uint64_t rows[3], uint32_t n,
                            uint32_t *deg, uint32_t *bad, hipStream_t s);
hipError_t launch_csr_scan(const uint32_t *deg, uint32_t n, uint32_t *row_ptr, uint32_t *cursor,
                           uint32_t *scratch /* (n + 1023) / 1024 + 1 */, hipStream_t s);
hipError_t launch_csr_fill_sort(const uint32_t *const p[3], const uint64_t rows[3], uint32_t n,
                                const uint32_t *row_ptr, uint32_t *cursor, uint32_t *ent,
                                uint32_t *long_count, uint32_t *long_list, uint32_t long_cap, hipStream_t s);

// Exact rescan of the nq listed rows with the candidate range split over S
// blocks per row; part holds nq x S x k partial keys.
hipError_t launch_knn_rescan(const float *xp, uint64_t n, const KnnLayout &L, uint32_t k,
                             const uint32_t *qlist, uint32_t nq, uint64_t *part, uint32_t S,
                             uint32_t *idx, float *dist, hipStream_t s);

// MFMA candidate path (knn_mfma.hip)
// The candidate kernels see the rows in an order perm (position -> row, see
// launch_knn_order); fragments, a and rinv are by position.
hipError_t launch_knn_frag(const float *x, uint64_t n, uint32_t d, const double *mean, uint32_t kch,
                           uint64_t ntiles, const uint32_t *perm, void *fhi, void *flo, hipStream_t s);
// FP8 E4M3 fragments of the centred rows scaled by 1 / rinv (128-feature chunks)
hipError_t launch_knn_frag8(const float *x, uint64_t n, uint32_t d, const double *mean, uint32_t kch,
                            uint64_t ntiles, const float *rinv, const uint32_t *perm, void *fa, void *fb,
                            hipStream_t s);
// arow = |c|^2 (1 - kappa) by position, arow_orig the same by row; rinv
// (optional) = max |c| / 448 by position
hipError_t launch_knn_rownorm(const float *x, uint64_t n, uint32_t d, const double *mean,
                              float one_minus_kappa, uint64_t rows_padded, const uint32_t *perm, float *arow,
                              float *arow_orig, uint32_t *nonfinite, float *rinv, hipStream_t s);
// perm: the rows sorted by the Morton code of their first 6 coordinates
// (mm: 12 words; key, key_sorted, idx: n words; temp: knn_order_temp_bytes(n))
size_t knn_order_temp_bytes(uint64_t n);
hipError_t launch_knn_order(const float *x, uint64_t n, uint32_t d, uint32_t *mm, uint32_t *key, uint32_t *key_sorted,
                            uint32_t *idx, uint32_t *perm, void *temp, size_t temp_bytes, hipStream_t s);
// Tile-pair pruning of the symmetric pass: box (16 floats per tile) bounds
// the first min(d, 8) coordinates of each tile's rows; after the tau
// sample, tmax (a float per tile) is the tile's largest tau and gtau_orig
// the taus by row; bits (words per query group) flags the candidate tiles
// a group can skip.
hipError_t launch_knn_tilebox(const float *x, uint64_t n, uint32_t d, const uint32_t *perm, uint64_t ntiles,
                              float *box, hipStream_t s);
hipError_t launch_knn_tiletau(const float *gtau, const uint32_t *perm, uint64_t n, uint64_t ntiles, float *tmax,
                              float *gtau_orig, hipStream_t s);
// gbox: 17 floats per query group (knn_cand_groups(ntiles)) of scratch;
// count (zeroed by the caller) accumulates the skipped (group, tile) pairs
hipError_t launch_knn_prune(const float *box, const float *tmax, uint64_t ntiles, uint32_t d, uint32_t words,
                            float *gbox, uint32_t *bits, unsigned long long *count, hipStream_t s);
// phase 0: tau per row from a 1/ss tile sample; phase 1: append candidates
// with lb <= tau to per-(row, split) lists (gcnt: n x nsplit, gcand:
// n x nsplit x capg words, twice that in mode 1, whose entries are
// (candidate, dot) pairs).  Query tiles are grouped knn_cand_groups(ntiles)
// ways.  Phase 2 (symmetric) computes each tile pair of the upper triangle
// once and appends to both rows' lists: gcnt n counters (zeroed first),
// gcand n x capg entries (nsplit = 1), one workgroup per item of
// knn_sym_plan.
// mode 0: bf16, 1: split bf16, 2: FP8 E4M3 with per-row scales (rinv).
uint32_t knn_cand_groups(uint32_t ntiles);
struct KnnSymPlan {
  uint32_t gt = 0;     // query tiles per group
  uint32_t L = 0;      // candidate tiles per window (m gt)
  uint32_t m = 0;
  uint32_t W = 0;      // windows
  uint32_t items = 0;  // workgroups (0: no plan)
};
KnnSymPlan knn_sym_plan(uint32_t ntiles, uint32_t target_items, uint32_t max_window);
// perm maps positions to rows (lists and counters are by row, gtau by
// position); prune (phase 2, optional) holds pwords words per query group.
// The tau sample's candidates: fragments and a values of `tiles` tiles of
// rows spread over the whole input (knn_sample_rows), so that the 16 rows
// of a sample tile are not neighbours in the candidate order.
struct KnnSampleSet {
  const uint4 *fhi = nullptr;
  const uint4 *flo = nullptr;
  const float *arow = nullptr;
  const float *rinv = nullptr;
  uint32_t tiles = 0;
};
struct KnnOrder {
  const uint32_t *perm = nullptr;
  const uint32_t *prune = nullptr;
  uint32_t pwords = 0;
  KnnSampleSet sample;
};
// rows[p] = (stride p) mod n for p < m: m rows spread over [0, n) (stride
// coprime to n, near 0.618 n)
uint32_t knn_sample_stride(uint64_t n);
hipError_t launch_knn_sample_rows(uint64_t n, uint64_t m, uint32_t stride, uint32_t *rows, hipStream_t s);
hipError_t launch_knn_cand(int phase, const void *fhi, const void *flo, const float *arow, const float *rinv,
                           uint32_t n, uint32_t kch, uint32_t ntiles, uint32_t ss, uint32_t capg,
                           uint32_t nsplit, uint32_t *gcnt, uint32_t *gcand, float *gtau,
                           int mode, hipStream_t s, const KnnSymPlan &sp, const KnnOrder &o);
// cert_scale < 0: accept every row with at least k listed candidates (no
// certificate).  dots: the lists hold (candidate, dot) pairs (mode 1); with
// ub_s1 > 0 as well, candidates whose lower bound exceeds the k-th smallest
// upper bound are dropped before the exact distances (arow: the rows' a).
// stats[2 s] / [2 s + 1], s = row mod 256, accumulate the listed and the
// exactly reranked candidates of the rows reranked (512 counters).
hipError_t launch_knn_rerank(const float *xp, const KnnLayout &L, uint32_t n, uint32_t k,
                             uint32_t capg, uint32_t nsplit, const uint32_t *gcnt,
                             const uint32_t *gcand, const float *gtau, double cert_scale,
                             const float *arow, bool dots, double ub_s1, double ub_s2,
                             uint32_t *idx, float *dist, uint32_t *fb_count, uint32_t *fb_list,
                             unsigned long long *stats, hipStream_t s);

hipError_t launch_sigma(const float *dist, uint64_t n, uint32_t kp, float *sig, hipStream_t s);
hipError_t launch_scale_nn(const float *dist, const uint32_t *idx, const float *sig, uint64_t n,
                           uint32_t kp, uint32_t knn, float *scaled, uint32_t *pairs,
                           hipStream_t s);

// CSR entry: bits 0..28 other index, bit 29 "owner" (the pair's first
// endpoint, counts the loss once), bits 30..31 pair type (0 NN, 1 MN, 2 FP).
constexpr uint32_t kIdxMask = (1u << 29) - 1u;
constexpr uint32_t kOwnerBit = 1u << 29;
constexpr uint32_t kMaxPoints = 1u << 29;

struct GradArgs {
  const uint32_t *row_ptr = nullptr;  // n + 1
  const uint32_t *ent = nullptr;      // row_ptr[n]
  const float *y_in = nullptr;        // n x dims
  float *y_out = nullptr;             // n x dims (Adam mode)
  float *m = nullptr;
  float *v = nullptr;
  float *grad_out = nullptr;          // n x dims (gradient mode)
  float *loss_part = nullptr;         // one float per block (may be null)
  float *snap = nullptr;              // optional copy of y_out
  uint32_t n = 0;
  float w[3] = {0, 0, 0};             // w_nn, w_mn, w_fp
  float lr_t = 0, omb1 = 0, omb2 = 0; // Adam scalars
  float mean_row_len = 0;             // row_ptr[n] / n (picks the load depth)
};
constexpr uint32_t kGradRowsPerBlock = 16;
uint32_t grad_blocks(uint32_t n);
// adam == false: write the gradient to grad_out instead of stepping.
hipError_t launch_grad_adam(const GradArgs &a, uint32_t dims, bool adam, hipStream_t s);
// Target-blocked gradient for large n (see kernels.hip): the CSR rows are
// re-laid out by tb_partition (entries grouped by kTargetBlocks index ranges,
// off: (kTargetBlocks) x n u16 range starts), then each iteration runs the
// per-range pass (loss_part: grad_tb_blocks(n) floats) and the combine+Adam
// pass (part: kTargetBlocks x n x dims floats).
constexpr uint32_t kTargetBlocks = 8;
uint32_t grad_tb_blocks(uint32_t n);
hipError_t launch_tb_partition(const uint32_t *row_ptr, const uint32_t *ent, uint32_t n, uint32_t *out,
                               uint16_t *off, uint32_t *bad, hipStream_t s);
hipError_t launch_grad_adam_tb(const GradArgs &a, const uint16_t *off, float *part, uint32_t dims, bool adam,
                               hipStream_t s);
// points per target range of the tb layout (a multiple of 4)
uint32_t tb_span(uint32_t n);
// pass 2 alone: part's kTargetBlocks range partials summed in range order,
// then Adam (or the gradient written to grad_out)
hipError_t launch_tb_combine(const GradArgs &a, const float *part, uint32_t dims, bool adam, hipStream_t s);
// Paged form of the target-blocked gradient (grad_paged.hip), for
// embeddings whose target range outgrows LDS: a workgroup owns a chunk of
// rows (their Y and gradient sums in LDS) and walks its target range in
// pages of Y staged through LDS.  Built once per fit from the tb_partition
// output: count (rsz: page_runs() run sizes), exclusive scan of rsz into
// roff (+1 total), fill (went: roff[last] words).  Each iteration runs
// grad_page (loss_part: page_grad_blocks() floats) and the tb combine.
struct PagePlan {
  uint32_t span = 0;       // points per target range (tb_span)
  uint32_t rc = 0;         // rows per chunk (a multiple of 4)
  uint32_t chunks = 0;
  uint32_t pp = 0;         // points per page (a multiple of 4)
  uint32_t pages = 0;      // pages per range
  uint32_t lds_bytes = 0;  // dynamic LDS per workgroup of the gradient pass
  uint32_t words = 0;      // entry words (set once the layout is built)
};
// max_rows / max_page_points: 0 = no cap beyond the LDS budget
bool page_plan(uint32_t n, uint32_t dims, uint32_t max_rows, uint32_t max_page_points, PagePlan &p);
uint64_t page_runs(const PagePlan &p);          // (range, chunk, page, wave) runs
uint32_t page_grad_blocks(const PagePlan &p);   // gradient pass workgroups (loss_part entries)
constexpr uint32_t kPageMaxWords = 1u << 29;    // entry words addressable by the gradient pass
constexpr uint32_t kPageRunSteps = 6;           // steps of a wave's run per page loaded ahead
hipError_t launch_page_count(const uint32_t *row_ptr, const uint32_t *tb_ent, const uint16_t *off, uint32_t n,
                             const PagePlan &p, uint32_t *rsz, hipStream_t s);
hipError_t launch_page_fill(const uint32_t *row_ptr, const uint32_t *tb_ent, const uint16_t *off, uint32_t n,
                            const PagePlan &p, const uint32_t *roff, uint32_t *went, hipStream_t s);
hipError_t launch_grad_adam_paged(const GradArgs &a, const PagePlan &p, const uint32_t *roff, const uint32_t *went,
                                  float *part, uint32_t dims, bool adam, hipStream_t s);
// LDS-staged variant of the target-blocked gradient (sliced ELL, see
// kernels.hip), used while a range of Y plus a row chunk fits in LDS.  Built
// once per fit from the tb_partition output: sort (perm: kTargetBlocks x n,
// ssz: sell_slices() slice sizes), exclusive scan of ssz into soff (+1
// total), fill (sent: soff[last] words).  Each iteration runs grad_sell
// (loss_part: sell_grad_blocks() floats) and a combine+Adam pass over the
// ranges' partials.
struct SellPlan {
  uint32_t ranges = 0;     // target ranges: 8 (the tb ranges) or 4 (tb ranges merged in pairs)
  uint32_t group = 0;      // tb ranges per SELL range
  uint32_t span = 0;       // points per target range
  uint32_t rc = 0;         // rows per chunk (multiple of 16)
  uint32_t chunks = 0;
  uint32_t smax = 0;       // slices per chunk
  uint32_t lds_bytes = 0;  // LDS per workgroup
  uint32_t words = 0;      // slice words (set once the layout is built)
};
bool sell_plan(uint32_t n, uint32_t dims, uint32_t ranges, SellPlan &p);
uint64_t sell_slices(const SellPlan &p);
uint32_t sell_blocks(const SellPlan &p);       // sort pass workgroups
uint32_t sell_grad_blocks(const SellPlan &p);  // gradient pass workgroups (loss_part entries)
hipError_t launch_sell_sort(const uint32_t *row_ptr, const uint16_t *off, uint32_t n, const SellPlan &p,
                            uint32_t *perm, uint32_t *ssz, hipStream_t s);
hipError_t launch_sell_fill(const uint32_t *row_ptr, const uint32_t *tb_ent, const uint16_t *off, uint32_t n,
                            const SellPlan &p, const uint32_t *perm, const uint32_t *soff, uint32_t *sent,
                            hipStream_t s);
hipError_t launch_grad_adam_sell(const GradArgs &a, const SellPlan &p, const uint32_t *perm, const uint32_t *soff,
                                 const uint32_t *sent, float *part, uint32_t dims, bool adam, hipStream_t s);
// loss[t] = sum of part[t * nb .. t * nb + nb) in a fixed order.
hipError_t launch_loss_reduce(const float *part, uint32_t nb, uint32_t iters, float *loss,
                              hipStream_t s);

// Preprocessing (lib.rs:486-511)
hipError_t launch_minmax(const float *x, uint64_t count, uint32_t *keys /*2*/, hipStream_t s);
hipError_t launch_normalize(const float *x, float *out, uint64_t count, const uint32_t *keys,
                            hipStream_t s);
hipError_t launch_colmean_seq(const float *x, uint64_t n, uint32_t d, float *mean, hipStream_t s);
hipError_t launch_sub_mean(float *x, uint64_t n, uint32_t d, const float *mean, hipStream_t s);

// PCA pieces (FP64 covariance, FP32 projection)
hipError_t launch_colmean_f64(const float *x, uint64_t n, uint32_t d, double *mean,
                              double *scratch, hipStream_t s);
uint64_t colmean_scratch_doubles(uint64_t n, uint32_t d);
// PCA statistics by row chunks of kCovChunkBytes (the upload pieces of a
// host X): chunks of `rows` rows, `per` covariance and `col_per` column-sum
// splits each.  launch_pca_stats_chunks sums chunks k0 .. k0 + nk - 1 (the
// first call must include chunk 0: it sets the shift); launch_pca_stats_finish
// then gives the FP64 mean and the covariance sum about it.  Launching every
// chunk at once or one chunk at a time gives the same bits.
constexpr uint64_t kCovChunkBytes = uint64_t{32} << 20;
struct CovChunks {
  uint64_t rows = 0;
  uint32_t chunks = 0, per = 0, col_per = 0;
};
CovChunks cov_chunks(uint64_t n, uint32_t d);
uint64_t cov_chunk_scratch_doubles(const CovChunks &c, uint32_t d);
uint64_t colsum_chunk_scratch_doubles(const CovChunks &c, uint32_t d);
hipError_t launch_pca_stats_chunks(const float *x, uint64_t n, uint32_t d, const CovChunks &c, uint32_t k0,
                                   uint32_t nk, double *colpart, double *shift, double *covpart, hipStream_t s);
hipError_t launch_pca_stats_finish(uint64_t n, uint32_t d, const CovChunks &c, const double *colpart,
                                   const double *shift, const double *covpart, double *mean, double *cov,
                                   hipStream_t s);
// out (n x q) = (x - mean) V; the workspace (project_workspace_floats) holds
// V and the mean packed in the order the FP32-MFMA kernel stages them
uint64_t project_workspace_floats(uint32_t d, uint32_t q);
hipError_t launch_project(const float *x, uint64_t n, uint32_t d, const float *mean,
                          const float *V, uint32_t q, float *workspace, float *out,
                          hipStream_t s);

hipError_t launch_adam(float *y, const float *g, float *m, float *v, uint64_t count, float lr_t,
                       float omb1, float omb2, hipStream_t s);
hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t s);
// dst = src, bytes a multiple of 16, both 16-byte aligned
hipError_t launch_hbm_copy(void *dst, const void *src, uint64_t bytes, hipStream_t s);
// dst (n x cols) = src[:, :cols] * scale, src with row stride ld
hipError_t launch_take_cols(const float *src, uint64_t n, uint32_t ld, uint32_t cols, float scale,
                            float *dst, hipStream_t s);

// Small dense FP64 linear algebra (linalg.hip)
// C = op(A) B, row-major FP64 (the K range is split inside each tile's
// workgroup and the partials summed in a fixed order)
hipError_t launch_gemm64(uint32_t M, uint32_t N, uint32_t K, const double *A, uint32_t lda, bool tA,
                         const double *B, uint32_t ldb, double *C, uint32_t ldc, hipStream_t s);
// G (b x b, b <= 128) = R^T R -> R^{-1}; status[0] = 1 when factoring fails
hipError_t launch_chol_inv(const double *G, uint32_t b, double *Rinv, uint32_t *status, hipStream_t s);
// q leading eigenvectors of symmetric A (m x m, m <= jacobi_max_size()) as
// columns of Vout (m x q)
// (work: jacobi_work_doubles(m) doubles of scratch)
uint32_t jacobi_max_size();
uint64_t jacobi_work_doubles(uint32_t m);
hipError_t launch_jacobi(const double *A, uint32_t m, uint32_t q, double *Vout, double *work, hipStream_t s);
// column signs of V (d x q), its FP32 copy Vf, and the FP32 copy of the mean
hipError_t launch_sign_fix(double *V, uint32_t d, uint32_t q, float *Vf, const double *mean, float *meanf,
                           hipStream_t s);

}  // namespace pm
