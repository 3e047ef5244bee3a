/*
 * pacmap_oracle.h -- CPU restatement of the PaCMAP hot path (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check and bench.py's cpu_baseline leg load it.  Every function restates one
 * function of beamform/pacmap-rs 0.2.6 in plain C with the reference's f32
 * operation order (no FMA contraction: build with -ffp-contract=off), citing
 * the reference file:line it follows.  FP64 twins exist where a result is
 * compared against an order-independent answer.
 *
 * Pinning status (what fixes each function besides its own formula):
 *   orc_find_weights        weights.rs:88-114 goldens (4 iterations)
 *   orc_grad / orc_grad_f64 gradient.rs:161-216 golden (10 points) + zero-sum
 *                           invariant + FP64 twin
 *   orc_adam / orc_powi     adam.rs:68-117 golden (t = 1); powi for t > 1 follows
 *                           compiler-rt's square-and-multiply (parity unpinned
 *                           beyond t = 1 except through the FP64 closed form)
 *   orc_euclid              distance.rs:230-268 properties + FP64 brute force
 *   orc_scale_dist          distance.rs:120-190 goldens (2x2 exact, 10x5 1e-6)
 *   orc_sigma               closed form on hand-made rows (mean of cols 3..5)
 *   orc_knn_triple_sort     knn.rs:215-324 KATs + FP64 brute force on tie-free data
 *   orc_knn_rowwise         equal to orc_knn_triple_sort on random inputs
 *   orc_decide_num_pairs    closed form (lib.rs:565-597) at hand-computed n
 *   orc_rng_next_u64        the xoshiro256++ reference-implementation vector for
 *                           the state {1, 2, 3, 4} (pins the generator and from_seed)
 *   orc_rng_seed_from_u64,  invariants of sampling.rs:323-538 only: the seed
 *   orc_sample_*            expansion and the stream values are PARITY UNPINNED
 *   orc_standard_normal,    closed forms of the ziggurat tables (layer areas, the
 *   orc_random_init         top layer meeting f(0) = 1) and the N(0, 1) law of a
 *                           large sample: PARITY UNPINNED (no reference vectors)
 *   orc_value_init          closed form (mean 0, std 1e-4) only
 *   orc_optimize            composition of the above
 */
#ifndef PACMAP_ORACLE_H
#define PACMAP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* weights.rs:52-81 */
void orc_find_weights(float w_mn_init, uint64_t itr, uint64_t p1, uint64_t p2,
                      float out[3] /* w_nn, w_mn, w_fp */);

/* gradient.rs:35-151.  Pair arrays are rows x 2 u32, row-major.  Output
 * (n+1) x dims, loss at [n][0].  Chunks of 131072 pairs are reduced in
 * array order (NN chunks, then MN, then FP), one valid order of the
 * reference's rayon reduce.  Returns 0, or -1 on an out-of-range index. */
int orc_grad(const float *y, uint64_t n, uint64_t dims,
             const uint32_t *nn, uint64_t n_nn, const uint32_t *mn, uint64_t n_mn,
             const uint32_t *fp, uint64_t n_fp, float w_nn, float w_mn, float w_fp,
             float *grad_out, int nthreads);
/* FP64 twin of orc_grad (same formulas in double, sequential order). */
int orc_grad_f64(const double *y, uint64_t n, uint64_t dims,
                 const uint32_t *nn, uint64_t n_nn, const uint32_t *mn, uint64_t n_mn,
                 const uint32_t *fp, uint64_t n_fp, double w_nn, double w_mn, double w_fp,
                 double *grad_out);

/* powi as lowered by LLVM for a runtime exponent (compiler-rt __powisf2:
 * binary square-and-multiply), used by adam.rs:47. */
float orc_powi(float a, int32_t b);
/* adam.rs:46-47 bias-corrected learning rate for 0-based itr. */
float orc_adam_lr_t(float lr, float beta1, float beta2, uint64_t itr);
/* adam.rs:35-60 */
void orc_adam(float *y, const float *grad, float *m, float *v, uint64_t n, uint64_t dims,
              float beta1, float beta2, float lr, uint64_t itr);

/* distance.rs:50-75 (8 strided lane partials, lanes summed 0..7, tail, sqrt) */
float orc_euclid(const float *a, const float *b, uint64_t d);
double orc_euclid_f64(const float *a, const float *b, uint64_t d);
/* distance.rs:27-37 */
void orc_scale_dist(const float *knn_dist, const float *sig, const uint32_t *nbrs,
                    uint64_t n, uint64_t k, float *out);
/* neighbors.rs:77-84: sigma_i = max(mean(dist[i, start..end]), 1e-10) */
void orc_sigma(const float *knn_dist, uint64_t n, uint64_t k, float *sig);

/* knn.rs:137-194, ties broken by (i, j) of the triple, i.e. by neighbour index
 * within a row.  k is clamped to n-1; outputs n x min(k, n-1).
 * Uses n(n-1)/2 12-byte triples: only for small n. */
int orc_knn_triple_sort(const float *x, uint64_t n, uint64_t d, uint64_t k,
                        uint32_t *idx, float *dist, int nthreads);
/* Same result, per-row selection, O(n*k) memory. */
int orc_knn_rowwise(const float *x, uint64_t n, uint64_t d, uint64_t k,
                    uint32_t *idx, float *dist, int nthreads);
/* The rows of orc_knn_rowwise for the nrows query rows listed in `rows`
 * (output row t is query rows[t]); -1 on an index >= n. */
int orc_knn_rows(const float *x, uint64_t n, uint64_t d, uint64_t k, const uint32_t *rows,
                 uint64_t nrows, uint32_t *idx, float *dist, int nthreads);

/* lib.rs:565-597.  Returns 0, 5 (InvalidNeighborCount), 6 (InvalidFarPointCount). */
int orc_decide_num_pairs(uint64_t n, int has_override, uint64_t override_k,
                         float mn_ratio, float fp_ratio, uint64_t out[3]);

/* rand 0.8.5 SmallRng (Xoshiro256++ on 64-bit) restatement -- PARITY UNPINNED. */
typedef struct { uint64_t s[4]; } orc_rng;
void orc_rng_seed_from_u64(orc_rng *r, uint64_t state);
/* SeedableRng::from_seed: four little-endian u64 words (all zero: SplitMix64(0)) */
void orc_rng_seed_from_bytes(orc_rng *r, const uint8_t seed[32]);
uint64_t orc_rng_next_u64(orc_rng *r);
uint32_t orc_rng_next_u32(orc_rng *r);
uint32_t orc_rng_gen_range_u32(orc_rng *r, uint32_t low, uint32_t high);

/* rand_distr 0.4.3 ziggurat tables (x, f: 257 each) and StandardNormal for
 * f64; lib.rs:769-778 Random(seed) initial Y (n x dims, row-major). */
void orc_zig_tables(double x[257], double f[257]);
double orc_standard_normal(orc_rng *r);
void orc_random_init(uint64_t n, uint64_t dims, uint64_t seed, float *y);

/* sampling.rs:37-61 */
uint64_t orc_sample_fp(uint64_t n_samples, uint32_t maximum, const uint32_t *reject,
                       uint64_t n_reject, uint64_t reject_stride, uint32_t self_ind,
                       orc_rng *rng, uint32_t *out);
/* sampling.rs:77-103 */
void orc_sample_fp_pair_det(uint64_t n, const uint32_t *pair_neighbors, uint64_t n_neighbors,
                            uint64_t n_fp, uint64_t seed, uint32_t *pair_fp /* (n*n_fp) x 2 */);
/* sampling.rs:118-142 + 296-313 */
void orc_sample_mn_pair_det(const float *x, uint64_t n, uint64_t d, uint64_t n_mn,
                            uint64_t seed, uint32_t *pair_mn /* (n*n_mn) x 2 */);
/* The same samplers for chosen points only: out holds nrows x n_mn (or
 * n_fp) rows x 2, the rows of point rows[t] at block t. */
void orc_sample_mn_rows(const float *x, uint64_t n, uint64_t d, uint64_t n_mn, uint64_t seed,
                        const uint32_t *rows, uint64_t nrows, uint32_t *out);
void orc_sample_fp_rows(uint64_t n, const uint32_t *pair_neighbors, uint64_t n_neighbors,
                        uint64_t n_fp, uint64_t seed, const uint32_t *rows, uint64_t nrows,
                        uint32_t *out);
/* sampling.rs:251-286 (ties by column: stable) */
void orc_sample_neighbors_pair(const float *scaled, const uint32_t *nbrs, uint64_t n,
                               uint64_t kprime, uint64_t n_neighbors, uint32_t *pairs);

/* lib.rs:753-766 with ndarray 0.16 mean_axis / var_axis orders. */
void orc_value_init(const float *y_in, uint64_t n, uint64_t dims, float *y_out);
/* lib.rs:486-511 (non-PCA preprocessing).  Returns 0 or 7 (EmptyArrayMinMax). */
int orc_normalize(const float *x, uint64_t n, uint64_t d, float *x_out);

/* lib.rs:781-837: the optimiser loop from a given Y (already initialised).
 * y is updated in place.  snap (may be NULL) gets ns x n x dims states with
 * the reference's slot rules (slot 0 pre-filled with the initial Y). */
int orc_optimize(float *y, uint64_t n, uint64_t dims,
                 const uint32_t *nn, uint64_t n_nn, const uint32_t *mn, uint64_t n_mn,
                 const uint32_t *fp, uint64_t n_fp, float lr, const uint64_t iters[3],
                 const uint64_t *snap_idx, uint64_t ns, float *snap, float *loss_out,
                 int nthreads);

#ifdef __cplusplus
}
#endif
#endif
