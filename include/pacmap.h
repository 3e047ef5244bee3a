/*
 * pacmap.h -- C ABI of the MI355X-native PaCMAP engine (libpacmap_mi355x.so).
 *
 * Drop-in for the public surface of beamform/pacmap-rs 0.2.6:
 *   pacmap::fit_transform(x, Configuration)            src/lib.rs:294-297
 *   pacmap::knn::find_k_nearest_neighbors(data, k)     src/knn.rs:137
 * plus test seams for the per-iteration pieces the reference keeps private
 * (gradient.rs:35, adam.rs:35, weights.rs:52, neighbors.rs:77-87).
 *
 * Conventions (all entry points):
 *   - row-major, contiguous arrays; plain pointers and element counts;
 *   - host pointers unless the name ends in _dev;
 *   - outputs are caller-allocated;
 *   - every call is blocking, owns its HIP stream and device buffers, and is
 *     safe to call from several host threads at once;
 *   - the return value is a pacmap_code; when `st` is non-NULL it receives the
 *     code plus details (shape mismatch, HIP error, message).
 */
#ifndef PACMAP_H
#define PACMAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One code per PaCMapError variant (lib.rs:870-922), then the new ones. */
typedef enum {
  PACMAP_OK = 0,
  PACMAP_ERR_SAMPLE_SIZE = 1,                 /* lib.rs:875-876  n <= 1 */
  PACMAP_ERR_INVALID_NEAREST_NEIGHBOR_SHAPE = 2, /* lib.rs:878-885 */
  PACMAP_ERR_EMPTY_ARRAY_MEAN = 3,            /* lib.rs:887-889 */
  PACMAP_ERR_NORMAL = 4,                      /* lib.rs:891-893 */
  PACMAP_ERR_INVALID_NEIGHBOR_COUNT = 5,      /* lib.rs:895-897 */
  PACMAP_ERR_INVALID_FAR_POINT_COUNT = 6,     /* lib.rs:899-901 */
  PACMAP_ERR_EMPTY_ARRAY_MIN_MAX = 7,         /* lib.rs:903-905 */
  PACMAP_ERR_ZERO_RANGE = 8,                  /* lib.rs:907-909 (declared, never returned) */
  PACMAP_ERR_PCA = 9,                         /* lib.rs:911-913 */
  PACMAP_ERR_NEIGHBORS = 10,                  /* lib.rs:915-917 */
  PACMAP_ERR_MISSING_TRANSFORM = 11,          /* lib.rs:919-921 */
  /* new: the reference panics (gradient.rs:123) where this returns */
  PACMAP_ERR_INVALID_PAIR_INDEX = 12,
  PACMAP_ERR_INVALID_ARGUMENT = 13,
  PACMAP_ERR_GPU = 14,                        /* HIP runtime error, see gpu_error */
  PACMAP_ERR_UNSUPPORTED = 15                 /* size beyond what a kernel supports */
} pacmap_code;

typedef struct {
  int32_t code;          /* pacmap_code */
  int32_t gpu_error;     /* hipError_t when code == PACMAP_ERR_GPU */
  uint64_t expected[2];  /* InvalidNearestNeighborShape: [n * n_neighbors, 2] */
  uint64_t actual[2];
  char message[256];
} pacmap_status;

/* Initialization (lib.rs:226-239) */
enum { PACMAP_INIT_PCA = 0, PACMAP_INIT_VALUE = 1, PACMAP_INIT_RANDOM = 2 };
/* PairConfiguration (lib.rs:241-267) */
enum { PACMAP_PAIRS_GENERATE = 0, PACMAP_PAIRS_NEIGHBORS_PROVIDED = 1, PACMAP_PAIRS_ALL_PROVIDED = 2 };
/* kNN path (MI355X extension).  FP32, BF16 and DIRECT return the exact FP32
 * neighbour sets of knn.rs:137-194 (see pacmap_knn_exact):
 *   FP32   (default) split-bf16 MFMA candidates, exact FP32 rerank, per-row
 *          certificate, exact rescan of uncertified rows;
 *   BF16   single-bf16 MFMA candidates, same rerank/certificate/rescan;
 *   DIRECT every distance in the reference's order on the vector ALUs;
 *   FP8    FP8 E4M3 candidates with per-row scaling (16x16x128 f8f6f4 MFMA),
 *          exact FP32 rerank of the candidates, no certificate: approximate
 *          neighbours (the E4M3 error bound is too wide to certify).
 * OR-ing PACMAP_KNN_APPROX into FP32 or BF16 also drops the certificate and
 * the rescan (approximate neighbours, the exact top-k of the candidates). */
enum { PACMAP_KNN_FP32 = 0, PACMAP_KNN_BF16 = 1, PACMAP_KNN_FP8 = 2, PACMAP_KNN_DIRECT = 3,
       PACMAP_KNN_APPROX = 0x100 };

/* Configuration (lib.rs:163-224).  Option<T> fields carry a has_ flag; enum
 * payloads are borrowed pointers that must outlive the call. */
typedef struct {
  uint64_t embedding_dimensions;      /* default 2 */
  int32_t init_kind;                  /* default PACMAP_INIT_PCA */
  int32_t init_random_has_seed;       /* Random(Some(seed)) */
  uint64_t init_random_seed;
  const float *init_value;            /* Value: n x embedding_dimensions */
  uint64_t init_value_rows, init_value_cols;
  float mid_near_ratio;               /* default 0.5 */
  float far_pair_ratio;               /* default 2.0 */
  int32_t has_override_neighbors;
  uint64_t override_neighbors;
  int32_t has_seed;
  uint64_t seed;
  int32_t pair_kind;                  /* default PACMAP_PAIRS_GENERATE */
  const uint32_t *pair_neighbors;     /* rows x cols (cols must be 2) */
  uint64_t pair_neighbors_rows, pair_neighbors_cols;
  const uint32_t *pair_mn;            /* AllProvided: rows x 2 */
  uint64_t pair_mn_rows;
  const uint32_t *pair_fp;            /* AllProvided: rows x 2 */
  uint64_t pair_fp_rows;
  float learning_rate;                /* default 1.0 */
  uint64_t num_iters[3];              /* default (100, 100, 250) */
  int32_t has_snapshots;
  const uint64_t *snapshots;          /* iteration indices */
  uint64_t num_snapshots;
  uint64_t approx_threshold;          /* default 8000 */
  /* MI355X extensions */
  int32_t knn_precision;              /* PACMAP_KNN_*, default FP32 */
  int32_t record_loss;                /* 1: fill loss_out when it is non-NULL (default 1) */
} pacmap_config;

/* Fills `cfg` with Configuration::default() (lib.rs:208-224). */
void pacmap_config_default(pacmap_config *cfg);

/* fit_transform (lib.rs:294-367).  y_out: n x dims.  snapshots_out: when the
 * config has snapshots, num_snapshots x n x dims (else may be NULL).
 * loss_out (may be NULL): total-iteration float array of losses.  With NULL (or
 * record_loss 0) the optimiser skips the loss terms; the embedding is the same. */
int pacmap_fit_transform(const float *x, uint64_t n, uint64_t d, const pacmap_config *cfg,
                         float *y_out, float *snapshots_out, float *loss_out,
                         pacmap_status *st);

/* Same, with x already resident in device memory (HBM); outputs on the host. */
int pacmap_fit_transform_dev(const float *x_dev, uint64_t n, uint64_t d,
                             const pacmap_config *cfg, float *y_out, float *snapshots_out,
                             float *loss_out, pacmap_status *st);

/* find_k_nearest_neighbors (knn.rs:137-194): per row the min(k, n-1) nearest
 * other rows by the distance of distance.rs:50-75, ascending, ties by index.
 * idx: n x min(k, n-1) u32, dist: n x min(k, n-1) f32.  n == 0 is a no-op. */
int pacmap_knn_exact(const float *x, uint64_t n, uint64_t d, uint64_t k, uint32_t *idx,
                     float *dist, pacmap_status *st);
/* Same with a candidate precision (PACMAP_KNN_*); results equal pacmap_knn_exact
 * except in the approximate modes (FP8, PACMAP_KNN_APPROX).
 * stats_out (may be NULL): [rows re-scanned exactly, candidates per row]. */
int pacmap_knn(const float *x, uint64_t n, uint64_t d, uint64_t k, int32_t precision,
               uint32_t *idx, float *dist, uint64_t *stats_out, pacmap_status *st);

/* Seam: pacmap_grad (gradient.rs:35-74).  grad_out: (n+1) x dims, loss at
 * [n][0].  Pair arrays rows x 2. */
int pacmap_grad(const float *y, uint64_t n, uint64_t dims, const uint32_t *pair_nn,
                uint64_t n_nn, const uint32_t *pair_mn, uint64_t n_mn, const uint32_t *pair_fp,
                uint64_t n_fp, float w_nn, float w_mn, float w_fp, float *grad_out,
                pacmap_status *st);

/* Seam: update_embedding_adam (adam.rs:35-60), in place on y, m, v (n x dims);
 * grad has at least n rows. */
int pacmap_adam_step(float *y, const float *grad, float *m, float *v, uint64_t n, uint64_t dims,
                     float beta1, float beta2, float lr, uint64_t itr, pacmap_status *st);

/* Seam: find_weights (weights.rs:52-81); out = {w_nn, w_mn, w_fp}.  Host only. */
void pacmap_weights(float w_mn_init, uint64_t itr, uint64_t phase_1_iters,
                    uint64_t phase_2_iters, float out[3]);

/* Seam: sigma and scale_dist (neighbors.rs:77-87, distance.rs:27-37) plus the
 * NN pair pick of sample_neighbors_pair (sampling.rs:251-286, ties by column).
 * knn_dist / knn_idx: n x kprime.  pairs_out (may be NULL): (n*n_neighbors) x 2. */
int pacmap_sigma_scale(const float *knn_dist, const uint32_t *knn_idx, uint64_t n,
                       uint64_t kprime, uint64_t n_neighbors, float *sigma_out,
                       float *scaled_out, uint32_t *pairs_out, pacmap_status *st);

/* Seam: scale_dist (distance.rs:27-37) with sigma given: scaled[i][c] =
 * knn_dist[i][c]^2 / (sig[i] * sig[nbrs[i][c]]).  knn_dist, nbrs, scaled_out:
 * n x k; sig: n. */
int pacmap_scale_dist(const float *knn_dist, const float *sig, const uint32_t *nbrs, uint64_t n,
                      uint64_t k, float *scaled_out, pacmap_status *st);

/* Seam: mid-near and far pair sampling of generate_pairs with a seed
 * (sampling.rs:77-142, 296-313; neighbors.rs:92-96) on the GPU.  x: n x d
 * (the preprocessed data), pair_nn: (n*n_neighbors) x 2 (i-major NN pairs,
 * the far-pair reject lists).  pair_mn_out: (n*n_mn) x 2, pair_fp_out:
 * (n*n_fp) x 2 (unfilled far pairs stay (0, 0)). */
int pacmap_sample_pairs(const float *x, uint64_t n, uint64_t d, const uint32_t *pair_nn,
                        uint64_t n_neighbors, uint64_t n_mn, uint64_t n_fp, uint64_t seed,
                        uint32_t *pair_mn_out, uint32_t *pair_fp_out, pacmap_status *st);

/* Seam: the optimiser loop of pacmap() (lib.rs:781-837) from an already
 * initialised Y (no Value renormalisation), as one hipGraph.  y: n x dims in/out. */
int pacmap_optimize(float *y, uint64_t n, uint64_t dims, const uint32_t *pair_nn,
                    uint64_t n_nn, const uint32_t *pair_mn, uint64_t n_mn,
                    const uint32_t *pair_fp, uint64_t n_fp, float lr,
                    const uint64_t num_iters[3], const uint64_t *snapshots,
                    uint64_t num_snapshots, float *snapshots_out, float *loss_out,
                    pacmap_status *st);

/* The pairs of PairConfiguration::Generate (lib.rs:629-650 with
 * neighbors.rs:56-108): preprocess x as fit_transform does (lib.rs:442-532),
 * then exact kNN, sigma/scale, NN pairs and seeded MN/FP sampling.  With
 * pair_nn_out NULL only counts_out is filled: NN, MN and FP row counts and
 * the preprocessed width d'.  Otherwise pair_*_out receive rows x 2 u32
 * (i-major) and x_pre_out (optional) the preprocessed data n x d'. */
int pacmap_generate_pairs(const float *x, uint64_t n, uint64_t d, const pacmap_config *cfg,
                          uint32_t *pair_nn_out, uint32_t *pair_mn_out, uint32_t *pair_fp_out,
                          float *x_pre_out, uint64_t counts_out[4], pacmap_status *st);

/* ---- Parametric encoder (no counterpart in the reference) ----------------
 * An MLP d -> hidden -> hidden -> dims with GELU, trained with the PaCMAP
 * NN/MN/FP loss (gradient.rs:98-151 terms) on the method's own pairs, for
 * out-of-sample transform and warm-start initialisation.  BF16 weights and
 * activations on MFMA with FP32 accumulation; FP32 master weights and Adam.
 * Parameters (FP32, row-major): W1 (hidden x d), b1, W2 (hidden x hidden), b2,
 * W3 (dims x hidden), b3. */
typedef struct {
  uint64_t hidden;            /* 256 (multiple of 64, at most 1024) */
  uint64_t epochs;            /* 20 */
  uint64_t batch_pairs;       /* 8192 pairs per step (16384 encoder rows) */
  float learning_rate;        /* Adam, 1e-3 */
  uint64_t seed;              /* pair order of each epoch */
  uint64_t schedule_iters[3]; /* the weight phases of weights.rs spread over training */
} pacmap_mlp_config;
void pacmap_mlp_config_default(pacmap_mlp_config *cfg);
uint64_t pacmap_mlp_param_count(uint64_t d, uint64_t hidden, uint64_t dims);
/* U(-1/sqrt(fan_in), 1/sqrt(fan_in)) initial parameters from a seed. */
int pacmap_mlp_init(uint64_t d, uint64_t hidden, uint64_t dims, uint64_t seed, float *params,
                    pacmap_status *st);
/* Trains params (in: initial, out: trained) on x (n x d) and the three pair
 * lists; epoch_loss (epochs floats, optional) gets the mean pair loss. */
int pacmap_mlp_fit(const float *x, uint64_t n, uint64_t d, uint64_t dims, const uint32_t *pair_nn,
                   uint64_t n_nn, const uint32_t *pair_mn, uint64_t n_mn, const uint32_t *pair_fp,
                   uint64_t n_fp, const pacmap_mlp_config *cfg, float *params, float *epoch_loss,
                   pacmap_status *st);
/* y_out (n x dims) = encoder(x); x on the host or (``_dev``) in device memory.
 * The converted weights of a thread's last call are kept and reused while
 * params stay byte-equal (pacmap_release_cached_memory frees them). */
int pacmap_mlp_transform(const float *x, uint64_t n, uint64_t d, uint64_t hidden, uint64_t dims,
                         const float *params, float *y_out, pacmap_status *st);
int pacmap_mlp_transform_dev(const float *x_dev, uint64_t n, uint64_t d, uint64_t hidden,
                             uint64_t dims, const float *params, float *y_out, pacmap_status *st);
/* Allocation-free pacmap_mlp_transform_dev.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (serving loops that
 * embed new points with a trained encoder, harnesses that forbid allocation).
 * x_dev (n x d), params_dev (pacmap_mlp_param_count floats) and y_dev
 * (n x dims) are all device pointers; the call makes no hipMalloc/hipFree.
 * The bf16 weight copies go to a static per-device workspace of 1,179,648
 * bytes (held by each GPU the library runs on until the process ends), so
 * hidden must be 256, d at most 2048 and dims 1..8 (else
 * PACMAP_ERR_UNSUPPORTED).  Runs on the calling thread's default stream and
 * returns when done.  Calls that overlap on one GPU (other host threads)
 * share the workspace and race: serialise them. */
int pacmap_mlp_transform_noalloc(const float *x_dev, uint64_t n, uint64_t d, uint64_t hidden,
                                 uint64_t dims, const float *params_dev, float *y_dev,
                                 pacmap_status *st);
/* Seam: mean loss and its gradient over one batch (all given pairs, at most
 * 65536) with weights (w_nn, w_mn, w_fp); grad_out has the parameter layout. */
int pacmap_mlp_grad(const float *x, uint64_t n, uint64_t d, uint64_t hidden, uint64_t dims,
                    const float *params, const uint32_t *pair_nn, uint64_t n_nn,
                    const uint32_t *pair_mn, uint64_t n_mn, const uint32_t *pair_fp, uint64_t n_fp,
                    const float weights[3], float *grad_out, float *loss_out, pacmap_status *st);

/* Device buffers for callers that keep inputs resident in HBM. */
int pacmap_dev_alloc(uint64_t bytes, void **ptr_out, pacmap_status *st);
int pacmap_dev_free(void *ptr);
int pacmap_dev_upload(void *dst_dev, const void *src_host, uint64_t bytes, pacmap_status *st);
int pacmap_dev_download(void *dst_host, const void *src_dev, uint64_t bytes, pacmap_status *st);
/* Device blocks the library allocated are cached per calling thread and
 * reused by its later calls (no hipMalloc / hipFree per call), and so is the
 * thread's last optimiser graph (replayed when a call's launches are
 * identical).  Uploads of 8 MiB or more go through two 32 MiB pinned host
 * buffers per calling thread, also kept between calls.  A thread's stream,
 * graph, pinned buffers and cached blocks are released when the thread
 * exits; an allocation that runs out of device memory first frees the idle
 * cached blocks of every thread and retries.  This frees the idle cached
 * blocks of every thread, and the calling thread's graph, pinned upload
 * buffers and cached encoder transform weights. */
void pacmap_release_cached_memory(void);
/* HBM reference: a plain streaming device-to-device copy kernel run `repeats`
 * times; ms_out gets the average time per copy (HIP events).  Its rocprof
 * FETCH_SIZE / WRITE_SIZE readings calibrate the traffic counters. */
int pacmap_dev_copy(void *dst_dev, const void *src_dev, uint64_t bytes, uint32_t repeats, double *ms_out,
                    pacmap_status *st);

/* Per-thread timings of the last call on this thread (HIP events on the call's
 * stream, milliseconds; host phases by steady clock). */
typedef struct {
  double total_ms;
  double upload_ms;
  double preprocess_ms;       /* normalise or PCA (covariance, eigensolve, projection) */
  double pca_cov_ms;          /* covariance kernel alone */
  double pca_eig_ms;          /* eigen-step kernels (subspace iteration, Jacobi) */
  double knn_ms;              /* kNN kernels */
  double knn_kernel_ms;       /* the dominant kNN kernel alone */
  uint64_t knn_kernel_launches;
  double sigma_ms;            /* sigma/scale/NN-pair kernels */
  double sampling_ms;         /* MN/FP sampling kernels */
  double csr_ms;              /* CSR build (degrees, scan, fill, row sort) */
  double graph_build_ms;      /* capture + instantiate */
  double optimize_ms;         /* the graph of all iterations */
  uint64_t iterations;
  uint64_t csr_entries;       /* 2 x non-self pairs */
  uint64_t knn_rescanned_rows; /* rows the certificate sent to the exact rescan */
  uint64_t knn_candidates;     /* candidate list capacity per row */
  double knn_rerank_ms;        /* exact rerank + certificate kernel */
  double knn_fallback_ms;      /* exact rescan of uncertified rows */
  uint64_t knn_path;           /* 0 direct, 1 split-bf16 MFMA, 2 bf16 MFMA, 3 FP8 MFMA */
  double knn_sample_ms;        /* MFMA threshold-sampling kernel */
  uint64_t knn_certified;      /* 1: exact neighbours (certificate + rescan), 0: approximate */
  uint64_t grad_target_blocks; /* 0: one-pass gradient; 8: target-blocked two-pass gradient */
  uint64_t grad_sell;          /* 1: LDS-staged target-blocked gradient (sliced ELL) */
  uint64_t grad_sell_entries;  /* sliced-ELL words including padding */
  uint64_t grad_sell_ranges;   /* target ranges of the sliced-ELL gradient (4 or 8) */
  uint64_t graph_reused;       /* 1: the previous call's optimiser graph was replayed */
  uint64_t knn_listed;         /* candidates listed, summed over the reranked rows */
  uint64_t knn_reranked;       /* of those, candidates given an exact distance */
  uint64_t grad_paged;         /* 1: paged target-blocked gradient (Y pages staged through LDS) */
  uint64_t grad_page_words;    /* its entry words including padding */
  uint64_t grad_page_pages;    /* its pages per target range */
  uint64_t knn_tile_pairs;     /* symmetric kNN pass: (query group, candidate tile) pairs past the group's own */
  uint64_t knn_tile_pairs_skipped; /* of those, skipped by their bounding boxes */
} pacmap_timings;
void pacmap_last_timings(pacmap_timings *out);
/* Seam: the rows (ascending) that the last kNN on this thread (pacmap_knn or
 * the kNN inside fit_transform / generate_pairs) recomputed by the exact
 * rescan because the certificate did not cover them.  Copies at most cap
 * rows; returns their count. */
uint64_t pacmap_last_rescanned_rows(uint32_t *rows_out, uint64_t cap);

/* Library and device information; returns 0 when a device is usable. */
int pacmap_device_info(char *name_out, uint64_t name_len, int32_t *cu_count_out);
const char *pacmap_version(void);

#ifdef __cplusplus
}
#endif
#endif
