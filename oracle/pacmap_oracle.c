/*
 * pacmap_oracle.c -- CPU restatement of the PaCMAP hot path (TEST INFRASTRUCTURE).
 *
 * The checker, never the product: only tests/, the smoke check and bench.py's
 * cpu_baseline leg load this library.  Each function follows one function of
 * beamform/pacmap-rs 0.2.6 (cited file:line) with the reference's f32
 * operation order.  Build with -ffp-contract=off so no multiply-add is fused
 * (Rust never contracts).  The pinning status of every function is listed in
 * pacmap_oracle.h; rand 0.8.5 stream values are parity unpinned.
 */
#include "pacmap_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small pthread parallel-for                                                */
/* ------------------------------------------------------------------------ */
typedef void (*orc_task_fn)(void *ctx, uint64_t begin, uint64_t end);
typedef struct {
  orc_task_fn fn;
  void *ctx;
  uint64_t begin, end;
} orc_task;

static void *orc_task_run(void *p) {
  orc_task *t = (orc_task *)p;
  t->fn(t->ctx, t->begin, t->end);
  return NULL;
}

/* Splits [0, total) into nthreads contiguous ranges. */
static void orc_parallel_for(uint64_t total, int nthreads, orc_task_fn fn, void *ctx) {
  if (nthreads <= 1 || total < 2) {
    fn(ctx, 0, total);
    return;
  }
  if ((uint64_t)nthreads > total) nthreads = (int)total;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  orc_task *tk = (orc_task *)malloc(sizeof(orc_task) * (size_t)nthreads);
  for (int t = 0; t < nthreads; t++) {
    tk[t].fn = fn;
    tk[t].ctx = ctx;
    tk[t].begin = total * (uint64_t)t / (uint64_t)nthreads;
    tk[t].end = total * (uint64_t)(t + 1) / (uint64_t)nthreads;
    pthread_create(&th[t], NULL, orc_task_run, &tk[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(tk);
}

/* Monotone map of f32 bits onto u32 so that unsigned order equals Rust's
 * f32::total_cmp order (used by knn.rs:167, sampling.rs:265,308). */
static inline uint32_t orc_total_key(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

/* ------------------------------------------------------------------------ */
/* weights.rs:52-81                                                          */
/* ------------------------------------------------------------------------ */
void orc_find_weights(float w_mn_init, uint64_t itr, uint64_t p1, uint64_t p2, float out[3]) {
  if (itr < p1) {
    /* weights.rs:58-63: progress in f32, (1-p)*init + p*3 */
    float progress = (float)itr / (float)p1;
    float a = (1.0f - progress) * w_mn_init;
    float b = progress * 3.0f;
    out[0] = 2.0f;
    out[1] = a + b;
    out[2] = 1.0f;
  } else if (itr < p1 + p2) {
    out[0] = 3.0f; /* weights.rs:65-70 */
    out[1] = 3.0f;
    out[2] = 1.0f;
  } else {
    out[0] = 1.0f; /* weights.rs:73-78 */
    out[1] = 0.0f;
    out[2] = 1.0f;
  }
}

/* ------------------------------------------------------------------------ */
/* gradient.rs:35-151                                                        */
/* ------------------------------------------------------------------------ */
#define ORC_CHUNK 131072u /* gradient.rs:57: 128 * 1024 pairs */
#define ORC_MAX_DIMS 64

typedef struct {
  const float *y;
  uint64_t n, dims;
  const uint32_t *pairs[3];
  uint64_t rows[3];
  float w[3];
  uint64_t nchunks[3];
  uint64_t total_chunks;
  float *chunk_grad; /* total_chunks x (n+1) x dims */
  float *chunk_loss;
} orc_grad_ctx;

/* gradient.rs:98-151: one chunk, sequential. */
static void orc_process_pairs(const float *y, uint64_t dims, const uint32_t *pairs, uint64_t rows,
                              float w, float denom_const, float w_const, int is_fp, float *grad,
                              float *loss_out) {
  float loss = 0.0f;
  float yij[ORC_MAX_DIMS];
  for (uint64_t r = 0; r < rows; r++) {
    uint64_t i = pairs[2 * r], j = pairs[2 * r + 1];
    if (i == j) continue; /* gradient.rs:116-118 */
    float d_ij = 1.0f;    /* gradient.rs:121-125 */
    for (uint64_t d = 0; d < dims; d++) {
      yij[d] = y[i * dims + d] - y[j * dims + d];
      float sq = yij[d] * yij[d];
      d_ij += sq;
    }
    if (is_fp) {
      /* gradient.rs:129-136 */
      float opd = 1.0f + d_ij;
      loss += w * (1.0f / opd);
      float w1 = w * (2.0f / (opd * opd));
      for (uint64_t d = 0; d < dims; d++) {
        float g = w1 * yij[d];
        grad[i * dims + d] -= g;
        grad[j * dims + d] += g;
      }
    } else {
      /* gradient.rs:139-146 */
      float cd = denom_const + d_ij;
      loss += w * (d_ij / cd);
      float w1 = w * (w_const / (cd * cd));
      for (uint64_t d = 0; d < dims; d++) {
        float g = w1 * yij[d];
        grad[i * dims + d] += g;
        grad[j * dims + d] -= g;
      }
    }
  }
  *loss_out = loss;
}

static void orc_grad_chunk_task(void *vctx, uint64_t begin, uint64_t end) {
  orc_grad_ctx *c = (orc_grad_ctx *)vctx;
  static const float denom[3] = {10.0f, 10000.0f, 1.0f};
  static const float wconst[3] = {20.0f, 20000.0f, 2.0f};
  uint64_t stride = (c->n + 1) * c->dims;
  for (uint64_t ch = begin; ch < end; ch++) {
    uint64_t t = 0, local = ch;
    while (local >= c->nchunks[t]) {
      local -= c->nchunks[t];
      t++;
    }
    uint64_t r0 = local * ORC_CHUNK;
    uint64_t rows = c->rows[t] - r0;
    if (rows > ORC_CHUNK) rows = ORC_CHUNK;
    float *g = c->chunk_grad + ch * stride;
    memset(g, 0, sizeof(float) * stride);
    orc_process_pairs(c->y, c->dims, c->pairs[t] + 2 * r0, rows, c->w[t], denom[t], wconst[t],
                      t == 2, g, &c->chunk_loss[ch]);
  }
}

static int orc_check_pairs(const uint32_t *p, uint64_t rows, uint64_t n) {
  for (uint64_t r = 0; r < 2 * rows; r++)
    if (p[r] >= n) return -1;
  return 0;
}

int orc_grad(const float *y, uint64_t n, uint64_t dims, const uint32_t *nn, uint64_t n_nn,
             const uint32_t *mn, uint64_t n_mn, const uint32_t *fp, uint64_t n_fp, float w_nn,
             float w_mn, float w_fp, float *grad_out, int nthreads) {
  if (dims > ORC_MAX_DIMS) return -2;
  if (orc_check_pairs(nn, n_nn, n) || orc_check_pairs(mn, n_mn, n) || orc_check_pairs(fp, n_fp, n))
    return -1;
  orc_grad_ctx c;
  memset(&c, 0, sizeof(c));
  c.y = y;
  c.n = n;
  c.dims = dims;
  c.pairs[0] = nn;
  c.pairs[1] = mn;
  c.pairs[2] = fp;
  c.rows[0] = n_nn;
  c.rows[1] = n_mn;
  c.rows[2] = n_fp;
  c.w[0] = w_nn;
  c.w[1] = w_mn;
  c.w[2] = w_fp;
  for (int t = 0; t < 3; t++) {
    c.nchunks[t] = (c.rows[t] + ORC_CHUNK - 1) / ORC_CHUNK;
    c.total_chunks += c.nchunks[t];
  }
  uint64_t stride = (n + 1) * dims;
  /* gradient.rs:65: reduce identity is a zero (n+1) x dims array and loss 0 */
  memset(grad_out, 0, sizeof(float) * stride);
  float total_loss = 0.0f;
  if (c.total_chunks > 0) {
    c.chunk_grad = (float *)malloc(sizeof(float) * stride * c.total_chunks);
    c.chunk_loss = (float *)malloc(sizeof(float) * c.total_chunks);
    orc_parallel_for(c.total_chunks, nthreads, orc_grad_chunk_task, &c);
    /* gradient.rs:66-69 in chunk order */
    for (uint64_t ch = 0; ch < c.total_chunks; ch++) {
      const float *g = c.chunk_grad + ch * stride;
      for (uint64_t e = 0; e < stride; e++) grad_out[e] += g[e];
      total_loss = total_loss + c.chunk_loss[ch];
    }
    free(c.chunk_grad);
    free(c.chunk_loss);
  }
  grad_out[n * dims] = total_loss; /* gradient.rs:72 */
  return 0;
}

int orc_grad_f64(const double *y, uint64_t n, uint64_t dims, const uint32_t *nn, uint64_t n_nn,
                 const uint32_t *mn, uint64_t n_mn, const uint32_t *fp, uint64_t n_fp,
                 double w_nn, double w_mn, double w_fp, double *grad_out) {
  if (dims > ORC_MAX_DIMS) return -2;
  if (orc_check_pairs(nn, n_nn, n) || orc_check_pairs(mn, n_mn, n) || orc_check_pairs(fp, n_fp, n))
    return -1;
  const uint32_t *P[3] = {nn, mn, fp};
  uint64_t R[3] = {n_nn, n_mn, n_fp};
  double W[3] = {w_nn, w_mn, w_fp};
  const double C[3] = {10.0, 10000.0, 1.0};
  const double C2[3] = {20.0, 20000.0, 2.0};
  memset(grad_out, 0, sizeof(double) * (n + 1) * dims);
  double loss = 0.0, u[ORC_MAX_DIMS];
  for (int t = 0; t < 3; t++) {
    for (uint64_t r = 0; r < R[t]; r++) {
      uint64_t i = P[t][2 * r], j = P[t][2 * r + 1];
      if (i == j) continue;
      double D = 1.0;
      for (uint64_t d = 0; d < dims; d++) {
        u[d] = y[i * dims + d] - y[j * dims + d];
        D += u[d] * u[d];
      }
      double w1, s;
      if (t == 2) {
        loss += W[t] / (1.0 + D);
        w1 = W[t] * 2.0 / ((1.0 + D) * (1.0 + D));
        s = -1.0;
      } else {
        loss += W[t] * D / (C[t] + D);
        w1 = W[t] * C2[t] / ((C[t] + D) * (C[t] + D));
        s = 1.0;
      }
      for (uint64_t d = 0; d < dims; d++) {
        grad_out[i * dims + d] += s * w1 * u[d];
        grad_out[j * dims + d] -= s * w1 * u[d];
      }
    }
  }
  grad_out[n * dims] = loss;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* adam.rs:35-60                                                             */
/* ------------------------------------------------------------------------ */
float orc_powi(float a, int32_t b) {
  /* binary square-and-multiply, the lowering of llvm.powi.f32 for a runtime
   * exponent (compiler-rt __powisf2) */
  int recip = b < 0;
  float r = 1.0f;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

float orc_adam_lr_t(float lr, float beta1, float beta2, uint64_t itr) {
  int32_t t = (int32_t)(itr + 1); /* adam.rs:46 */
  float num = sqrtf(1.0f - orc_powi(beta2, t));
  float lr_num = lr * num;
  return lr_num / (1.0f - orc_powi(beta1, t)); /* adam.rs:47 */
}

void orc_adam(float *y, const float *grad, float *m, float *v, uint64_t n, uint64_t dims,
              float beta1, float beta2, float lr, uint64_t itr) {
  float lr_t = orc_adam_lr_t(lr, beta1, beta2, itr);
  float omb1 = 1.0f - beta1, omb2 = 1.0f - beta2;
  for (uint64_t e = 0; e < n * dims; e++) {
    float g = grad[e];
    m[e] = m[e] + omb1 * (g - m[e]);       /* adam.rs:57 */
    v[e] = v[e] + omb2 * (g * g - v[e]);   /* adam.rs:58 */
    float step = (lr_t * m[e]) / (sqrtf(v[e]) + 1e-7f);
    y[e] = y[e] - step;                    /* adam.rs:59 */
  }
}

/* ------------------------------------------------------------------------ */
/* distance.rs                                                               */
/* ------------------------------------------------------------------------ */
float orc_euclid(const float *a, const float *b, uint64_t d) {
  /* distance.rs:53-64: f32x8 accumulation over 8-element chunks */
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t chunks = d / 8;
  for (uint64_t c = 0; c < chunks; c++)
    for (int l = 0; l < 8; l++) {
      float diff = a[8 * c + l] - b[8 * c + l];
      float sq = diff * diff;
      acc[l] = acc[l] + sq;
    }
  /* distance.rs:66: lanes summed left to right */
  float s = 0.0f;
  for (int l = 0; l < 8; l++) s = s + acc[l];
  /* distance.rs:69-72: remainder, sequentially */
  for (uint64_t e = chunks * 8; e < d; e++) {
    float diff = a[e] - b[e];
    float sq = diff * diff;
    s = s + sq;
  }
  return sqrtf(s); /* distance.rs:74 */
}

double orc_euclid_f64(const float *a, const float *b, uint64_t d) {
  double s = 0.0;
  for (uint64_t e = 0; e < d; e++) {
    double diff = (double)a[e] - (double)b[e];
    s += diff * diff;
  }
  return sqrt(s);
}

void orc_scale_dist(const float *knn_dist, const float *sig, const uint32_t *nbrs, uint64_t n,
                    uint64_t k, float *out) {
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t c = 0; c < k; c++) {
      float dd = knn_dist[i * k + c] * knn_dist[i * k + c]; /* distance.rs:35 */
      float ss = sig[i] * sig[nbrs[i * k + c]];
      out[i * k + c] = dd / ss;
    }
}

void orc_sigma(const float *knn_dist, uint64_t n, uint64_t k, float *sig) {
  /* neighbors.rs:79-80 */
  uint64_t start = (k >= 1 ? k - 1 : 0);
  if (start > 3) start = 3;
  uint64_t end = k < 6 ? k : 6;
  for (uint64_t i = 0; i < n; i++) {
    if (end <= start) { /* mean_axis of an empty lane: None -> 1e-10 (neighbors.rs:84) */
      sig[i] = 1e-10f;
      continue;
    }
    float s = 0.0f; /* ndarray unrolled_fold over < 8 elements: sequential from 0 */
    for (uint64_t c = start; c < end; c++) s = s + knn_dist[i * k + c];
    float mean = s / (float)(end - start);
    sig[i] = fmaxf(mean, 1e-10f); /* f32::max ignores NaN */
  }
}

/* ------------------------------------------------------------------------ */
/* knn.rs:137-194                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t key; /* total_key(dist) << 32 | i */
  uint32_t j;
  float dist;
} orc_triple;

static int orc_triple_cmp(const void *pa, const void *pb) {
  const orc_triple *a = (const orc_triple *)pa, *b = (const orc_triple *)pb;
  if (a->key != b->key) return a->key < b->key ? -1 : 1;
  if (a->j != b->j) return a->j < b->j ? -1 : 1;
  return 0;
}

typedef struct {
  const float *x;
  uint64_t n, d;
  orc_triple *tri;
} orc_tri_ctx;

static inline uint64_t orc_tri_offset(uint64_t i, uint64_t n) {
  /* number of triples with first index < i */
  return i * n - i * (i + 1) / 2;
}

static void orc_tri_fill(void *vc, uint64_t b, uint64_t e) {
  orc_tri_ctx *c = (orc_tri_ctx *)vc;
  for (uint64_t i = b; i < e; i++) {
    orc_triple *t = c->tri + orc_tri_offset(i, c->n);
    for (uint64_t j = i + 1; j < c->n; j++, t++) {
      float dist = orc_euclid(c->x + i * c->d, c->x + j * c->d, c->d); /* knn.rs:158 */
      t->key = ((uint64_t)orc_total_key(dist) << 32) | (uint64_t)i;
      t->j = (uint32_t)j;
      t->dist = dist;
    }
  }
}

typedef struct {
  orc_triple *a;
  uint64_t *bounds;
} orc_sort_ctx;

static void orc_sort_seg(void *vc, uint64_t b, uint64_t e) {
  orc_sort_ctx *c = (orc_sort_ctx *)vc;
  for (uint64_t s = b; s < e; s++)
    qsort(c->a + c->bounds[s], c->bounds[s + 1] - c->bounds[s], sizeof(orc_triple), orc_triple_cmp);
}

typedef struct {
  orc_triple *src, *dst;
  uint64_t *bounds;
  uint64_t nseg, width;
} orc_merge_ctx;

static void orc_merge_task(void *vc, uint64_t b, uint64_t e) {
  orc_merge_ctx *c = (orc_merge_ctx *)vc;
  for (uint64_t p = b; p < e; p++) {
    uint64_t s0 = p * 2 * c->width;
    uint64_t s1 = s0 + c->width, s2 = s0 + 2 * c->width;
    if (s1 > c->nseg) s1 = c->nseg;
    if (s2 > c->nseg) s2 = c->nseg;
    uint64_t i = c->bounds[s0], ie = c->bounds[s1], j = c->bounds[s1], je = c->bounds[s2];
    uint64_t o = c->bounds[s0];
    while (i < ie && j < je) {
      if (orc_triple_cmp(&c->src[j], &c->src[i]) < 0) c->dst[o++] = c->src[j++];
      else c->dst[o++] = c->src[i++];
    }
    while (i < ie) c->dst[o++] = c->src[i++];
    while (j < je) c->dst[o++] = c->src[j++];
  }
}

int orc_knn_triple_sort(const float *x, uint64_t n, uint64_t d, uint64_t k, uint32_t *idx,
                        float *dist, int nthreads) {
  if (n == 0) return 0; /* knn.rs:141-143 */
  if (k > n - 1) k = n - 1; /* knn.rs:146 */
  uint64_t ntri = n * (n - 1) / 2;
  orc_triple *tri = (orc_triple *)malloc(sizeof(orc_triple) * (ntri ? ntri : 1));
  if (!tri) return -1;
  orc_tri_ctx fc = {x, n, d, tri};
  orc_parallel_for(n, nthreads, orc_tri_fill, &fc);
  /* knn.rs:167: global sort by distance (parallel segment sort + merges) */
  uint64_t nseg = nthreads > 1 ? (uint64_t)nthreads : 1;
  if (nseg > ntri) nseg = ntri ? ntri : 1;
  uint64_t *bounds = (uint64_t *)malloc(sizeof(uint64_t) * (nseg + 1));
  for (uint64_t s = 0; s <= nseg; s++) bounds[s] = ntri * s / nseg;
  orc_sort_ctx sc = {tri, bounds};
  orc_parallel_for(nseg, nthreads, orc_sort_seg, &sc);
  if (nseg > 1) {
    orc_triple *tmp = (orc_triple *)malloc(sizeof(orc_triple) * ntri);
    orc_triple *src = tri, *dst = tmp;
    for (uint64_t w = 1; w < nseg; w *= 2) {
      orc_merge_ctx mc = {src, dst, bounds, nseg, w};
      uint64_t npairs = (nseg + 2 * w - 1) / (2 * w);
      orc_parallel_for(npairs, nthreads, orc_merge_task, &mc);
      orc_triple *t = src;
      src = dst;
      dst = t;
    }
    if (src != tri) memcpy(tri, src, sizeof(orc_triple) * ntri);
    free(tmp);
  }
  free(bounds);
  /* knn.rs:170-191: greedy fill */
  for (uint64_t e = 0; e < n * k; e++) {
    idx[e] = 0;
    dist[e] = 3.40282347e38f; /* f32::MAX */
  }
  uint64_t *counts = (uint64_t *)calloc(n, sizeof(uint64_t));
  for (uint64_t t = 0; t < ntri; t++) {
    uint64_t i = tri[t].key & 0xffffffffu, j = tri[t].j;
    if (counts[i] < k) {
      idx[i * k + counts[i]] = (uint32_t)j;
      dist[i * k + counts[i]] = tri[t].dist;
      counts[i]++;
    }
    if (counts[j] < k) {
      idx[j * k + counts[j]] = (uint32_t)i;
      dist[j * k + counts[j]] = tri[t].dist;
      counts[j]++;
    }
  }
  free(counts);
  free(tri);
  return 0;
}

typedef struct {
  const float *x;
  uint64_t n, d, k;
  uint32_t *idx;
  float *dist;
  const uint32_t *rows; /* query rows (NULL: every row); output row t is query rows[t] */
} orc_row_ctx;

static void orc_heap_sift_down(uint64_t *h, uint64_t sz, uint64_t p) {
  for (;;) {
    uint64_t l = 2 * p + 1, r = l + 1, m = p;
    if (l < sz && h[l] > h[m]) m = l;
    if (r < sz && h[r] > h[m]) m = r;
    if (m == p) return;
    uint64_t t = h[p];
    h[p] = h[m];
    h[m] = t;
    p = m;
  }
}

static int orc_u64_cmp(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static void orc_row_task(void *vc, uint64_t b, uint64_t e) {
  orc_row_ctx *c = (orc_row_ctx *)vc;
  uint64_t k = c->k;
  uint64_t *heap = (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1));
  float *dj = (float *)malloc(sizeof(float) * c->n);
  for (uint64_t t = b; t < e; t++) {
    const uint64_t i = c->rows ? c->rows[t] : t;
    uint64_t sz = 0;
    for (uint64_t j = 0; j < c->n; j++) {
      if (j == i) continue;
      float dd = orc_euclid(c->x + i * c->d, c->x + j * c->d, c->d);
      dj[j] = dd;
      uint64_t key = ((uint64_t)orc_total_key(dd) << 32) | j;
      if (sz < k) {
        heap[sz++] = key;
        for (uint64_t p = sz - 1; p > 0;) { /* sift up */
          uint64_t q = (p - 1) / 2;
          if (heap[q] >= heap[p]) break;
          uint64_t t = heap[p];
          heap[p] = heap[q];
          heap[q] = t;
          p = q;
        }
      } else if (k && key < heap[0]) {
        heap[0] = key;
        orc_heap_sift_down(heap, sz, 0);
      }
    }
    qsort(heap, sz, sizeof(uint64_t), orc_u64_cmp);
    for (uint64_t c2 = 0; c2 < k; c2++) {
      uint32_t j = (uint32_t)(heap[c2] & 0xffffffffu);
      c->idx[t * k + c2] = j;
      c->dist[t * k + c2] = dj[j];
    }
  }
  free(heap);
  free(dj);
}

int orc_knn_rowwise(const float *x, uint64_t n, uint64_t d, uint64_t k, uint32_t *idx, float *dist,
                    int nthreads) {
  if (n == 0) return 0;
  if (k > n - 1) k = n - 1;
  orc_row_ctx c = {x, n, d, k, idx, dist, NULL};
  orc_parallel_for(n, nthreads, orc_row_task, &c);
  return 0;
}

int orc_knn_rows(const float *x, uint64_t n, uint64_t d, uint64_t k, const uint32_t *rows,
                 uint64_t nrows, uint32_t *idx, float *dist, int nthreads) {
  if (n == 0 || nrows == 0) return 0;
  if (k > n - 1) k = n - 1;
  for (uint64_t t = 0; t < nrows; t++)
    if (rows[t] >= n) return -1;
  orc_row_ctx c = {x, n, d, k, idx, dist, rows};
  orc_parallel_for(nrows, nthreads, orc_row_task, &c);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* lib.rs:565-597                                                            */
/* ------------------------------------------------------------------------ */
int orc_decide_num_pairs(uint64_t n, int has_override, uint64_t override_k, float mn_ratio,
                         float fp_ratio, uint64_t out[3]) {
  uint64_t k;
  if (has_override) {
    k = override_k;
  } else if (n <= 10000) {
    k = 10;
  } else {
    /* lib.rs:576: f32 log10, round half away from zero, saturating cast */
    float lg = log10f((float)n);
    float v = 10.0f + 15.0f * (lg - 4.0f);
    float r = roundf(v);
    k = r <= 0.0f ? 0 : (uint64_t)r;
  }
  float mn = roundf((float)k * mn_ratio);
  float fp = roundf((float)k * fp_ratio);
  /* Rust `as usize` saturates: negative and NaN become 0 */
  uint64_t n_mn = (mn > 0.0f) ? (uint64_t)mn : 0;
  uint64_t n_fp = (fp > 0.0f) ? (uint64_t)fp : 0;
  out[0] = k;
  out[1] = n_mn;
  out[2] = n_fp;
  if (k < 1) return 5;
  if (n_fp < 1) return 6;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* rand 0.8.5 restatement (PARITY UNPINNED)                                  */
/* ------------------------------------------------------------------------ */
static uint32_t orc_pcg32(uint64_t *state) {
  /* rand_core 0.6 SeedableRng::seed_from_u64 default: PCG32 output stream */
  const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
  *state = *state * MUL + INC;
  uint64_t st = *state;
  uint32_t xorshifted = (uint32_t)(((st >> 18) ^ st) >> 27);
  uint32_t rot = (uint32_t)(st >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

static void orc_rng_from_seed(orc_rng *r, const uint8_t seed[32]) {
  int all_zero = 1;
  for (int b = 0; b < 32; b++) all_zero &= seed[b] == 0;
  if (all_zero) { /* Xoshiro256PlusPlus::from_seed: zero seed -> SplitMix64(0) */
    uint64_t state = 0;
    for (int w = 0; w < 4; w++) {
      state += 0x9e3779b97f4a7c15ull;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      r->s[w] = z ^ (z >> 31);
    }
    return;
  }
  for (int w = 0; w < 4; w++) {
    uint64_t v = 0;
    for (int b = 7; b >= 0; b--) v = (v << 8) | seed[8 * w + b];
    r->s[w] = v;
  }
}

void orc_rng_seed_from_bytes(orc_rng *r, const uint8_t seed[32]) { orc_rng_from_seed(r, seed); }

void orc_rng_seed_from_u64(orc_rng *r, uint64_t state) {
  uint8_t seed[32];
  for (int c = 0; c < 8; c++) {
    uint32_t x = orc_pcg32(&state);
    seed[4 * c + 0] = (uint8_t)(x);
    seed[4 * c + 1] = (uint8_t)(x >> 8);
    seed[4 * c + 2] = (uint8_t)(x >> 16);
    seed[4 * c + 3] = (uint8_t)(x >> 24);
  }
  orc_rng_from_seed(r, seed);
}

static inline uint64_t orc_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t orc_rng_next_u64(orc_rng *r) {
  uint64_t *s = r->s;
  uint64_t result = orc_rotl(s[0] + s[3], 23) + s[0];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = orc_rotl(s[3], 45);
  return result;
}

uint32_t orc_rng_next_u32(orc_rng *r) { return (uint32_t)(orc_rng_next_u64(r) >> 32); }

uint32_t orc_rng_gen_range_u32(orc_rng *r, uint32_t low, uint32_t high) {
  /* UniformInt<u32>::sample_single -> sample_single_inclusive(low, high-1):
   * widening multiply with the zone (range << lz(range)) - 1 */
  uint32_t range = high - low;
  if (range == 0) return orc_rng_next_u32(r);
  uint32_t zone = (range << __builtin_clz(range)) - 1u;
  for (;;) {
    uint32_t v = orc_rng_next_u32(r);
    uint64_t m = (uint64_t)v * (uint64_t)range;
    uint32_t hi = (uint32_t)(m >> 32), lo = (uint32_t)m;
    if (lo <= zone) return low + hi;
  }
}

/* rand 0.8.5 Standard for f64 (53 high bits scaled by 2^-53) and Open01
 * (52 high bits as a [1, 2) mantissa, minus 1 - 2^-53). */
static double orc_rng_f64(orc_rng *r) {
  return (double)(orc_rng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0);
}
static double orc_f64_with_exponent(uint64_t mant52, int e) {
  uint64_t b = mant52 | ((uint64_t)(1023 + e) << 52);
  double v;
  memcpy(&v, &b, sizeof(v));
  return v;
}
static double orc_rng_open01(orc_rng *r) {
  return orc_f64_with_exponent(orc_rng_next_u64(r) >> 12, 0) - (1.0 - 2.220446049250313e-16 / 2.0);
}

/* rand_distr 0.4.3 ziggurat tables for N(0, 1): the ZIGNOR construction
 * (Doornik 2005) with 256 layers, R = 3.6541528853610088, V = 0.00492867323399:
 * x[0] = V / f(R), x[1] = R, x[i] = f^-1(V / x[i-1] + f(x[i-1])), x[256] = 0,
 * F[i] = f(x[i]) with f(x) = exp(-x^2 / 2). */
#define ORC_ZIG_R 3.6541528853610088
#define ORC_ZIG_V 0.00492867323399
static double orc_zig_x[257], orc_zig_f[257];
static pthread_once_t orc_zig_once = PTHREAD_ONCE_INIT;
static double orc_norm_pdf(double x) { return exp(-x * x / 2.0); }
static void orc_zig_init(void) {
  orc_zig_x[0] = ORC_ZIG_V / orc_norm_pdf(ORC_ZIG_R);
  orc_zig_x[1] = ORC_ZIG_R;
  for (int i = 2; i < 256; i++) {
    double last = orc_zig_x[i - 1];
    orc_zig_x[i] = sqrt(-2.0 * log(ORC_ZIG_V / last + orc_norm_pdf(last)));
  }
  orc_zig_x[256] = 0.0;
  for (int i = 0; i < 257; i++) orc_zig_f[i] = orc_norm_pdf(orc_zig_x[i]);
}

void orc_zig_tables(double x[257], double f[257]) {
  pthread_once(&orc_zig_once, orc_zig_init);
  memcpy(x, orc_zig_x, sizeof(orc_zig_x));
  memcpy(f, orc_zig_f, sizeof(orc_zig_f));
}

/* rand_distr 0.4.3 StandardNormal for f64: utils.rs ziggurat(symmetric) with
 * the normal pdf and the tail sampler of normal.rs (zero_case). */
double orc_standard_normal(orc_rng *r) {
  pthread_once(&orc_zig_once, orc_zig_init);
  for (;;) {
    uint64_t bits = orc_rng_next_u64(r);
    unsigned i = (unsigned)(bits & 0xff);
    double u = orc_f64_with_exponent(bits >> 12, 1) - 3.0; /* [2, 4) - 3 = [-1, 1) */
    double x = u * orc_zig_x[i];
    if (fabs(x) < orc_zig_x[i + 1]) return x;
    if (i == 0) { /* the tail beyond R */
      double xt = 1.0, yt = 0.0;
      while (-2.0 * yt < xt * xt) {
        double x_ = orc_rng_open01(r), y_ = orc_rng_open01(r);
        xt = log(x_) / ORC_ZIG_R;
        yt = log(y_);
      }
      return u < 0.0 ? xt - ORC_ZIG_R : ORC_ZIG_R - xt;
    }
    if (orc_zig_f[i + 1] + (orc_zig_f[i] - orc_zig_f[i + 1]) * orc_rng_f64(r) < orc_norm_pdf(x)) return x;
  }
}

/* lib.rs:769-778: Normal<f32>(0, 1) draws in row-major order from
 * SmallRng::seed_from_u64(seed); each is mean + std_dev * (z as f32). */
void orc_random_init(uint64_t n, uint64_t dims, uint64_t seed, float *y) {
  orc_rng r;
  orc_rng_seed_from_u64(&r, seed);
  for (uint64_t e = 0; e < n * dims; e++) y[e] = 0.0f + 1.0f * (float)orc_standard_normal(&r);
}

/* ------------------------------------------------------------------------ */
/* sampling.rs                                                               */
/* ------------------------------------------------------------------------ */
uint64_t orc_sample_fp(uint64_t n_samples, uint32_t maximum, const uint32_t *reject,
                       uint64_t n_reject, uint64_t reject_stride, uint32_t self_ind, orc_rng *rng,
                       uint32_t *out) {
  /* sampling.rs:44-46 */
  int self_absent = 1;
  for (uint64_t q = 0; q < n_reject; q++)
    if (reject[q * reject_stride] == self_ind) self_absent = 0;
  uint64_t avail = maximum;
  avail = avail > n_reject ? avail - n_reject : 0;
  avail = avail > (uint64_t)self_absent ? avail - (uint64_t)self_absent : 0;
  if (n_samples > avail) n_samples = avail;
  uint64_t len = 0;
  while (len < n_samples) { /* sampling.rs:50-59 */
    uint32_t j = orc_rng_gen_range_u32(rng, 0, maximum);
    if (j == self_ind) continue;
    int dup = 0;
    for (uint64_t q = 0; q < len && !dup; q++) dup = out[q] == j;
    for (uint64_t q = 0; q < n_reject && !dup; q++) dup = reject[q * reject_stride] == j;
    if (!dup) out[len++] = j;
  }
  return len;
}

/* The n_fp far-pair rows of point i (sampling.rs:90-101): rows (out) start
 * zeroed (sampling.rs:84); buf holds n_fp words. */
static void orc_fp_point(uint64_t n, const uint32_t *pair_neighbors, uint64_t n_neighbors,
                         uint64_t n_fp, uint64_t seed, uint64_t i, uint32_t *buf, uint32_t *out) {
  orc_rng rng;
  orc_rng_seed_from_u64(&rng, seed + i); /* sampling.rs:97 */
  /* reject = pair_neighbors[i*k .. (i+1)*k, 1] (sampling.rs:94-95) */
  uint64_t got = orc_sample_fp(n_fp, (uint32_t)n, pair_neighbors + 2 * i * n_neighbors + 1,
                               n_neighbors, 2, (uint32_t)i, &rng, buf);
  for (uint64_t q = 0; q < got; q++) { /* sampling.rs:124-132 */
    out[2 * q] = (uint32_t)i;
    out[2 * q + 1] = buf[q];
  }
}

void orc_sample_fp_pair_det(uint64_t n, const uint32_t *pair_neighbors, uint64_t n_neighbors,
                            uint64_t n_fp, uint64_t seed, uint32_t *pair_fp) {
  memset(pair_fp, 0, sizeof(uint32_t) * 2 * n * n_fp); /* sampling.rs:84 zeros */
  uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (n_fp ? n_fp : 1));
  for (uint64_t i = 0; i < n; i++)
    orc_fp_point(n, pair_neighbors, n_neighbors, n_fp, seed, i, buf, pair_fp + 2 * i * n_fp);
  free(buf);
}

void orc_sample_fp_rows(uint64_t n, const uint32_t *pair_neighbors, uint64_t n_neighbors,
                        uint64_t n_fp, uint64_t seed, const uint32_t *rows, uint64_t nrows,
                        uint32_t *out) {
  memset(out, 0, sizeof(uint32_t) * 2 * nrows * n_fp);
  uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (n_fp ? n_fp : 1));
  for (uint64_t t = 0; t < nrows; t++)
    orc_fp_point(n, pair_neighbors, n_neighbors, n_fp, seed, rows[t], buf, out + 2 * t * n_fp);
  free(buf);
}

typedef struct {
  float d;
  uint32_t idx;
} orc_di;

/* The n_mn mid-near rows of point i (sampling.rs:128-139), rows zeroed. */
static void orc_mn_point(const float *x, uint64_t n, uint64_t d, uint64_t n_mn, uint64_t seed,
                         uint64_t i, uint32_t *rows) {
  orc_rng rng;
  orc_rng_seed_from_u64(&rng, seed + i); /* sampling.rs:133 */
  for (uint64_t j = 0; j < n_mn; j++) {
    uint32_t sampled[6];
    /* reject = rows[0..j, 1] (sampling.rs:135) */
    uint64_t got = orc_sample_fp(6, (uint32_t)n, rows + 1, j, 2, (uint32_t)i, &rng, sampled);
    /* sampling.rs:302-309: six (dist, idx) slots initialised to (0.0, 0) */
    orc_di e[6];
    for (int q = 0; q < 6; q++) {
      e[q].d = 0.0f;
      e[q].idx = 0;
    }
    for (uint64_t q = 0; q < got; q++) {
      e[q].d = orc_euclid(x + i * d, x + (uint64_t)sampled[q] * d, d);
      e[q].idx = sampled[q];
    }
    /* sort_unstable_by(total_cmp) on six entries: insertion order kept for ties */
    for (int a = 1; a < 6; a++) {
      orc_di t = e[a];
      int b = a - 1;
      while (b >= 0 && orc_total_key(e[b].d) > orc_total_key(t.d)) {
        e[b + 1] = e[b];
        b--;
      }
      e[b + 1] = t;
    }
    rows[2 * j] = (uint32_t)i; /* sampling.rs:310-312 */
    rows[2 * j + 1] = e[1].idx;
  }
}

void orc_sample_mn_pair_det(const float *x, uint64_t n, uint64_t d, uint64_t n_mn, uint64_t seed,
                            uint32_t *pair_mn) {
  memset(pair_mn, 0, sizeof(uint32_t) * 2 * n * n_mn);
  for (uint64_t i = 0; i < n; i++) orc_mn_point(x, n, d, n_mn, seed, i, pair_mn + 2 * i * n_mn);
}

void orc_sample_mn_rows(const float *x, uint64_t n, uint64_t d, uint64_t n_mn, uint64_t seed,
                        const uint32_t *rows, uint64_t nrows, uint32_t *out) {
  memset(out, 0, sizeof(uint32_t) * 2 * nrows * n_mn);
  for (uint64_t t = 0; t < nrows; t++) orc_mn_point(x, n, d, n_mn, seed, rows[t], out + 2 * t * n_mn);
}

void orc_sample_neighbors_pair(const float *scaled, const uint32_t *nbrs, uint64_t n,
                               uint64_t kprime, uint64_t n_neighbors, uint32_t *pairs) {
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (kprime ? kprime : 1));
  for (uint64_t i = 0; i < n; i++) {
    for (uint64_t c = 0; c < kprime; c++)
      keys[c] = ((uint64_t)orc_total_key(scaled[i * kprime + c]) << 32) | c;
    qsort(keys, kprime, sizeof(uint64_t), orc_u64_cmp); /* sampling.rs:265 */
    for (uint64_t j = 0; j < n_neighbors; j++) {        /* sampling.rs:276-283 */
      uint64_t col = keys[j] & 0xffffffffu;
      pairs[2 * (i * n_neighbors + j)] = (uint32_t)i;
      pairs[2 * (i * n_neighbors + j) + 1] = nbrs[i * kprime + col];
    }
  }
  free(keys);
}

/* ------------------------------------------------------------------------ */
/* lib.rs:753-766 and 486-511                                                */
/* ------------------------------------------------------------------------ */
void orc_value_init(const float *y_in, uint64_t n, uint64_t dims, float *y_out) {
  float *mean = (float *)calloc(dims, sizeof(float));
  float *wm = (float *)calloc(dims, sizeof(float));
  float *ssq = (float *)calloc(dims, sizeof(float));
  /* mean_axis(Axis(0)): rows summed in order from zero, then / n */
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t d = 0; d < dims; d++) mean[d] = mean[d] + y_in[i * dims + d];
  for (uint64_t d = 0; d < dims; d++) mean[d] = mean[d] / (float)n;
  /* std_axis(Axis(0), 0): ndarray var_axis running mean with mul_add */
  for (uint64_t i = 0; i < n; i++) {
    float count = (float)(i + 1);
    for (uint64_t d = 0; d < dims; d++) {
      float x = y_in[i * dims + d];
      float delta = x - wm[d];
      wm[d] = wm[d] + delta / count;
      ssq[d] = fmaf(x - wm[d], delta, ssq[d]);
    }
  }
  for (uint64_t d = 0; d < dims; d++) ssq[d] = sqrtf(ssq[d] / (float)n);
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t d = 0; d < dims; d++) {
      float c = y_in[i * dims + d] - mean[d];
      float s = c * 0.0001f;
      y_out[i * dims + d] = s / ssq[d]; /* lib.rs:762 */
    }
  free(mean);
  free(wm);
  free(ssq);
}

int orc_normalize(const float *x, uint64_t n, uint64_t d, float *x_out) {
  uint64_t total = n * d;
  if (total == 0) return 7;
  uint32_t kmin = 0xffffffffu, kmax = 0;
  float xmin = 0, xmax = 0;
  for (uint64_t e = 0; e < total; e++) { /* lib.rs:489-497, total_cmp */
    uint32_t kk = orc_total_key(x[e]);
    if (e == 0 || kk < kmin) {
      kmin = kk;
      xmin = x[e];
    }
    if (e == 0 || kk > kmax) {
      kmax = kk;
      xmax = x[e];
    }
  }
  for (uint64_t e = 0; e < total; e++) x_out[e] = (x[e] - xmin) / xmax; /* lib.rs:500-503 */
  float *mean = (float *)calloc(d, sizeof(float));
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t c = 0; c < d; c++) mean[c] = mean[c] + x_out[i * d + c];
  for (uint64_t c = 0; c < d; c++) mean[c] = mean[c] / (float)n; /* lib.rs:506-508 */
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t c = 0; c < d; c++) x_out[i * d + c] = x_out[i * d + c] - mean[c];
  free(mean);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* lib.rs:781-837                                                            */
/* ------------------------------------------------------------------------ */
int orc_optimize(float *y, uint64_t n, uint64_t dims, const uint32_t *nn, uint64_t n_nn,
                 const uint32_t *mn, uint64_t n_mn, const uint32_t *fp, uint64_t n_fp, float lr,
                 const uint64_t iters[3], const uint64_t *snap_idx, uint64_t ns, float *snap,
                 float *loss_out, int nthreads) {
  const float w_mn_init = 1000.0f, beta1 = 0.9f, beta2 = 0.999f; /* lib.rs:782-784 */
  uint64_t nd = n * dims;
  float *m = (float *)calloc(nd ? nd : 1, sizeof(float));
  float *v = (float *)calloc(nd ? nd : 1, sizeof(float));
  float *g = (float *)malloc(sizeof(float) * (n + 1) * dims);
  if (snap && ns) { /* lib.rs:789-791 */
    memset(snap, 0, sizeof(float) * ns * nd);
    memcpy(snap, y, sizeof(float) * nd);
  }
  uint64_t total = iters[0] + iters[1] + iters[2];
  int rc = 0;
  for (uint64_t itr = 0; itr < total; itr++) {
    float w[3];
    orc_find_weights(w_mn_init, itr, iters[0], iters[1], w);
    rc = orc_grad(y, n, dims, nn, n_nn, mn, n_mn, fp, n_fp, w[0], w[1], w[2], g, nthreads);
    if (rc) break;
    if (loss_out) loss_out[itr] = g[n * dims];
    orc_adam(y, g, m, v, n, dims, beta1, beta2, lr, itr);
    if (snap) /* lib.rs:834-836: first position equal to itr+1 */
      for (uint64_t s = 0; s < ns; s++)
        if (snap_idx[s] == itr + 1) {
          memcpy(snap + s * nd, y, sizeof(float) * nd);
          break;
        }
  }
  free(m);
  free(v);
  free(g);
  return rc;
}
