/*
 * mvref.c — CPU oracle (TEST INFRASTRUCTURE ONLY) for the MojoVec batched
 * vector-scan path. See mvref.h for the contract. Only tests/, the smoke
 * check and the bench's CPU-baseline leg load this library.
 *
 * Floating-point rules of the restatement (build with -ffp-contract=off so
 * that only the explicit fmaf calls fuse):
 *   - `fma(a, b, acc)` in the reference is fmaf here; `a * b` then `+` stays
 *     two roundings.
 *   - SIMD[float32, W].reduce_add() is restated as the halving tree
 *     v[j] += v[j + W/2], then W/4, ... down to 1 lane. The Mojo stdlib lane
 *     order is not vendored with the reference: parity unpinned for the last
 *     ulp, pinned against exact integer-valued data where every order agrees.
 *   - math.round is restated as roundf (ties away from zero): parity unpinned
 *     at exact .5 ties.
 */
#include "mvref.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define IVF_PAD_DIST 1e38f

static int nthreads_or_default(int nthreads) {
  return nthreads > 0 ? nthreads : omp_get_max_threads();
}

/* Halving-tree reduce_add restatement (see file header). */
static float reduce_add(float *v, int w) {
  for (int h = w / 2; h >= 1; h >>= 1)
    for (int j = 0; j < h; j++) v[j] = v[j] + v[j + h];
  return v[0];
}

/* l2_distance_simd[W]: mojovec/utils/distances.mojo:43-143. */
float mvref_l2_simd(const float *x, const float *y, int64_t d, int W) {
  if (d < W) { /* :57-62 scalar path */
    float r = 0.0f;
    for (int64_t i = 0; i < d; i++) {
      float df = x[i] - y[i];
      r = r + df * df;
    }
    return r;
  }
  float a0[64], a1[64], a2[64], a3[64];
  for (int j = 0; j < W; j++) a0[j] = a1[j] = a2[j] = a3[j] = 0.0f;
  int64_t i = 0;
  while (i <= d - 4 * W) { /* :75-96 four accumulators */
    for (int j = 0; j < W; j++) {
      float d0 = x[i + j] - y[i + j];
      float d1 = x[i + W + j] - y[i + W + j];
      float d2 = x[i + 2 * W + j] - y[i + 2 * W + j];
      float d3 = x[i + 3 * W + j] - y[i + 3 * W + j];
      a0[j] = fmaf(d0, d0, a0[j]);
      a1[j] = fmaf(d1, d1, a1[j]);
      a2[j] = fmaf(d2, d2, a2[j]);
      a3[j] = fmaf(d3, d3, a3[j]);
    }
    i += 4 * W;
  }
  float dist[64];
  for (int j = 0; j < W; j++) dist[j] = ((a0[j] + a1[j]) + a2[j]) + a3[j]; /* :98 */
  while (i <= d - W) { /* :101-106 */
    for (int j = 0; j < W; j++) {
      float df = x[i + j] - y[i + j];
      dist[j] = fmaf(df, df, dist[j]);
    }
    i += W;
  }
  float res = reduce_add(dist, W); /* :108 */
  float r[16];
  if (W >= 32) /* :110-117 */
    while (i <= d - 16) {
      for (int j = 0; j < 16; j++) { float df = x[i + j] - y[i + j]; r[j] = df * df; }
      res += reduce_add(r, 16);
      i += 16;
    }
  if (W >= 16) /* :119-126 */
    while (i <= d - 8) {
      for (int j = 0; j < 8; j++) { float df = x[i + j] - y[i + j]; r[j] = df * df; }
      res += reduce_add(r, 8);
      i += 8;
    }
  if (W >= 8) /* :128-135 */
    while (i <= d - 4) {
      for (int j = 0; j < 4; j++) { float df = x[i + j] - y[i + j]; r[j] = df * df; }
      res += reduce_add(r, 4);
      i += 4;
    }
  while (i < d) { /* :138-141 */
    float df = x[i] - y[i];
    res += df * df;
    i++;
  }
  return res;
}

/* inner_product_simd[W]: mojovec/utils/distances.mojo:146-228. */
float mvref_ip_simd(const float *x, const float *y, int64_t d, int W) {
  if (d < W) {
    float r = 0.0f;
    for (int64_t i = 0; i < d; i++) r = r + x[i] * y[i];
    return r;
  }
  float p0[64], p1[64], p2[64], p3[64];
  for (int j = 0; j < W; j++) p0[j] = p1[j] = p2[j] = p3[j] = 0.0f;
  int64_t i = 0;
  while (i <= d - 4 * W) {
    for (int j = 0; j < W; j++) {
      p0[j] = fmaf(x[i + j], y[i + j], p0[j]);
      p1[j] = fmaf(x[i + W + j], y[i + W + j], p1[j]);
      p2[j] = fmaf(x[i + 2 * W + j], y[i + 2 * W + j], p2[j]);
      p3[j] = fmaf(x[i + 3 * W + j], y[i + 3 * W + j], p3[j]);
    }
    i += 4 * W;
  }
  float prod[64];
  for (int j = 0; j < W; j++) prod[j] = ((p0[j] + p1[j]) + p2[j]) + p3[j];
  while (i <= d - W) {
    for (int j = 0; j < W; j++) prod[j] = fmaf(x[i + j], y[i + j], prod[j]);
    i += W;
  }
  float res = reduce_add(prod, W);
  float r[16];
  if (W >= 32)
    while (i <= d - 16) {
      for (int j = 0; j < 16; j++) r[j] = fmaf(x[i + j], y[i + j], 0.0f);
      res += reduce_add(r, 16);
      i += 16;
    }
  if (W >= 16)
    while (i <= d - 8) {
      for (int j = 0; j < 8; j++) r[j] = fmaf(x[i + j], y[i + j], 0.0f);
      res += reduce_add(r, 8);
      i += 8;
    }
  if (W >= 8)
    while (i <= d - 4) {
      for (int j = 0; j < 4; j++) r[j] = fmaf(x[i + j], y[i + j], 0.0f);
      res += reduce_add(r, 4);
      i += 4;
    }
  while (i < d) {
    res += x[i] * y[i];
    i++;
  }
  return res;
}

/* l2_distance_short4: mojovec/utils/distances.mojo:9-23. */
float mvref_l2_short4(const float *x, const float *y, int64_t d) {
  if (d == 4) {
    float r[4];
    for (int j = 0; j < 4; j++) { float df = x[j] - y[j]; r[j] = df * df; }
    return reduce_add(r, 4);
  }
  float res = 0.0f;
  for (int64_t i = 0; i < d; i++) {
    float df = x[i] - y[i];
    res += df * df;
  }
  return res;
}

/* inner_product_short4: mojovec/utils/distances.mojo:26-39. */
float mvref_ip_short4(const float *x, const float *y, int64_t d) {
  if (d == 4) {
    float r[4];
    for (int j = 0; j < 4; j++) r[j] = x[j] * y[j];
    return reduce_add(r, 4);
  }
  float res = 0.0f;
  for (int64_t i = 0; i < d; i++) res += x[i] * y[i];
  return res;
}

/* sq8_dot_product_simd (generic widening path): distances.mojo:286-340.
 * Exact integer arithmetic, so any summation order gives the same result. */
uint32_t mvref_dot_u8(const uint8_t *x, const uint8_t *y, int64_t d) {
  uint32_t s = 0;
  for (int64_t i = 0; i < d; i++) s += (uint32_t)x[i] * (uint32_t)y[i];
  return s;
}

/* sq8_signed_dot_product_simd: distances.mojo:344-434 (exact). */
int32_t mvref_dot_i8(const int8_t *x, const int8_t *y, int64_t d) {
  int32_t s = 0;
  for (int64_t i = 0; i < d; i++) s += (int32_t)x[i] * (int32_t)y[i];
  return s;
}

/* sq8_l2_from_dot: distances.mojo:437-451. */
uint32_t mvref_l2_from_dot(uint32_t na, uint32_t nb, uint32_t dot) {
  int64_t v = (int64_t)na + (int64_t)nb - 2 * (int64_t)dot;
  return (uint32_t)(v > 0 ? v : 0);
}

void mvref_l2_simd_batch(const float *x, const float *y, int64_t n, int64_t d,
                         int W, float *out) {
  for (int64_t i = 0; i < n; i++) out[i] = mvref_l2_simd(x + i * d, y + i * d, d, W);
}

void mvref_ip_simd_batch(const float *x, const float *y, int64_t n, int64_t d,
                         int W, float *out) {
  for (int64_t i = 0; i < n; i++) out[i] = mvref_ip_simd(x + i * d, y + i * d, d, W);
}

/* l2_distance_simd_batch4[W], one of its four rows: distances.mojo:541-605.
 * One W-lane accumulator over W-element blocks (fma(diff, diff, acc), diff =
 * x - y), the halving tree, then a scalar tail res += diff * diff. */
float mvref_l2_simd_b4(const float *x, const float *y, int64_t d, int W) {
  float acc[64];
  for (int j = 0; j < W; j++) acc[j] = 0.0f;
  int64_t i = 0;
  for (; i + W <= d; i += W) /* :556-576 */
    for (int j = 0; j < W; j++) {
      float df = x[i + j] - y[i + j];
      acc[j] = fmaf(df, df, acc[j]);
    }
  float res = reduce_add(acc, W); /* :578-581 */
  for (; i < d; i++) { /* :583-597 */
    float df = x[i] - y[i];
    res += df * df;
  }
  return res;
}

/* inner_product_simd_batch4[W], one row: distances.mojo:608-660 (acc =
 * fma(x, y, acc) over W-element blocks, halving tree, res += x * y tail). */
float mvref_ip_simd_b4(const float *x, const float *y, int64_t d, int W) {
  float acc[64];
  for (int j = 0; j < W; j++) acc[j] = 0.0f;
  int64_t i = 0;
  for (; i + W <= d; i += W)
    for (int j = 0; j < W; j++) acc[j] = fmaf(x[i + j], y[i + j], acc[j]);
  float res = reduce_add(acc, W);
  for (; i < d; i++) res += x[i] * y[i];
  return res;
}

/* FlatDistanceComputer (index_flat.mojo:36-115, SIMD_WIDTH = 8 at :9) over a
 * batch: out[p] = distance(q[p / per_query], row ids[p]). order 0 is
 * distance() (:60-68: l2_distance_simd[8] / -inner_product_simd[8]), order 1
 * distance_batch4() (:70-90: l2_distance_simd_batch4[8] /
 * -inner_product_simd_batch4[8]). */
void mvref_flat_distances(const float *base, int d, int metric, const float *q,
                          int64_t nq, const int64_t *ids, int64_t per_query,
                          int order, float *out) {
  for (int64_t p = 0; p < nq * per_query; p++) {
    const float *qv = q + (p / per_query) * d, *xv = base + ids[p] * d;
    float v;
    if (order == 0)
      v = metric == MVREF_L2 ? mvref_l2_simd(qv, xv, d, 8) : -mvref_ip_simd(qv, xv, d, 8);
    else
      v = metric == MVREF_L2 ? mvref_l2_simd_b4(xv, qv, d, 8) : -mvref_ip_simd_b4(xv, qv, d, 8);
    out[p] = v;
  }
}

/* FlatDistanceComputer.symmetric_distance (index_flat.mojo:92-99). */
void mvref_flat_pair_distances(const float *base, int d, int metric, const int64_t *a,
                               const int64_t *b, int64_t npairs, float *out) {
  for (int64_t p = 0; p < npairs; p++) {
    const float *xa = base + a[p] * d, *xb = base + b[p] * d;
    out[p] = metric == MVREF_L2 ? mvref_l2_simd(xa, xb, d, 8) : -mvref_ip_simd(xa, xb, d, 8);
  }
}

/* ------------------------------------------------------------------------ */
/* Reference max-heap (mojovec/utils/heap.mojo:17-213), literal restatement. */

static void heap_swap(float *hd, int64_t *hl, int64_t a, int64_t b) {
  float td = hd[a]; hd[a] = hd[b]; hd[b] = td;
  int64_t tl = hl[a]; hl[a] = hl[b]; hl[b] = tl;
}

static void ref_max_heap_push(float *hd, int64_t *hl, int64_t size, float dist, int64_t label) {
  int64_t i = size; /* heap.mojo:56-84 */
  hd[i] = dist; hl[i] = label;
  while (i > 0) {
    int64_t parent = (i - 1) / 2;
    if (hd[parent] < hd[i]) { heap_swap(hd, hl, i, parent); i = parent; }
    else break;
  }
}

static void ref_max_heap_replace_top(float *hd, int64_t *hl, int64_t k, float dist, int64_t label) {
  hd[0] = dist; hl[0] = label; /* heap.mojo:17-53 */
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = 2 * i + 2, largest = i;
    if (l < k && hd[l] > hd[largest]) largest = l;
    if (r < k && hd[r] > hd[largest]) largest = r;
    if (largest != i) { heap_swap(hd, hl, i, largest); i = largest; }
    else break;
  }
}

static void ref_max_heap_pop(float *hd, int64_t *hl, int64_t size, float *od, int64_t *ol) {
  *od = hd[0]; *ol = hl[0]; /* heap.mojo:167-213 */
  int64_t last = size - 1;
  if (last == 0) return;
  hd[0] = hd[last]; hl[0] = hl[last];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = 2 * i + 2, largest = i;
    if (l < last && hd[l] > hd[largest]) largest = l;
    if (r < last && hd[r] > hd[largest]) largest = r;
    if (largest != i) { heap_swap(hd, hl, i, largest); i = largest; }
    else break;
  }
}

/* Scan + pop loop of index_flat.mojo:285-311 with ids 0..n-1: the
 * reference's own (not lower-id-stable) tie behaviour. */
void mvref_heap_topk_reference(const float *dist, int64_t n, int k,
                               float *out_d, int64_t *out_i, int *found) {
  int64_t hs = 0;
  for (int64_t j = 0; j < n; j++) {
    if (hs < k) { ref_max_heap_push(out_d, out_i, hs, dist[j], j); hs++; }
    else if (dist[j] < out_d[0]) ref_max_heap_replace_top(out_d, out_i, k, dist[j], j);
  }
  int64_t cur = hs;
  for (int64_t j = 0; j < cur; j++) {
    float pd; int64_t pl;
    ref_max_heap_pop(out_d, out_i, hs, &pd, &pl);
    hs--;
    out_d[cur - 1 - j] = pd;
    out_i[cur - 1 - j] = pl;
  }
  *found = (int)cur;
}

/* ------------------------------------------------------------------------ */
/* Canonical top-k: the k smallest (dist, id) pairs in lexicographic order,
 * i.e. the reference heap's selection with exact ties broken by lower id. */

typedef struct { float d; int64_t id; } cand_t;

static int cand_less(cand_t a, cand_t b) {
  return a.d < b.d || (a.d == b.d && a.id < b.id);
}

typedef struct { cand_t *h; int size; int cap; } cheap_t;

static void ch_init(cheap_t *c, cand_t *buf, int cap) { c->h = buf; c->size = 0; c->cap = cap; }

static void ch_offer(cheap_t *c, float d, int64_t id) {
  cand_t x = {d, id};
  if (c->cap <= 0) return;
  if (c->size < c->cap) {
    int i = c->size++;
    c->h[i] = x;
    while (i > 0) {
      int p = (i - 1) / 2;
      if (cand_less(c->h[p], c->h[i])) { cand_t t = c->h[p]; c->h[p] = c->h[i]; c->h[i] = t; i = p; }
      else break;
    }
  } else if (cand_less(x, c->h[0])) {
    c->h[0] = x;
    int i = 0, n = c->size;
    for (;;) {
      int l = 2 * i + 1, r = 2 * i + 2, g = i;
      if (l < n && cand_less(c->h[g], c->h[l])) g = l;
      if (r < n && cand_less(c->h[g], c->h[r])) g = r;
      if (g == i) break;
      cand_t t = c->h[g]; c->h[g] = c->h[i]; c->h[i] = t; i = g;
    }
  }
}

static int cand_cmp(const void *a, const void *b) {
  cand_t x = *(const cand_t *)a, y = *(const cand_t *)b;
  if (cand_less(x, y)) return -1;
  if (cand_less(y, x)) return 1;
  return 0;
}

static void ch_sorted(cheap_t *c) { qsort(c->h, (size_t)c->size, sizeof(cand_t), cand_cmp); }

void mvref_topk_canonical(const float *dist, const int64_t *ids, int64_t n,
                          int k, float *out_d, int64_t *out_i, int *found) {
  cand_t *buf = (cand_t *)malloc(sizeof(cand_t) * (size_t)(k > 0 ? k : 1));
  cheap_t c;
  ch_init(&c, buf, k);
  for (int64_t j = 0; j < n; j++) ch_offer(&c, dist[j], ids ? ids[j] : j);
  ch_sorted(&c);
  for (int j = 0; j < c.size; j++) { out_d[j] = c.h[j].d; out_i[j] = c.h[j].id; }
  *found = c.size;
  free(buf);
}

/* ------------------------------------------------------------------------ */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), key = 64-bit seed,
 * counter = (lo, hi, 0, 0). Restatement of std.random.philox.Random as used
 * by kmeans.mojo:70-82; the stdlib's counter/key layout is not vendored, so
 * the k-means initial picks are parity unpinned. */
void mvref_philox4x32_10(uint64_t seed, uint64_t counter, uint32_t out[4]) {
  uint32_t c0 = (uint32_t)counter, c1 = (uint32_t)(counter >> 32), c2 = 0, c3 = 0;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Initial centroids: kmeans.mojo:67-82. Returns sample_offset (:74). */
int64_t mvref_kmeans_init(const float *x, int64_t n, int d, int k,
                          uint64_t seed, float *centroids) {
  uint32_t w[4];
  uint64_t ctr = 0;
  mvref_philox4x32_10(seed, ctr++, w);
  int64_t sample_offset = (int64_t)(w[0] % (uint32_t)n);
  for (int i = 0; i < k; i++) {
    if (i != 0 && i % 4 == 0) mvref_philox4x32_10(seed, ctr++, w);
    int64_t src = (int64_t)(w[i % 4] % (uint32_t)n);
    memcpy(centroids + (int64_t)i * d, x + src * d, sizeof(float) * (size_t)d);
  }
  return sample_offset;
}

static float kmeans_dist(const float *x, const float *c, int d) {
  return d <= 4 ? mvref_l2_short4(x, c, d) : mvref_l2_simd(x, c, d, 8); /* kmeans.mojo:131-138 */
}

/* Lloyd iterations: kmeans.mojo:84-220 with 32 fixed chunks and an ordered
 * chunk reduction. fp64_sums selects the build's FP64 accumulation (sums are
 * rounded to FP32 once, then scaled by the FP32 reciprocal of the count as
 * the reference does); otherwise the reference's FP32 sums are restated. */
void mvref_kmeans_train(const float *x, int64_t n, int d, int k, int niter,
                        const float *init, int64_t sample_offset,
                        int fp64_sums, float *centroids, int32_t *assign,
                        int nthreads) {
  if (n == 0) return;
  const int nchunks = 32;
  int64_t chunk_size = (n + nchunks - 1) / nchunks;
  memcpy(centroids, init, sizeof(float) * (size_t)k * d);
  size_t acc_n = (size_t)nchunks * k * d;
  double *s64 = fp64_sums ? (double *)malloc(sizeof(double) * acc_n) : NULL;
  float *s32 = fp64_sums ? NULL : (float *)malloc(sizeof(float) * acc_n);
  int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * (size_t)nchunks * k);
  int nt = nthreads_or_default(nthreads);
  for (int it = 0; it < niter; it++) {
    if (s64) memset(s64, 0, sizeof(double) * acc_n);
    if (s32) memset(s32, 0, sizeof(float) * acc_n);
    memset(cnt, 0, sizeof(int64_t) * (size_t)nchunks * k);
#pragma omp parallel for schedule(static, 1) num_threads(nt)
    for (int ch = 0; ch < nchunks; ch++) {
      int64_t start = ch * chunk_size, end = start + chunk_size < n ? start + chunk_size : n;
      for (int64_t i = start; i < end; i++) {
        const float *xi = x + i * d;
        float best = 1e38f;
        int bc = -1;
        for (int c = 0; c < k; c++) {
          float dd = kmeans_dist(xi, centroids + (int64_t)c * d, d);
          if (dd < best) { best = dd; bc = c; }
        }
        int c = bc > 0 ? bc : 0; /* :147 */
        if (assign) assign[i] = c;
        cnt[(int64_t)ch * k + c]++;
        if (s64) {
          double *acc = s64 + ((int64_t)ch * k + c) * d;
          for (int j = 0; j < d; j++) acc[j] += (double)xi[j];
        } else {
          float *acc = s32 + ((int64_t)ch * k + c) * d;
          for (int j = 0; j < d; j++) acc[j] = acc[j] + xi[j];
        }
      }
    }
    for (int c = 0; c < k; c++) {
      int64_t count = 0;
      for (int ch = 0; ch < nchunks; ch++) count += cnt[(int64_t)ch * k + c];
      float *cp = centroids + (int64_t)c * d;
      if (count > 0) {
        float inv = 1.0f / (float)count; /* :201 */
        for (int j = 0; j < d; j++) {
          float sum;
          if (s64) {
            double t = 0.0;
            for (int ch = 0; ch < nchunks; ch++) t += s64[((int64_t)ch * k + c) * d + j];
            sum = (float)t;
          } else {
            float t = 0.0f;
            for (int ch = 0; ch < nchunks; ch++) t = t + s32[((int64_t)ch * k + c) * d + j];
            sum = t;
          }
          cp[j] = sum * inv;
        }
      } else { /* :211-220 empty-cluster reseed */
        int64_t src = (sample_offset + (int64_t)(it + 1) * k + c) % n;
        memcpy(cp, x + src * d, sizeof(float) * (size_t)d);
      }
    }
  }
  free(s64); free(s32); free(cnt);
}

/* ------------------------------------------------------------------------ */
/* ProductQuantizer: mojovec/quantization/pq.mojo. */

#define PQ_KSUB 256

/* train: pq.mojo:125-176; each subspace runs KMeans(dsub, 256, niter) with the
 * same seeded init (_train_pq_subspace :45-69). */
void mvref_pq_train(const float *x, int64_t n, int d, int M, int niter,
                    uint64_t seed, int fp64_sums, float *codebooks,
                    int nthreads) {
  int dsub = d / M;
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int m = 0; m < M; m++) {
    float *sub = (float *)malloc(sizeof(float) * (size_t)n * dsub);
    for (int64_t i = 0; i < n; i++)
      memcpy(sub + i * dsub, x + i * d + (int64_t)m * dsub, sizeof(float) * (size_t)dsub);
    float *init = (float *)malloc(sizeof(float) * PQ_KSUB * (size_t)dsub);
    int64_t off = mvref_kmeans_init(sub, n, dsub, PQ_KSUB, seed, init);
    mvref_kmeans_train(sub, n, dsub, PQ_KSUB, niter, init, off, fp64_sums,
                       codebooks + (int64_t)m * PQ_KSUB * dsub, NULL, 1);
    free(init);
    free(sub);
  }
}

static float pq_sub_l2(const float *a, const float *b, int dsub) {
  return dsub <= 4 ? mvref_l2_short4(a, b, dsub) : mvref_l2_simd(a, b, dsub, 8);
}

/* compute_codes: pq.mojo:178-219 (strict <, start 1e38, UInt8(best)). */
void mvref_pq_encode(const float *x, int64_t n, int d, int M,
                     const float *codebooks, uint8_t *codes, int nthreads) {
  int dsub = d / M;
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < n; i++) {
    for (int m = 0; m < M; m++) {
      const float *sx = x + i * d + (int64_t)m * dsub;
      const float *cm = codebooks + (int64_t)m * PQ_KSUB * dsub;
      float best = 1e38f;
      int bk = -1;
      for (int kk = 0; kk < PQ_KSUB; kk++) {
        float dd = pq_sub_l2(sx, cm + (int64_t)kk * dsub, dsub);
        if (dd < best) { best = dd; bk = kk; }
      }
      codes[i * M + m] = (uint8_t)bk;
    }
  }
}

/* decode: pq.mojo:221-250. */
void mvref_pq_decode(const uint8_t *codes, int64_t n, int d, int M,
                     const float *codebooks, float *x) {
  int dsub = d / M;
  for (int64_t i = 0; i < n; i++)
    for (int m = 0; m < M; m++)
      memcpy(x + i * d + (int64_t)m * dsub,
             codebooks + ((int64_t)m * PQ_KSUB + codes[i * M + m]) * dsub,
             sizeof(float) * (size_t)dsub);
}

/* compute_distance_table: pq.mojo:252-292. */
void mvref_pq_lut(const float *q, int d, int M, const float *codebooks,
                  int metric, float *lut) {
  int dsub = d / M;
  for (int m = 0; m < M; m++) {
    const float *sq = q + (int64_t)m * dsub;
    const float *cm = codebooks + (int64_t)m * PQ_KSUB * dsub;
    for (int kk = 0; kk < PQ_KSUB; kk++) {
      const float *c = cm + (int64_t)kk * dsub;
      float v;
      if (metric == MVREF_L2)
        v = dsub <= 4 ? mvref_l2_short4(sq, c, dsub) : mvref_l2_simd(sq, c, dsub, 8);
      else
        v = -(dsub <= 4 ? mvref_ip_short4(sq, c, dsub) : mvref_ip_simd(sq, c, dsub, 4));
      lut[m * PQ_KSUB + kk] = v;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* IndexFlat._search_impl (index_flat.mojo:260-325) with canonical ties.
 * Slots past the found count are padded with (0.0, -1). */
void mvref_flat_search(const float *base, int64_t n, int d, int metric,
                       const float *q, int64_t nq, int k,
                       const uint8_t *exclude, int64_t exclude_len,
                       float *D, int64_t *I, int nthreads) {
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
  {
    cand_t *buf = (cand_t *)malloc(sizeof(cand_t) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < nq; qi++) {
      const float *qp = q + qi * d;
      cheap_t c;
      ch_init(&c, buf, k);
      for (int64_t j = 0; j < n; j++) {
        if (exclude && j < exclude_len && exclude[j] > 0) continue;
        float dd = metric == MVREF_L2 ? mvref_l2_simd(qp, base + j * d, d, 8)
                                      : -mvref_ip_simd(qp, base + j * d, d, 8);
        ch_offer(&c, dd, j);
      }
      ch_sorted(&c);
      for (int j = 0; j < k; j++) {
        if (j < c.size) {
          float v = c.h[j].d;
          D[qi * k + j] = metric == MVREF_IP ? -v : v;
          I[qi * k + j] = c.h[j].id;
        } else {
          D[qi * k + j] = 0.0f;
          I[qi * k + j] = -1;
        }
      }
    }
    free(buf);
  }
}

/* ------------------------------------------------------------------------ */
/* SQ8: sq8.mojo:13-79 and index_flat_sq8.mojo:286-526. */

uint32_t mvref_sq8_encode_vec(const float *x, int64_t d, int symmetric,
                              float minimum, float inverse_scale,
                              uint8_t *code) {
  uint32_t norm = 0;
  if (symmetric) { /* sq8.mojo:22-29 */
    for (int64_t i = 0; i < d; i++) {
      float s = roundf(x[i] * inverse_scale);
      s = s < -127.0f ? -127.0f : (s > 127.0f ? 127.0f : s);
      int8_t c = (int8_t)s;
      code[i] = (uint8_t)c;
      int32_t w = c;
      norm += (uint32_t)(w * w);
    }
  } else { /* sq8.mojo:30-37 */
    for (int64_t i = 0; i < d; i++) {
      float s = roundf((x[i] - minimum) * inverse_scale);
      s = s < 0.0f ? 0.0f : (s > 255.0f ? 255.0f : s);
      uint8_t c = (uint8_t)s;
      code[i] = c;
      norm += (uint32_t)c * (uint32_t)c;
    }
  }
  return norm;
}

struct mvref_sq8 {
  int d, metric;
  int64_t ntotal, cap;
  float *f32;
  uint8_t *codes;
  uint32_t *norms;
  float gmin, gmax, scale;
};

mvref_sq8 *mvref_sq8_create(int d, int metric) {
  mvref_sq8 *h = (mvref_sq8 *)calloc(1, sizeof(mvref_sq8));
  h->d = d;
  h->metric = metric;
  h->gmin = FLT_MAX; /* index_flat_sq8.mojo:234-236 */
  h->gmax = -FLT_MAX;
  h->scale = 1.0f;
  return h;
}

void mvref_sq8_destroy(mvref_sq8 *h) {
  if (!h) return;
  free(h->f32); free(h->codes); free(h->norms); free(h);
}

/* IndexFlatSQ8.add: index_flat_sq8.mojo:286-397. */
int mvref_sq8_add(mvref_sq8 *h, const float *x, int64_t n) {
  if (n == 0) return 0;
  int d = h->d;
  if (h->ntotal + n > h->cap) {
    int64_t nc = h->cap * 2 > h->ntotal + n ? h->cap * 2 : h->ntotal + n;
    h->f32 = (float *)realloc(h->f32, sizeof(float) * (size_t)nc * d);
    h->codes = (uint8_t *)realloc(h->codes, (size_t)nc * d);
    h->norms = (uint32_t *)realloc(h->norms, sizeof(uint32_t) * (size_t)nc);
    h->cap = nc;
  }
  float bmin = FLT_MAX, bmax = -FLT_MAX; /* :317-331 */
  for (int64_t i = 0; i < n * d; i++) {
    float v = x[i];
    if (v < bmin) bmin = v;
    if (v > bmax) bmax = v;
  }
  int requant = 0;
  if (bmin < h->gmin) { h->gmin = bmin; requant = 1; }
  if (bmax > h->gmax) { h->gmax = bmax; requant = 1; }
  if (h->metric == MVREF_L2) { /* :333-342 */
    h->scale = h->gmax == h->gmin ? 1.0f : (h->gmax - h->gmin) / 255.0f;
  } else {
    float a = fabsf(h->gmin), b = fabsf(h->gmax);
    float mx = a > b ? a : b;
    h->scale = mx > 0.0f ? mx / 127.0f : 1.0f;
  }
  float inv = 1.0f / h->scale;
  int sym = h->metric != MVREF_L2;
  float mn = sym ? 0.0f : h->gmin;
  if (requant && h->ntotal > 0) /* :346-367 */
    for (int64_t i = 0; i < h->ntotal; i++)
      h->norms[i] = mvref_sq8_encode_vec(h->f32 + i * d, d, sym, mn, inv, h->codes + i * d);
  memcpy(h->f32 + h->ntotal * d, x, sizeof(float) * (size_t)n * d);
  for (int64_t i = 0; i < n; i++) /* :369-395 */
    h->norms[h->ntotal + i] =
        mvref_sq8_encode_vec(x + i * d, d, sym, mn, inv, h->codes + (h->ntotal + i) * d);
  h->ntotal += n;
  return 0;
}

void mvref_sq8_params(const mvref_sq8 *h, float *gmin, float *gmax,
                      float *scale, int64_t *ntotal) {
  *gmin = h->gmin; *gmax = h->gmax; *scale = h->scale; *ntotal = h->ntotal;
}

void mvref_sq8_codes(const mvref_sq8 *h, uint8_t *codes, uint32_t *norms) {
  memcpy(codes, h->codes, (size_t)h->ntotal * h->d);
  memcpy(norms, h->norms, sizeof(uint32_t) * (size_t)h->ntotal);
}

/* IndexFlatSQ8._search_impl (index_flat_sq8.mojo:428-526), canonical ties;
 * with k_factor > 1 the k*k_factor code-distance candidates are reranked with
 * exact FP32 distances as the SQ8 rerank in index_hnsw.mojo:433-496 does
 * (l2_distance_simd[64] / inner_product_simd[64]). Pad (0.0, -1). */
void mvref_sq8_search(const mvref_sq8 *h, const float *q, int64_t nq, int k,
                      int k_factor, const uint8_t *exclude,
                      int64_t exclude_len, float *D, int64_t *I,
                      int nthreads) {
  int d = h->d, ip = h->metric != MVREF_L2;
  int kk = k * (k_factor > 1 ? k_factor : 1);
  float inv = 1.0f / h->scale, ssq = h->scale * h->scale;
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
  {
    cand_t *buf = (cand_t *)malloc(sizeof(cand_t) * (size_t)kk);
    cand_t *buf2 = (cand_t *)malloc(sizeof(cand_t) * (size_t)k);
    uint8_t *qc = (uint8_t *)malloc((size_t)d);
#pragma omp for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < nq; qi++) {
      const float *qp = q + qi * d;
      uint32_t qn = mvref_sq8_encode_vec(qp, d, ip, ip ? 0.0f : h->gmin, inv, qc);
      cheap_t c;
      ch_init(&c, buf, kk);
      for (int64_t j = 0; j < h->ntotal; j++) {
        if (exclude && j < exclude_len && exclude[j] > 0) continue;
        float dd;
        if (ip) {
          int32_t dot = mvref_dot_i8((const int8_t *)qc, (const int8_t *)(h->codes + j * d), d);
          dd = -(float)dot * ssq;
        } else {
          uint32_t dot = mvref_dot_u8(qc, h->codes + j * d, d);
          dd = (float)mvref_l2_from_dot(qn, h->norms[j], dot) * ssq;
        }
        ch_offer(&c, dd, j);
      }
      ch_sorted(&c);
      cheap_t *res = &c;
      cheap_t r2;
      if (k_factor > 1) {
        ch_init(&r2, buf2, k);
        for (int j = 0; j < c.size; j++) {
          const float *xv = h->f32 + c.h[j].id * d;
          float ex = ip ? -mvref_ip_simd(qp, xv, d, 64) : mvref_l2_simd(qp, xv, d, 64);
          ch_offer(&r2, ex, c.h[j].id);
        }
        ch_sorted(&r2);
        res = &r2;
      }
      for (int j = 0; j < k; j++) {
        if (j < res->size) {
          float v = res->h[j].d;
          D[qi * k + j] = ip ? -v : v;
          I[qi * k + j] = res->h[j].id;
        } else {
          D[qi * k + j] = 0.0f;
          I[qi * k + j] = -1;
        }
      }
    }
    free(buf); free(buf2); free(qc);
  }
}

/* SQ8DistanceComputer (index_flat_sq8.mojo:21-190) over a batch: the query
 * is encoded as get_distance_computer does (:527-560), then distance() /
 * distance_batch4() (:80-146, the same values: code dots are exact):
 * IP -Float32(dot_i8) * scale^2, L2 Float32(l2_from_dot) * scale^2
 * (sq8.mojo:62-79). */
void mvref_sq8_distances(const mvref_sq8 *h, const float *q, int64_t nq, const int64_t *ids,
                         int64_t per_query, float *out) {
  int d = h->d, ip = h->metric != MVREF_L2;
  float inv = 1.0f / h->scale, ssq = h->scale * h->scale;
  uint8_t *qc = (uint8_t *)malloc((size_t)d);
  for (int64_t qi = 0; qi < nq; qi++) {
    uint32_t qn = mvref_sq8_encode_vec(q + qi * d, d, ip, ip ? 0.0f : h->gmin, inv, qc);
    for (int64_t j = 0; j < per_query; j++) {
      const uint8_t *xc = h->codes + ids[qi * per_query + j] * d;
      float v;
      if (ip) {
        v = -(float)mvref_dot_i8((const int8_t *)qc, (const int8_t *)xc, d) * ssq;
      } else {
        uint32_t xn = h->norms[ids[qi * per_query + j]];
        v = (float)mvref_l2_from_dot(qn, xn, mvref_dot_u8(qc, xc, d)) * ssq;
      }
      out[qi * per_query + j] = v;
    }
  }
  free(qc);
}

/* SQ8DistanceComputer.symmetric_distance (index_flat_sq8.mojo:148-177). */
void mvref_sq8_pair_distances(const mvref_sq8 *h, const int64_t *a, const int64_t *b,
                              int64_t npairs, float *out) {
  int d = h->d, ip = h->metric != MVREF_L2;
  float ssq = h->scale * h->scale;
  for (int64_t p = 0; p < npairs; p++) {
    const uint8_t *ca = h->codes + a[p] * d, *cb = h->codes + b[p] * d;
    if (ip)
      out[p] = -(float)mvref_dot_i8((const int8_t *)ca, (const int8_t *)cb, d) * ssq;
    else
      out[p] = (float)mvref_l2_from_dot(h->norms[a[p]], h->norms[b[p]], mvref_dot_u8(ca, cb, d)) * ssq;
  }
}

/* ------------------------------------------------------------------------ */
/* IndexIVFPQ: mojovec/index/index_ivf_pq.mojo. */

struct mvref_ivfpq {
  int d, nlist, M, metric, trained;
  float *coarse;  /* nlist x d */
  float *pq;      /* M x 256 x dsub */
  int64_t ntotal;
  int64_t *lsize, *lcap;
  int64_t **lids;
  uint8_t **lcodes;
  float *vectors; /* FP32 rows by id (vcap rows), kept for the exact rerank */
  int64_t vcap;
};

mvref_ivfpq *mvref_ivfpq_create(int d, int nlist, int M, int metric) {
  mvref_ivfpq *h = (mvref_ivfpq *)calloc(1, sizeof(mvref_ivfpq));
  h->d = d; h->nlist = nlist; h->M = M; h->metric = metric;
  h->coarse = (float *)calloc((size_t)nlist * d, sizeof(float));
  h->pq = (float *)calloc((size_t)d * PQ_KSUB, sizeof(float));
  h->lsize = (int64_t *)calloc((size_t)nlist, sizeof(int64_t));
  h->lcap = (int64_t *)calloc((size_t)nlist, sizeof(int64_t));
  h->lids = (int64_t **)calloc((size_t)nlist, sizeof(int64_t *));
  h->lcodes = (uint8_t **)calloc((size_t)nlist, sizeof(uint8_t *));
  return h;
}

void mvref_ivfpq_destroy(mvref_ivfpq *h) {
  if (!h) return;
  for (int l = 0; l < h->nlist; l++) { free(h->lids[l]); free(h->lcodes[l]); }
  free(h->lids); free(h->lcodes); free(h->lsize); free(h->lcap);
  free(h->coarse); free(h->pq); free(h->vectors); free(h);
}

void mvref_ivfpq_set_codebooks(mvref_ivfpq *h, const float *coarse,
                               const float *pq) {
  memcpy(h->coarse, coarse, sizeof(float) * (size_t)h->nlist * h->d);
  memcpy(h->pq, pq, sizeof(float) * (size_t)h->d * PQ_KSUB);
  h->trained = 1;
}

void mvref_ivfpq_get_codebooks(const mvref_ivfpq *h, float *coarse, float *pq) {
  if (coarse) memcpy(coarse, h->coarse, sizeof(float) * (size_t)h->nlist * h->d);
  if (pq) memcpy(pq, h->pq, sizeof(float) * (size_t)h->d * PQ_KSUB);
}

/* Coarse assignment = quantizer.search(x, 1) (index_ivf_pq.mojo:115,172;
 * index_ivf_flat.mojo:115): the canonical nearest centroid. */
static void coarse_assign(const float *coarse, int nlist, int d, int metric,
                          const float *x, int64_t n, int32_t *labels, int nthreads) {
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
  for (int64_t i = 0; i < n; i++) {
    float best = 0.0f;
    int bl = -1;
    for (int l = 0; l < nlist; l++) {
      const float *c = coarse + (int64_t)l * d;
      float dd = metric == MVREF_L2 ? mvref_l2_simd(x + i * d, c, d, 8)
                                    : -mvref_ip_simd(x + i * d, c, d, 8);
      if (bl < 0 || dd < best) { best = dd; bl = l; }
    }
    labels[i] = bl;
  }
}

static void ivf_assign(const mvref_ivfpq *h, const float *x, int64_t n,
                       int32_t *labels, int nthreads) {
  coarse_assign(h->coarse, h->nlist, h->d, h->metric, x, n, labels, nthreads);
}

/* train: index_ivf_pq.mojo:87-136 (coarse k-means, residuals, PQ). */
void mvref_ivfpq_train(mvref_ivfpq *h, const float *x, int64_t n,
                       int niter_coarse, int niter_pq, uint64_t seed,
                       int nthreads) {
  if (h->trained) return;
  int d = h->d;
  float *init = (float *)malloc(sizeof(float) * (size_t)h->nlist * d);
  int64_t off = mvref_kmeans_init(x, n, d, h->nlist, seed, init);
  mvref_kmeans_train(x, n, d, h->nlist, niter_coarse, init, off, 1, h->coarse, NULL, nthreads);
  free(init);
  int32_t *lab = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  ivf_assign(h, x, n, lab, nthreads);
  float *res = (float *)malloc(sizeof(float) * (size_t)n * d);
  for (int64_t i = 0; i < n; i++) {
    int l = lab[i] < 0 || lab[i] >= h->nlist ? 0 : lab[i];
    for (int j = 0; j < d; j++) res[i * d + j] = x[i * d + j] - h->coarse[(int64_t)l * d + j];
  }
  mvref_pq_train(res, n, d, h->M, niter_pq, seed, 1, h->pq, nthreads);
  free(res); free(lab);
  h->trained = 1;
}

/* add / add_with_ids: index_ivf_pq.mojo:138-213. ids NULL: ntotal+i (add,
 * :146-147); otherwise the caller's ids, appended in input order. The kept
 * FP32 rows (the k_factor rerank's input, an addition of this build) are
 * stored by id, so the rerank reads the row added under a candidate's id. */
void mvref_ivfpq_add_ids(mvref_ivfpq *h, const float *x, int64_t n, const int64_t *ids,
                         int keep_vectors, int nthreads) {
  if (!h->trained || n == 0) return; /* :161 */
  int d = h->d, M = h->M;
  int32_t *lab = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  ivf_assign(h, x, n, lab, nthreads);
  float *res = (float *)malloc(sizeof(float) * (size_t)n * d);
  for (int64_t i = 0; i < n; i++) { /* :179-192 residual vs the assigned centroid */
    int l = lab[i] < 0 || lab[i] >= h->nlist ? 0 : lab[i];
    for (int j = 0; j < d; j++) res[i * d + j] = x[i * d + j] - h->coarse[(int64_t)l * d + j];
  }
  uint8_t *codes = (uint8_t *)malloc((size_t)n * M);
  mvref_pq_encode(res, n, d, M, h->pq, codes, nthreads);
  for (int64_t i = 0; i < n; i++) { /* :199-211 */
    int l = lab[i];
    if (l < 0 || l >= h->nlist) continue;
    if (h->lsize[l] == h->lcap[l]) {
      int64_t nc = h->lcap[l] ? h->lcap[l] * 2 : 16;
      h->lids[l] = (int64_t *)realloc(h->lids[l], sizeof(int64_t) * (size_t)nc);
      h->lcodes[l] = (uint8_t *)realloc(h->lcodes[l], (size_t)nc * M);
      h->lcap[l] = nc;
    }
    h->lids[l][h->lsize[l]] = ids ? ids[i] : h->ntotal + i;
    memcpy(h->lcodes[l] + h->lsize[l] * M, codes + i * M, (size_t)M);
    h->lsize[l]++;
  }
  if (keep_vectors) {
    int64_t top = h->ntotal + n; /* rows needed: one past the largest id */
    if (ids)
      for (int64_t i = 0; i < n; i++)
        if (ids[i] + 1 > top) top = ids[i] + 1;
    if (top > h->vcap) {
      int64_t nc = h->vcap * 2 > top ? h->vcap * 2 : top;
      h->vectors = (float *)realloc(h->vectors, sizeof(float) * (size_t)nc * d);
      memset(h->vectors + h->vcap * d, 0, sizeof(float) * (size_t)(nc - h->vcap) * d);
      h->vcap = nc;
    }
    for (int64_t i = 0; i < n; i++) {
      int64_t id = ids ? ids[i] : h->ntotal + i;
      memcpy(h->vectors + id * d, x + i * d, sizeof(float) * (size_t)d);
    }
  }
  h->ntotal += n; /* :213 */
  free(codes); free(res); free(lab);
}

void mvref_ivfpq_add(mvref_ivfpq *h, const float *x, int64_t n,
                     int keep_vectors, int nthreads) {
  mvref_ivfpq_add_ids(h, x, n, NULL, keep_vectors, nthreads);
}

/* Installs inverted lists built elsewhere (CSR: list_off[nlist+1], ids and
 * M-byte codes in list order) and optionally the FP32 rows for the rerank, so
 * the CPU baseline times the search alone. */
void mvref_ivfpq_load_lists(mvref_ivfpq *h, const int64_t *list_off, const int32_t *ids,
                            const uint8_t *codes, const float *vectors, int64_t ntotal) {
  int M = h->M;
  for (int l = 0; l < h->nlist; l++) {
    int64_t L = list_off[l + 1] - list_off[l];
    free(h->lids[l]); free(h->lcodes[l]);
    h->lids[l] = (int64_t *)malloc(sizeof(int64_t) * (size_t)(L > 0 ? L : 1));
    h->lcodes[l] = (uint8_t *)malloc((size_t)(L > 0 ? L : 1) * M);
    for (int64_t j = 0; j < L; j++) h->lids[l][j] = ids[list_off[l] + j];
    memcpy(h->lcodes[l], codes + list_off[l] * M, (size_t)L * M);
    h->lsize[l] = h->lcap[l] = L;
  }
  free(h->vectors);
  h->vectors = NULL;
  if (vectors) {
    h->vectors = (float *)malloc(sizeof(float) * (size_t)ntotal * h->d);
    memcpy(h->vectors, vectors, sizeof(float) * (size_t)ntotal * h->d);
  }
  h->vcap = ntotal;
  h->ntotal = ntotal;
  h->trained = 1;
}

int64_t mvref_ivfpq_list_size(const mvref_ivfpq *h, int list) { return h->lsize[list]; }

void mvref_ivfpq_list(const mvref_ivfpq *h, int list, int64_t *ids, uint8_t *codes) {
  memcpy(ids, h->lids[list], sizeof(int64_t) * (size_t)h->lsize[list]);
  memcpy(codes, h->lcodes[list], (size_t)h->lsize[list] * h->M);
}

/* Coarse top-np of one query: IndexFlat search over the centroids (:264). */
static void ivfpq_probe(const mvref_ivfpq *h, const float *qp, int np, cheap_t *pc, cand_t *pbuf) {
  int d = h->d, ip = h->metric != MVREF_L2;
  ch_init(pc, pbuf, np);
  for (int l = 0; l < h->nlist; l++) {
    const float *c = h->coarse + (int64_t)l * d;
    ch_offer(pc, ip ? -mvref_ip_simd(qp, c, d, 8) : mvref_l2_simd(qp, c, d, 8), l);
  }
  ch_sorted(pc);
}

/* Per-query part of the search after the coarse step: LUT, ADC scan of the
 * probed lists, optional exact rerank, output row (:276-360). */
static void ivfpq_scan_query(const mvref_ivfpq *h, const float *qp, const cheap_t *pc, int k,
                             int kf, const uint8_t *exclude, int64_t exclude_len, cand_t *buf,
                             cand_t *buf2, float *lut, float *r, float *Drow, int64_t *Irow) {
  int d = h->d, M = h->M, ip = h->metric != MVREF_L2;
  int kk = k * kf;
  if (ip) mvref_pq_lut(qp, d, M, h->pq, MVREF_IP, lut); /* :288-293 */
  cheap_t c;
  ch_init(&c, buf, kk);
  for (int p = 0; p < pc->size; p++) {
    int l = (int)pc->h[p].id;
    int64_t L = h->lsize[l];
    if (L == 0) continue;
    const float *cen = h->coarse + (int64_t)l * d;
    float coarse_d = 0.0f;
    if (!ip) { /* :306-315 residual LUT */
      for (int j = 0; j < d; j++) r[j] = qp[j] - cen[j];
      mvref_pq_lut(r, d, M, h->pq, MVREF_L2, lut);
    } else { /* :316-324 */
      coarse_d = -mvref_ip_simd(qp, cen, d, 8);
    }
    const uint8_t *codes = h->lcodes[l];
    const int64_t *ids = h->lids[l];
    for (int64_t j = 0; j < L; j++) { /* :326-347 ADC scan */
      int64_t id = ids[j];
      if (exclude && id >= 0 && id < exclude_len && exclude[id] != 0) continue;
      const uint8_t *cp = codes + j * M;
      float dist = coarse_d;
      for (int m = 0; m < M; m++) dist += lut[m * PQ_KSUB + cp[m]];
      ch_offer(&c, dist, id);
    }
  }
  ch_sorted(&c);
  cheap_t *res = &c;
  cheap_t r2;
  if (kf > 1 && h->vectors) {
    ch_init(&r2, buf2, k);
    for (int j = 0; j < c.size; j++) {
      const float *xv = h->vectors + c.h[j].id * d;
      float ex = ip ? -mvref_ip_simd(qp, xv, d, 64) : mvref_l2_simd(qp, xv, d, 64);
      ch_offer(&r2, ex, c.h[j].id);
    }
    ch_sorted(&r2);
    res = &r2;
  }
  for (int j = 0; j < k; j++) {
    if (j < res->size) {
      float v = res->h[j].d;
      Drow[j] = ip ? -v : v;
      Irow[j] = res->h[j].id;
    } else {
      Drow[j] = IVF_PAD_DIST;
      Irow[j] = -1;
    }
  }
}

static int ivfpq_np(const mvref_ivfpq *h, int nprobe, const int32_t *npq, int64_t qi) {
  int np = npq ? npq[qi] : nprobe;
  if (np > h->nlist) np = h->nlist; /* :253-254 */
  return np < 1 ? 1 : np;
}

/* search: index_ivf_pq.mojo:225-360 with canonical ties, optional per-query
 * nprobe and the k_factor exact rerank (l2_distance_simd[64] /
 * inner_product_simd[64], as the SQ8 rerank of index_hnsw.mojo:449-482).
 * Pad (1e38, -1); IP distances un-negated over found entries (:357-360).
 * Parallel over queries (coarse step and scan). */
void mvref_ivfpq_search(const mvref_ivfpq *h, const float *q, int64_t nq,
                        int k, int nprobe, const int32_t *nprobe_per_query,
                        int k_factor, const uint8_t *exclude,
                        int64_t exclude_len, float *D, int64_t *I,
                        int nthreads) {
  int d = h->d, M = h->M, nlist = h->nlist;
  int kf = k_factor > 1 ? k_factor : 1;
  int kk = k * kf;
  if (!h->trained || h->ntotal == 0) { /* :247-251 */
    for (int64_t i = 0; i < nq * k; i++) { D[i] = IVF_PAD_DIST; I[i] = -1; }
    return;
  }
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
  {
    cand_t *pbuf = (cand_t *)malloc(sizeof(cand_t) * (size_t)nlist);
    cand_t *buf = (cand_t *)malloc(sizeof(cand_t) * (size_t)kk);
    cand_t *buf2 = (cand_t *)malloc(sizeof(cand_t) * (size_t)k);
    float *lut = (float *)malloc(sizeof(float) * (size_t)M * PQ_KSUB);
    float *r = (float *)malloc(sizeof(float) * (size_t)d);
#pragma omp for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < nq; qi++) {
      cheap_t pc;
      ivfpq_probe(h, q + qi * d, ivfpq_np(h, nprobe, nprobe_per_query, qi), &pc, pbuf);
      ivfpq_scan_query(h, q + qi * d, &pc, k, kf, exclude, exclude_len, buf, buf2, lut, r,
                       D + qi * k, I + qi * k);
    }
    free(pbuf); free(buf); free(buf2); free(lut); free(r);
  }
}

/* The same search with the reference's threading (index_ivf_pq.mojo:
 * 256-276): the coarse IndexFlat search runs parallel over queries (as
 * IndexFlat._search_impl's parallelize), then the LUT + ADC loop over the
 * queries runs serially on one thread (`for i in range(n)`, :276). Same
 * results as mvref_ivfpq_search; a CPU-baseline timing variant. */
void mvref_ivfpq_search_refstruct(const mvref_ivfpq *h, const float *q, int64_t nq, int k,
                                  int nprobe, int k_factor, float *D, int64_t *I, int nthreads) {
  int d = h->d, M = h->M, nlist = h->nlist;
  int kf = k_factor > 1 ? k_factor : 1;
  int kk = k * kf;
  if (!h->trained || h->ntotal == 0) {
    for (int64_t i = 0; i < nq * k; i++) { D[i] = IVF_PAD_DIST; I[i] = -1; }
    return;
  }
  int np = ivfpq_np(h, nprobe, NULL, 0);
  cand_t *probes = (cand_t *)malloc(sizeof(cand_t) * (size_t)nq * np);
  int *counts = (int *)malloc(sizeof(int) * (size_t)nq);
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
  {
    cand_t *pbuf = (cand_t *)malloc(sizeof(cand_t) * (size_t)nlist);
#pragma omp for schedule(dynamic, 16)
    for (int64_t qi = 0; qi < nq; qi++) {
      cheap_t pc;
      ivfpq_probe(h, q + qi * d, np, &pc, pbuf);
      memcpy(probes + qi * np, pc.h, sizeof(cand_t) * (size_t)pc.size);
      counts[qi] = pc.size;
    }
    free(pbuf);
  }
  cand_t *buf = (cand_t *)malloc(sizeof(cand_t) * (size_t)kk);
  cand_t *buf2 = (cand_t *)malloc(sizeof(cand_t) * (size_t)k);
  float *lut = (float *)malloc(sizeof(float) * (size_t)M * PQ_KSUB);
  float *r = (float *)malloc(sizeof(float) * (size_t)d);
  for (int64_t qi = 0; qi < nq; qi++) {
    cheap_t pc;
    pc.h = probes + qi * np;
    pc.size = counts[qi];
    pc.cap = np;
    ivfpq_scan_query(h, q + qi * d, &pc, k, kf, NULL, 0, buf, buf2, lut, r, D + qi * k, I + qi * k);
  }
  free(buf); free(buf2); free(lut); free(r); free(probes); free(counts);
}

/* ------------------------------------------------------------------------ */
/* IndexIVFFlat: mojovec/index/index_ivf_flat.mojo. Lists hold the raw FP32
 * rows (code_size = 4d, :52) and their ids in append order. */

struct mvref_ivfflat {
  int d, nlist, metric, qmetric, trained; /* qmetric: the quantizer IndexFlat's */
  float *coarse; /* nlist x d */
  int64_t ntotal;
  int64_t *lsize, *lcap;
  int64_t **lids;
  float **lrows;
};

/* IndexIVFFlat[IndexFlat](IndexFlat(d, quantizer_metric), d, nlist, metric)
 * (:29-52): the quantizer keeps its own metric. */
mvref_ivfflat *mvref_ivfflat_create(int d, int nlist, int metric, int quantizer_metric) {
  mvref_ivfflat *h = (mvref_ivfflat *)calloc(1, sizeof(mvref_ivfflat));
  h->d = d; h->nlist = nlist; h->metric = metric; h->qmetric = quantizer_metric;
  h->coarse = (float *)calloc((size_t)nlist * d, sizeof(float));
  h->lsize = (int64_t *)calloc((size_t)nlist, sizeof(int64_t));
  h->lcap = (int64_t *)calloc((size_t)nlist, sizeof(int64_t));
  h->lids = (int64_t **)calloc((size_t)nlist, sizeof(int64_t *));
  h->lrows = (float **)calloc((size_t)nlist, sizeof(float *));
  return h;
}

void mvref_ivfflat_destroy(mvref_ivfflat *h) {
  if (!h) return;
  for (int l = 0; l < h->nlist; l++) { free(h->lids[l]); free(h->lrows[l]); }
  free(h->lids); free(h->lrows); free(h->lsize); free(h->lcap); free(h->coarse); free(h);
}

/* quantizer.add(centroids) + is_trained = True (:75-77). */
void mvref_ivfflat_set_centroids(mvref_ivfflat *h, const float *coarse) {
  memcpy(h->coarse, coarse, sizeof(float) * (size_t)h->nlist * h->d);
  h->trained = 1;
}

void mvref_ivfflat_get_centroids(const mvref_ivfflat *h, float *coarse) {
  memcpy(coarse, h->coarse, sizeof(float) * (size_t)h->nlist * h->d);
}

/* train: KMeans(d, nlist, 15).train(x) (:65-77). */
void mvref_ivfflat_train(mvref_ivfflat *h, const float *x, int64_t n, int niter,
                         uint64_t seed, int nthreads) {
  if (h->trained) return; /* :71 */
  int d = h->d;
  float *init = (float *)malloc(sizeof(float) * (size_t)h->nlist * d);
  int64_t off = mvref_kmeans_init(x, n, d, h->nlist, seed, init);
  mvref_kmeans_train(x, n, d, h->nlist, niter, init, off, 1, h->coarse, NULL, nthreads);
  free(init);
  h->trained = 1;
}

/* add / add_with_ids (:79-137): nearest centroid, one append per row in
 * input order (ArrayInvertedLists capacity: 16, then doubling); ids default
 * to ntotal + i; rows of an untrained index are dropped silently (:102-104). */
void mvref_ivfflat_add(mvref_ivfflat *h, const float *x, int64_t n, const int64_t *ids,
                       int nthreads) {
  if (!h->trained || n == 0) return;
  int d = h->d;
  int32_t *lab = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  coarse_assign(h->coarse, h->nlist, d, h->qmetric, x, n, lab, nthreads);
  for (int64_t i = 0; i < n; i++) {
    int l = lab[i];
    if (l < 0 || l >= h->nlist) continue; /* :125 */
    if (h->lsize[l] == h->lcap[l]) {
      int64_t nc = h->lcap[l] ? h->lcap[l] * 2 : 16;
      h->lids[l] = (int64_t *)realloc(h->lids[l], sizeof(int64_t) * (size_t)nc);
      h->lrows[l] = (float *)realloc(h->lrows[l], sizeof(float) * (size_t)nc * d);
      h->lcap[l] = nc;
    }
    h->lids[l][h->lsize[l]] = ids ? ids[i] : h->ntotal + i;
    memcpy(h->lrows[l] + h->lsize[l] * d, x + i * d, sizeof(float) * (size_t)d);
    h->lsize[l]++;
  }
  h->ntotal += n; /* :137 */
  free(lab);
}

/* Installs lists built elsewhere (CSR: list_off[nlist+1], int32 ids and FP32
 * rows in list order), so the CPU baseline times the search alone. */
void mvref_ivfflat_load_lists(mvref_ivfflat *h, const int64_t *list_off, const int32_t *ids,
                              const float *rows, int64_t ntotal) {
  int d = h->d;
  for (int l = 0; l < h->nlist; l++) {
    int64_t L = list_off[l + 1] - list_off[l];
    free(h->lids[l]); free(h->lrows[l]);
    h->lids[l] = (int64_t *)malloc(sizeof(int64_t) * (size_t)(L > 0 ? L : 1));
    h->lrows[l] = (float *)malloc(sizeof(float) * (size_t)(L > 0 ? L : 1) * d);
    for (int64_t j = 0; j < L; j++) h->lids[l][j] = ids[list_off[l] + j];
    memcpy(h->lrows[l], rows + list_off[l] * d, sizeof(float) * (size_t)L * d);
    h->lsize[l] = h->lcap[l] = L;
  }
  h->ntotal = ntotal;
  h->trained = 1;
}

int64_t mvref_ivfflat_list_size(const mvref_ivfflat *h, int list) { return h->lsize[list]; }

void mvref_ivfflat_list(const mvref_ivfflat *h, int list, int64_t *ids, float *rows) {
  memcpy(ids, h->lids[list], sizeof(int64_t) * (size_t)h->lsize[list]);
  memcpy(rows, h->lrows[list], sizeof(float) * (size_t)h->lsize[list] * h->d);
}

/* search (:139-244). Untrained or empty: (1e38, -1) everywhere (:171-175).
 * nprobe clamped to nlist (:177-178); coarse = IndexFlat top-nprobe (:189);
 * per probed list every row is scored with l2_distance_simd[8] or
 * -inner_product_simd[4] (:222-225) into a top-k kept in the canonical
 * (distance, lower id) order; the tail stays (1e38, -1) (:198-202) and IP
 * distances are un-negated over the found entries (:241-243). The reference
 * accepts a filter and ignores it; here a non-null `exclude` drops ids with
 * exclude[id] != 0, as the IVF-PQ search does (index_ivf_pq.mojo:328-334). */
void mvref_ivfflat_search(const mvref_ivfflat *h, const float *q, int64_t nq, int k,
                          int nprobe, const uint8_t *exclude, int64_t exclude_len,
                          float *D, int64_t *I, int nthreads) {
  int d = h->d, nlist = h->nlist, ip = h->metric != MVREF_L2, qip = h->qmetric != MVREF_L2;
  if (!h->trained || h->ntotal == 0) {
    for (int64_t i = 0; i < nq * k; i++) { D[i] = IVF_PAD_DIST; I[i] = -1; }
    return;
  }
  int np = nprobe > nlist ? nlist : nprobe;
  if (np < 1) np = 1;
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
  {
    cand_t *pbuf = (cand_t *)malloc(sizeof(cand_t) * (size_t)np);
    cand_t *buf = (cand_t *)malloc(sizeof(cand_t) * (size_t)k);
#pragma omp for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < nq; qi++) {
      const float *qp = q + qi * d;
      cheap_t pc;
      ch_init(&pc, pbuf, np);
      for (int l = 0; l < nlist; l++) {
        const float *c = h->coarse + (int64_t)l * d;
        ch_offer(&pc, qip ? -mvref_ip_simd(qp, c, d, 8) : mvref_l2_simd(qp, c, d, 8), l);
      }
      ch_sorted(&pc);
      cheap_t c;
      ch_init(&c, buf, k);
      for (int p = 0; p < pc.size; p++) {
        int l = (int)pc.h[p].id;
        const float *rows = h->lrows[l];
        const int64_t *ids = h->lids[l];
        for (int64_t j = 0; j < h->lsize[l]; j++) {
          int64_t id = ids[j];
          if (exclude && id >= 0 && id < exclude_len && exclude[id] != 0) continue;
          const float *xv = rows + j * d;
          ch_offer(&c, ip ? -mvref_ip_simd(qp, xv, d, 4) : mvref_l2_simd(qp, xv, d, 8), id);
        }
      }
      ch_sorted(&c);
      for (int j = 0; j < k; j++) {
        if (j < c.size) {
          float v = c.h[j].d;
          D[qi * k + j] = ip ? -v : v;
          I[qi * k + j] = c.h[j].id;
        } else {
          D[qi * k + j] = IVF_PAD_DIST;
          I[qi * k + j] = -1;
        }
      }
    }
    free(pbuf); free(buf);
  }
}
