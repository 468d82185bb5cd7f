/*
 * mvref — CPU oracle for the MojoVec batched vector-scan path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithms (MojoVec v0.8.2, Mojo sources) used as the checker for the HIP
 * library and as the timed CPU baseline. Nothing in the product library or the
 * Python package may link, load or call it.
 *
 * Every routine cites the reference file:line it restates. Where the Mojo
 * standard library fixes an arithmetic detail that is not vendored with the
 * reference (SIMD.reduce_add lane order, math.round ties, the Philox stream),
 * the chosen restatement is named at the routine; those details are "parity
 * unpinned" against the reference and pinned only against this restatement.
 */
#ifndef MVREF_H
#define MVREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MVREF_L2 = 0, MVREF_IP = 1 };

/* ---- distances (mojovec/utils/distances.mojo) ---- */
float mvref_l2_simd(const float *x, const float *y, int64_t d, int W);
float mvref_ip_simd(const float *x, const float *y, int64_t d, int W);
/* l2_distance_simd_batch4[W] / inner_product_simd_batch4[W] for one row
 * (distances.mojo:541-660), W <= 64. */
float mvref_l2_simd_b4(const float *x, const float *y, int64_t d, int W);
float mvref_ip_simd_b4(const float *x, const float *y, int64_t d, int W);
/* FlatDistanceComputer.distance (order 0) / distance_batch4 (order 1) over
 * nq x per_query (query, row id) pairs, and symmetric_distance over row
 * pairs (index_flat.mojo:36-115). */
void mvref_flat_distances(const float *base, int d, int metric, const float *q,
                          int64_t nq, const int64_t *ids, int64_t per_query,
                          int order, float *out);
void mvref_flat_pair_distances(const float *base, int d, int metric, const int64_t *a,
                               const int64_t *b, int64_t npairs, float *out);
float mvref_l2_short4(const float *x, const float *y, int64_t d);
float mvref_ip_short4(const float *x, const float *y, int64_t d);
uint32_t mvref_dot_u8(const uint8_t *x, const uint8_t *y, int64_t d);
int32_t mvref_dot_i8(const int8_t *x, const int8_t *y, int64_t d);
uint32_t mvref_l2_from_dot(uint32_t na, uint32_t nb, uint32_t dot);

/* batched helpers for tests: out[i] = f(x_i, y_i) */
void mvref_l2_simd_batch(const float *x, const float *y, int64_t n, int64_t d,
                         int W, float *out);
void mvref_ip_simd_batch(const float *x, const float *y, int64_t n, int64_t d,
                         int W, float *out);

/* ---- heaps (mojovec/utils/heap.mojo) ---- */
void mvref_heap_topk_reference(const float *dist, int64_t n, int k,
                               float *out_d, int64_t *out_i, int *found);
void mvref_topk_canonical(const float *dist, const int64_t *ids, int64_t n,
                          int k, float *out_d, int64_t *out_i, int *found);

/* ---- Philox init + k-means (mojovec/clustering/kmeans.mojo) ---- */
void mvref_philox4x32_10(uint64_t seed, uint64_t counter, uint32_t out[4]);
int64_t mvref_kmeans_init(const float *x, int64_t n, int d, int k,
                          uint64_t seed, float *centroids);
/* fp64_sums=1: the build's FP64 chunk accumulation; 0: the reference FP32. */
void mvref_kmeans_train(const float *x, int64_t n, int d, int k, int niter,
                        const float *init, int64_t sample_offset,
                        int fp64_sums, float *centroids, int32_t *assign,
                        int nthreads);

/* ---- product quantizer (mojovec/quantization/pq.mojo) ---- */
void mvref_pq_train(const float *x, int64_t n, int d, int M, int niter,
                    uint64_t seed, int fp64_sums, float *codebooks,
                    int nthreads);
void mvref_pq_encode(const float *x, int64_t n, int d, int M,
                     const float *codebooks, uint8_t *codes, int nthreads);
void mvref_pq_decode(const uint8_t *codes, int64_t n, int d, int M,
                     const float *codebooks, float *x);
void mvref_pq_lut(const float *q, int d, int M, const float *codebooks,
                  int metric, float *lut);

/* ---- exact Flat (mojovec/index/index_flat.mojo) ---- */
void mvref_flat_search(const float *base, int64_t n, int d, int metric,
                       const float *q, int64_t nq, int k,
                       const uint8_t *exclude, int64_t exclude_len,
                       float *D, int64_t *I, int nthreads);

/* ---- SQ8 (mojovec/quantization/sq8.mojo, index_flat_sq8.mojo) ---- */
uint32_t mvref_sq8_encode_vec(const float *x, int64_t d, int symmetric,
                              float minimum, float inverse_scale,
                              uint8_t *code);
typedef struct mvref_sq8 mvref_sq8;
mvref_sq8 *mvref_sq8_create(int d, int metric);
void mvref_sq8_destroy(mvref_sq8 *h);
int mvref_sq8_add(mvref_sq8 *h, const float *x, int64_t n);
void mvref_sq8_params(const mvref_sq8 *h, float *gmin, float *gmax,
                      float *scale, int64_t *ntotal);
void mvref_sq8_codes(const mvref_sq8 *h, uint8_t *codes, uint32_t *norms);
/* SQ8DistanceComputer.distance / symmetric_distance over a batch
 * (index_flat_sq8.mojo:21-190, :527-560). */
void mvref_sq8_distances(const mvref_sq8 *h, const float *q, int64_t nq, const int64_t *ids,
                         int64_t per_query, float *out);
void mvref_sq8_pair_distances(const mvref_sq8 *h, const int64_t *a, const int64_t *b,
                              int64_t npairs, float *out);
void mvref_sq8_search(const mvref_sq8 *h, const float *q, int64_t nq, int k,
                      int k_factor, const uint8_t *exclude,
                      int64_t exclude_len, float *D, int64_t *I,
                      int nthreads);

/* ---- IVF-PQ (mojovec/index/index_ivf_pq.mojo) ---- */
typedef struct mvref_ivfpq mvref_ivfpq;
mvref_ivfpq *mvref_ivfpq_create(int d, int nlist, int M, int metric);
void mvref_ivfpq_destroy(mvref_ivfpq *h);
void mvref_ivfpq_set_codebooks(mvref_ivfpq *h, const float *coarse,
                               const float *pq);
void mvref_ivfpq_train(mvref_ivfpq *h, const float *x, int64_t n,
                       int niter_coarse, int niter_pq, uint64_t seed,
                       int nthreads);
void mvref_ivfpq_get_codebooks(const mvref_ivfpq *h, float *coarse,
                               float *pq);
void mvref_ivfpq_add_ids(mvref_ivfpq *h, const float *x, int64_t n, const int64_t *ids,
                         int keep_vectors, int nthreads);
void mvref_ivfpq_add(mvref_ivfpq *h, const float *x, int64_t n,
                     int keep_vectors, int nthreads);
void mvref_ivfpq_load_lists(mvref_ivfpq *h, const int64_t *list_off, const int32_t *ids,
                            const uint8_t *codes, const float *vectors, int64_t ntotal);
int64_t mvref_ivfpq_list_size(const mvref_ivfpq *h, int list);
void mvref_ivfpq_list(const mvref_ivfpq *h, int list, int64_t *ids,
                      uint8_t *codes);
void mvref_ivfpq_search(const mvref_ivfpq *h, const float *q, int64_t nq,
                        int k, int nprobe, const int32_t *nprobe_per_query,
                        int k_factor, const uint8_t *exclude,
                        int64_t exclude_len, float *D, int64_t *I,
                        int nthreads);
/* The search with the reference's threading: coarse step parallel over
 * queries, LUT + ADC loop serial (a CPU-baseline timing variant). */
void mvref_ivfpq_search_refstruct(const mvref_ivfpq *h, const float *q, int64_t nq, int k,
                                  int nprobe, int k_factor, float *D, int64_t *I, int nthreads);

/* ---- IVF-Flat (mojovec/index/index_ivf_flat.mojo) ---- */
typedef struct mvref_ivfflat mvref_ivfflat;
mvref_ivfflat *mvref_ivfflat_create(int d, int nlist, int metric, int quantizer_metric);
void mvref_ivfflat_destroy(mvref_ivfflat *h);
void mvref_ivfflat_set_centroids(mvref_ivfflat *h, const float *coarse);
void mvref_ivfflat_get_centroids(const mvref_ivfflat *h, float *coarse);
void mvref_ivfflat_train(mvref_ivfflat *h, const float *x, int64_t n, int niter,
                         uint64_t seed, int nthreads);
void mvref_ivfflat_add(mvref_ivfflat *h, const float *x, int64_t n, const int64_t *ids,
                       int nthreads);
void mvref_ivfflat_load_lists(mvref_ivfflat *h, const int64_t *list_off, const int32_t *ids,
                              const float *rows, int64_t ntotal);
int64_t mvref_ivfflat_list_size(const mvref_ivfflat *h, int list);
void mvref_ivfflat_list(const mvref_ivfflat *h, int list, int64_t *ids, float *rows);
void mvref_ivfflat_search(const mvref_ivfflat *h, const float *q, int64_t nq, int k,
                          int nprobe, const uint8_t *exclude, int64_t exclude_len,
                          float *D, int64_t *I, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
