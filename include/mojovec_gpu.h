/*
 * mojovec_gpu.h — C ABI of the MI355X (gfx950) batched vector-scan library
 * (libmojovec_gpu.so).
 *
 * The entry points replace, on one GPU, the native calls that MojoVec's host
 * code makes for exact Flat / SQ8 search, IVF coarse assignment, PQ
 * LUT + inverted-list scan and the Lloyd k-means that trains IVF centroids and
 * PQ codebooks. Each declaration cites the reference interface it replaces
 * (paths relative to the MojoVec v0.8.2 source tree).
 *
 * Conventions
 *   - Plain pointers and sizes only. Host pointers are borrowed for the call;
 *     the library owns all device memory of an index handle.
 *   - Vectors are row-major float32 [n x d]; results are row-major [nq x k],
 *     ascending distance, exact ties broken by the lower id.
 *   - Metrics: 0 = L2 (squared), 1 = inner product (core/types.mojo:5-7);
 *     collections also accept "cosine".
 *   - Every function returning int returns MVG_OK (0) or an MVG_ERR_* code;
 *     mvg_last_error() then holds the message (thread-local).
 *   - A handle may be searched from several threads; mutations need external
 *     exclusion, as in the reference (README "Thread safety").
 */
#ifndef MOJOVEC_GPU_H
#define MOJOVEC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MVG_OK 0
#define MVG_ERR_INVALID 1 /* bad argument (the reference raises Error) */
#define MVG_ERR_DEVICE 2  /* HIP runtime failure */
#define MVG_ERR_ALLOC 3   /* device or host allocation failure */
#define MVG_ERR_STATE 4   /* call not valid in the handle's state */

#define MVG_METRIC_L2 0
#define MVG_METRIC_IP 1

typedef struct mvg_flat mvg_flat;
typedef struct mvg_sq8 mvg_sq8;
typedef struct mvg_ivfpq mvg_ivfpq;
typedef struct mvg_ivfflat mvg_ivfflat;
typedef struct mvg_collection mvg_collection;

/* ---- library ---------------------------------------------------------- */
const char *mvg_last_error(void);
const char *mvg_version(void);
/* Name and compute-unit count of the current HIP device. */
int mvg_device_info(char *name, int name_len, int *compute_units);

/* ---- exact Flat: IndexFlat (mojovec/index/index_flat.mojo:117-330) ---- */
/* IndexFlat.__init__ (:131-151) */
int mvg_flat_create(int d, int metric, mvg_flat **out);
/* IndexFlat.add (:185-210) */
int mvg_flat_add(mvg_flat *h, const float *x, int64_t n);
/* IndexFlat.search (:229-258, _search_impl :260-325). exclude[j] > 0 skips
 * internal id j (the `filter` span of core/index.mojo:21-31); may be NULL.
 * Slots past the number of candidates hold (0.0, -1). */
int mvg_flat_search(mvg_flat *h, const float *q, int64_t nq, int k, const uint8_t *exclude,
                    int64_t exclude_len, float *distances, int64_t *labels);
/* Search form (no reference counterpart). precision 0: plain FP32 in the
 * reference's operation order; 1: mixed, candidates from a BF16 GEMM-form
 * distance on MFMA (FP32 accumulate); 2: mixed with FP8 E4M3 operands (each
 * row scaled by a power of two, FP32 accumulate; inner products are scored
 * as c_q + c_x - q~.x~ with c = (|v~|^2 - |v|^2) / 2 over the dequantized
 * row v~). Mixed candidates are reranked with exact FP32 distances when
 * k_factor > 1; with k_factor 1 the candidate distances are returned. */
int mvg_flat_set_search_mode(mvg_flat *h, int precision, int k_factor);
/* FP8 E4M3 row encoding of the mixed Flat form (no reference counterpart):
 * codes [n x roundup(d, 64)] (zero padded), inv_scale [n] = 2^-e where 2^e
 * puts the row's largest magnitude in [224, 448], norms [n] (nullable) of the
 * dequantized rows. */
int mvg_fp8_encode(const float *x, int64_t n, int d, uint8_t *codes, float *inv_scale,
                   float *norms);
/* Batched distance service for a CPU graph search (HNSW build and search
 * stay on the CPU; their candidate evaluations can be sent here in batches).
 * FlatDistanceComputer (mojovec/index/index_flat.mojo:36-115) for nq x
 * per_query pairs: out[i * per_query + j] = distance of query i to row
 * ids[i * per_query + j], bit-identical to distance() (order 0:
 * l2_distance_simd[8] / -inner_product_simd[8], :60-68) or distance_batch4()
 * (order 1: l2_distance_simd_batch4[8] / -inner_product_simd_batch4[8],
 * :70-90; the form hnsw_graph.mojo:315 calls). ids must lie in [0, ntotal). */
int mvg_flat_distances(mvg_flat *h, const float *q, int64_t nq, const int64_t *ids,
                       int64_t per_query, int order, float *out);
/* FlatDistanceComputer.symmetric_distance (index_flat.mojo:92-99) for row
 * pairs (a[p], b[p]). */
int mvg_flat_pair_distances(mvg_flat *h, const int64_t *a, const int64_t *b, int64_t npairs,
                            float *out);
int64_t mvg_flat_ntotal(const mvg_flat *h);
/* IndexFlat.get_vector / get_vector_span (index_flat.mojo:212-227): the
 * stored row `id` (d floats) copied to out; ids must lie in [0, ntotal).
 * get_vectors gathers rows ids[0..n) into out [n x d]. */
int mvg_flat_get_vector(mvg_flat *h, int64_t id, float *out);
int mvg_flat_get_vectors(mvg_flat *h, const int64_t *ids, int64_t n, float *out);
void mvg_flat_destroy(mvg_flat *h);

/* ---- SQ8 Flat: IndexFlatSQ8 (mojovec/index/index_flat_sq8.mojo:192-582) */
/* IndexFlatSQ8.__init__ (:209-237) */
int mvg_sq8_create(int d, int metric, mvg_sq8 **out);
/* IndexFlatSQ8.add (:286-397): global min/max calibration, re-encode on
 * range growth, FP32 rows kept for the rerank. */
int mvg_sq8_add(mvg_sq8 *h, const float *x, int64_t n);
/* IndexFlatSQ8.search (:399-526). k_factor > 1 reranks the k*k_factor code
 * candidates with exact FP32 distances, as the SQ8 rerank of
 * index_hnsw.mojo:433-496 (fixed at 2 there). Pad (0.0, -1). */
int mvg_sq8_search(mvg_sq8 *h, const float *q, int64_t nq, int k, int k_factor,
                   const uint8_t *exclude, int64_t exclude_len, float *distances,
                   int64_t *labels);
/* Calibration state: global_min, global_max, scale (:203-206). */
/* SQ8DistanceComputer (index_flat_sq8.mojo:21-190) for nq x per_query
 * pairs: the query is encoded as get_distance_computer does (:527-560), the
 * distance is the code distance of distance() / distance_batch4() (equal:
 * exact code dots), sq8.mojo:62-79. ids must lie in [0, ntotal). */
int mvg_sq8_distances(mvg_sq8 *h, const float *q, int64_t nq, const int64_t *ids,
                      int64_t per_query, float *out);
/* SQ8DistanceComputer.symmetric_distance (index_flat_sq8.mojo:148-177). */
int mvg_sq8_pair_distances(mvg_sq8 *h, const int64_t *a, const int64_t *b, int64_t npairs,
                           float *out);
int mvg_sq8_params(const mvg_sq8 *h, float *gmin, float *gmax, float *scale, int64_t *ntotal);
/* Codes [ntotal x d] and code norms [ntotal] (codes_u8 / norms_u32). */
int mvg_sq8_get_codes(mvg_sq8 *h, uint8_t *codes, uint32_t *norms);
/* IndexFlatSQ8.get_vector / get_vector_span (index_flat_sq8.mojo:567-582):
 * the retained FP32 row `id` (codes_f32); get_vectors gathers rows
 * ids[0..n) into out [n x d]. ids must lie in [0, ntotal). */
int mvg_sq8_get_vector(mvg_sq8 *h, int64_t id, float *out);
int mvg_sq8_get_vectors(mvg_sq8 *h, const int64_t *ids, int64_t n, float *out);
/* Code-dot engine (no reference counterpart): 1 (default) contracts the
 * codes on int8 MFMA (affine uint8 rebased to int8 with an exact correction),
 * 0 uses dot4 VALU instructions. Both give bit-identical results. */
int mvg_sq8_set_search_mode(mvg_sq8 *h, int use_mfma);
void mvg_sq8_destroy(mvg_sq8 *h);

/* ---- k-means: KMeans (mojovec/clustering/kmeans.mojo:8-220) ------------- */
/* KMeans(d, k, niter).train(x). init == NULL selects the seeded pick of
 * kmeans.mojo:67-82 (seed 0x4D4F4A4F564543 there); otherwise init is
 * [k x d] and sample_offset drives the empty-cluster reseed (:211-220).
 * assign (nullable) receives the last iteration's assignment. */
int mvg_kmeans_train(const float *x, int64_t n, int d, int k, int niter, const float *init,
                     int64_t sample_offset, uint64_t seed, float *centroids, int32_t *assign);

/* ---- PQ: ProductQuantizer (mojovec/quantization/pq.mojo:80-292) -------- */
/* ProductQuantizer.train (:125-176): codebooks [M x 256 x d/M]. */
int mvg_pq_train(const float *x, int64_t n, int d, int M, int niter, uint64_t seed,
                 float *codebooks);
/* ProductQuantizer.compute_codes (:178-219): codes [n x M]. */
int mvg_pq_encode(const float *x, int64_t n, int d, int M, const float *codebooks,
                  uint8_t *codes);
/* ProductQuantizer.decode (:221-250): codes [n x M] -> rows [n x d], each
 * subvector the codebook entry its code names. */
int mvg_pq_decode(const uint8_t *codes, int64_t n, int d, int M, const float *codebooks,
                  float *out);
/* ProductQuantizer.compute_distance_table (:252-292) for nq vectors:
 * lut [nq x M x 256]. */
int mvg_pq_lut(const float *q, int64_t nq, int d, int M, const float *codebooks, int metric,
               float *lut);

/* ---- IVF-PQ: IndexIVFPQ (mojovec/index/index_ivf_pq.mojo:12-360) ------- */
/* IndexIVFPQ.__init__ (:35-71); nprobe starts at 1 (:63). */
int mvg_ivfpq_create(int d, int nlist, int M, int metric, mvg_ivfpq **out);
/* IndexIVFPQ.train (:87-136). The reference fixes both k-means iteration
 * counts at 15 (:98, pq.mojo:64); they are explicit here. No-op once trained. */
int mvg_ivfpq_train(mvg_ivfpq *h, const float *x, int64_t n, int niter_coarse, int niter_pq,
                    uint64_t seed);
/* Installs trained codebooks (coarse [nlist x d], pq [M x 256 x d/M]) and
 * marks the index trained: the state IndexIVFPQ.train leaves behind. */
int mvg_ivfpq_set_codebooks(mvg_ivfpq *h, const float *coarse, const float *pq);
int mvg_ivfpq_get_codebooks(mvg_ivfpq *h, float *coarse, float *pq);
/* IndexIVFPQ.add (:138-148): internal ids ntotal..ntotal+n-1. Silently does
 * nothing while untrained (:161). FP32 rows are kept on the device for the
 * k_factor rerank unless disabled with mvg_ivfpq_set_keep_vectors(h, 0)
 * before the first add. */
int mvg_ivfpq_add(mvg_ivfpq *h, const float *x, int64_t n);
/* IndexIVFPQ.add_with_ids (:150-213): as add, with the caller's ids stored in
 * the lists (ids NULL = add). Ids must lie in [0, 2^31 - 1] (the device id
 * width); ntotal grows by n and the FP32 rows stay in add order. While rows
 * are kept for the rerank the ids must be unique across the index (the
 * rerank finds a row by its id); a duplicate is refused before anything
 * changes. Searches, the exclusion mask and ties use these ids. */
int mvg_ivfpq_add_with_ids(mvg_ivfpq *h, const float *x, int64_t n, const int64_t *ids);
int mvg_ivfpq_set_keep_vectors(mvg_ivfpq *h, int keep);
/* The public `nprobe` field (:22). Values above nlist are clamped at search
 * time (:253-254). */
int mvg_ivfpq_set_nprobe(mvg_ivfpq *h, int nprobe);
/* IndexIVFPQ.search (:215-360). nprobe_per_query (nullable, [nq]) overrides
 * nprobe per query; k_factor > 1 reranks k*k_factor ADC candidates with exact
 * FP32 distances; exclude[id] != 0 skips internal id (:328-334). Pad
 * (1e38, -1); inner-product results are un-negated (:357-360). */
int mvg_ivfpq_search(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                     const int32_t *nprobe_per_query, int k_factor, const uint8_t *exclude,
                     int64_t exclude_len, float *distances, int64_t *labels);
/* Same search with every pointer in device memory, enqueued on `stream`
 * (a hipStream_t; NULL = the handle's stream) without synchronising. */
int mvg_ivfpq_search_device(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                            const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                            const uint8_t *exclude, int64_t exclude_len, float *distances,
                            int64_t *labels, void *stream);
/* Batched search with a k per query (no reference counterpart; the
 * reference's search takes one k per call, :215-224): query i returns
 * k_per_query[i] results (device [nq], values in 1..max_k) in output rows of
 * stride max_k, the slots past its own k padded (1e38, -1). Each query's
 * results equal those of this search with k = k_per_query[i] alone; the
 * batch runs as one scan for all queries. */
int mvg_ivfpq_search_device_k(mvg_ivfpq *h, const float *q, int64_t nq, const int32_t *k_per_query,
                              int max_k, const int32_t *nprobe_per_query, int max_nprobe,
                              int k_factor, const uint8_t *exclude, int64_t exclude_len,
                              float *distances, int64_t *labels, void *stream);
/* Allocation-free form of mvg_ivfpq_search_device. No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results; its workspace (probe lists, candidate keys,
 * pools) comes from one fixed 64 MiB device array in the library, one per
 * GPU, held until the process ends, and the call makes no device allocation
 * or free. A request whose workspace exceeds 64 MiB returns MVG_ERR_INVALID
 * (checked before anything is enqueued for the probe lists and candidate
 * keys, then piece by piece; results are then not written). The index must
 * be prepared: after a change, run one mvg_ivfpq_search or
 * mvg_ivfpq_search_device first, else MVG_ERR_STATE. Every call on one GPU
 * shares the array: calls that overlap on a GPU (other threads or streams)
 * race. */
int mvg_ivfpq_search_device_noalloc(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                                    const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                                    const uint8_t *exclude, int64_t exclude_len, float *distances,
                                    int64_t *labels, void *stream);
/* Search form (no reference counterpart). precision 0: plain FP32, the
 * reference's residual-LUT arithmetic in its operation order (bit-exact ids);
 * 1: mixed, L2 tables decomposed as |q - c_l|^2 + T[l] + Q[q] (precomputed
 * per list / per query), candidates then reranked exactly when k_factor > 1.
 * split_scan 1 selects one workgroup per (query, list) with a merge pass
 * instead of one workgroup per query (same results). */
int mvg_ivfpq_set_search_mode(mvg_ivfpq *h, int precision, int split_scan);
/* Mixed-form pre-rerank (no reference counterpart): with stored rows, the
 * mixed form scores its k * k_factor candidates by the SQ8 code distance of
 * the rows (global range and scale as IndexFlatSQ8 computes them,
 * index_flat_sq8.mojo:317-367) and reranks only the best max(64, per_k * k)
 * in exact FP32. per_k 0 reranks every candidate in FP32. Default 8. */
int mvg_ivfpq_set_prefilter(mvg_ivfpq *h, int per_k);
/* Mixed-form scan tables (no reference counterpart; pq.mojo:252-292 builds
 * the FP32 tables this quantizes): bits 8 makes the mixed L2 deep-pool scan
 * (k * k_factor at or above the deep-pool threshold, M in {8, 12, 16}) that
 * is followed by the exact rerank (k_factor > 1) sum 8-bit tables: each
 * subquantizer's FP32 table minus its minimum, in steps of one scale per
 * query (the largest range over 255), rounded, summed exactly in integers.
 * Its candidate distances are then within M * scale / 2 of the FP32 table
 * sums; the returned distances are the rerank's. bits 16 runs the mixed L2
 * candidate scan that the exact rerank follows (k_factor > 1; PQ with
 * 8-dimensional subspaces, M in {8, 12, 16}) on FP16 MFMA tiles of 16
 * queries: the scaled query times the PQ-decoded entry, FP32 accumulation,
 * plus the FP32 list and entry terms; batches below the tile minimum
 * (mvg_ivfpq_set_tile_min_batch) take the 8-bit scan. Every other form and
 * shape keeps FP32 tables. bits 32 (default): FP32 tables everywhere. */
int mvg_ivfpq_set_table_bits(mvg_ivfpq *h, int bits);
/* Tile-scan batch minimum (no reference counterpart): with table bits 16, a
 * batch of at least min_queries queries takes the FP16 MFMA tile scan
 * (0, the default: 16 queries per compute unit, one tile each; below that
 * each tile's union of lists runs on one compute unit and the per-query
 * scan is faster). Candidates of a query do not depend on the batch. */
int mvg_ivfpq_set_tile_min_batch(mvg_ivfpq *h, int64_t min_queries);
/* Skewed-list load balance (no reference counterpart): entries_per_item > 0
 * runs the scan as persistent workgroups over a device-built queue of
 * (query, entry-range) items of about that many list entries (at most 8 per
 * query), merged per query afterwards; 0 (default) scans one query per
 * workgroup. Results are identical in the plain form; in the mixed form the
 * queue's kernel sums each ADC distance from per-list tables rather than the
 * per-entry terms, so distances agree up to FP32 rounding. */
int mvg_ivfpq_set_work_queue(mvg_ivfpq *h, int entries_per_item);
/* Deep candidate pools (no reference counterpart): from k * k_factor >=
 * min_candidates (0 = default 513) the scan selects its candidates without
 * a sorted running top set. pool 0 (default): every visited entry's key is
 * stored in HBM and the top k * k_factor selected once per query by radix
 * select; pool 1: a running candidate set in LDS, re-thresholded by radix
 * select when its append buffer fills (inner-product and mixed forms; the
 * plain L2 form always uses pool 0). Results are identical in every case. */
int mvg_ivfpq_set_deep_select(mvg_ivfpq *h, int min_candidates, int pool);
/* Stage timing of the next searches (no reference counterpart; used by the
 * benchmark): when on, each search records HIP events on its stream around
 * the coarse, scan and merge/rerank stages, counts the list entries the scan
 * visits and the keys the deep-pool stream scan writes to its HBM pools.
 * read_profile returns the last search's stage_ms[3] and those two counts
 * (each output nullable). */
int mvg_ivfpq_set_profiling(mvg_ivfpq *h, int on);
int mvg_ivfpq_read_profile(mvg_ivfpq *h, float *stage_ms, int64_t *scanned_entries,
                           int64_t *pool_writes);
/* State: ntotal, trained flag, nprobe. */
int mvg_ivfpq_info(mvg_ivfpq *h, int64_t *ntotal, int *trained, int *nprobe);
/* Construction parameters: d, nlist, M, metric (IndexIVFPQ fields d, nlist,
 * M, metric_type, index_ivf_pq.mojo:19-33). */
int mvg_ivfpq_shape(mvg_ivfpq *h, int *d, int *nlist, int *M, int *metric);
/* Inverted-list sizes [nlist] and one list's ids/codes in append order
 * (ArrayInvertedLists.list_size/get_ids/get_codes, inverted_lists.mojo). */
int mvg_ivfpq_list_sizes(mvg_ivfpq *h, int64_t *sizes);
int mvg_ivfpq_get_list(mvg_ivfpq *h, int list, int64_t *ids, uint8_t *codes);
void mvg_ivfpq_destroy(mvg_ivfpq *h);

/* ---- IVF-Flat: IndexIVFFlat (mojovec/index/index_ivf_flat.mojo:10-244) -- */
/* IndexIVFFlat.__init__ (:29-52) with an IndexFlat quantizer of the same
 * metric; nprobe starts at 1 (:46). */
int mvg_ivfflat_create(int d, int nlist, int metric, mvg_ivfflat **out);
/* The same with the quantizer's own metric (IndexFlat(d) is L2 whatever the
 * index metric): coarse assignment and probe selection use quantizer_metric,
 * list scoring uses metric. */
int mvg_ivfflat_create_with_quantizer(int d, int nlist, int metric, int quantizer_metric,
                                      mvg_ivfflat **out);
/* IndexIVFFlat.train (:65-77): coarse k-means, which the reference fixes at
 * 15 iterations (:73); explicit here. No-op once trained (:71). */
int mvg_ivfflat_train(mvg_ivfflat *h, const float *x, int64_t n, int niter, uint64_t seed);
/* Installs coarse centroids [nlist x d] and marks the index trained (the
 * quantizer.add + is_trained of :75-77). */
int mvg_ivfflat_set_centroids(mvg_ivfflat *h, const float *coarse);
int mvg_ivfflat_get_centroids(mvg_ivfflat *h, float *coarse);
/* IndexIVFFlat.add / add_with_ids (:79-137): ids NULL gives ntotal..
 * ntotal+n-1; explicit ids must lie in [0, 2^31 - 1]. Rows go to their
 * nearest centroid's list in input order; nothing happens while untrained
 * (:102-104). */
int mvg_ivfflat_add(mvg_ivfflat *h, const float *x, int64_t n, const int64_t *ids);
/* The public `nprobe` field (:19); clamped to nlist at search time (:178).
 * nprobe above 4096 is rejected at search time. */
int mvg_ivfflat_set_nprobe(mvg_ivfflat *h, int nprobe);
/* IndexIVFFlat.search (:139-244): L2 = l2_distance_simd[8], IP =
 * inner_product_simd[4] (:222-225); ascending, ties by lower id; pad
 * (1e38, -1); IP un-negated over found entries (:241-243). k <= 4096. The
 * reference accepts a filter and ignores it (:149-156); exclude (nullable)
 * here skips ids with exclude[id] != 0, as IVF-PQ does. */
int mvg_ivfflat_search(mvg_ivfflat *h, const float *q, int64_t nq, int k, const uint8_t *exclude,
                       int64_t exclude_len, float *distances, int64_t *labels);
/* Same search with every pointer in device memory, enqueued on `stream`
 * (NULL = the handle's stream) without synchronising. */
int mvg_ivfflat_search_device(mvg_ivfflat *h, const float *q, int64_t nq, int k,
                              const uint8_t *exclude, int64_t exclude_len, float *distances,
                              int64_t *labels, void *stream);
/* State (ntotal, is_trained, nprobe) and construction parameters. */
int mvg_ivfflat_info(mvg_ivfflat *h, int64_t *ntotal, int *trained, int *nprobe);
int mvg_ivfflat_shape(mvg_ivfflat *h, int *d, int *nlist, int *metric, int *quantizer_metric);
/* Inverted-list sizes [nlist] and one list's ids / FP32 rows in append order
 * (ArrayInvertedLists with code_size 4d, :52). */
int mvg_ivfflat_list_sizes(mvg_ivfflat *h, int64_t *sizes);
int mvg_ivfflat_get_list(mvg_ivfflat *h, int list, int64_t *ids, float *rows);
void mvg_ivfflat_destroy(mvg_ivfflat *h);
/* write_index_ivf_flat / read_index_ivf_flat (mojovec/io/serialization.mojo:
 * 980-1020) to / from a file, the reference's byte layout (no checksum
 * trailer at index level). Loading also rejects a quantizer that contradicts
 * the header, nprobe <= 0 and ids outside [0, 2^31 - 1]. */
int mvg_ivfflat_save(mvg_ivfflat *h, const char *path);
int mvg_ivfflat_load(const char *path, mvg_ivfflat **out);

/* ---- captured search batch (no reference counterpart) ------------------ */
/* Records one whole device-pointer search (coarse, work-queue build, scan,
 * merge, rerank) of a fixed batch shape into a HIP graph with its own
 * workspace; mvg_graph_launch replays it on `stream` (a hipStream_t; NULL =
 * the null stream) against the same query/output buffers, whose contents may
 * change between replays. A graph replays only while its index is unchanged:
 * after an add, train or codebook change mvg_graph_launch returns
 * MVG_ERR_STATE (capture again). Destroy graphs before their index. */
typedef struct mvg_graph mvg_graph;
int mvg_ivfpq_capture(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                      const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                      float *distances, int64_t *labels, mvg_graph **out);
/* The same with a k per query (device k_per_query [nq], output stride max_k),
 * as mvg_ivfpq_search_device_k: one graph for a batch of mixed k. */
int mvg_ivfpq_capture_k(mvg_ivfpq *h, const float *q, int64_t nq, const int32_t *k_per_query,
                        int max_k, const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                        float *distances, int64_t *labels, mvg_graph **out);
int mvg_graph_launch(mvg_graph *g, void *stream);
/* Times the scan schedules (one workgroup per query; work queue with 1024,
 * 4096 or 16384 entries per item; one workgroup per (query, list)) on this
 * batch shape (device pointers) and keeps the fastest on the handle.
 * choice: 0..4 in that order; best_ms: its time per search. */
int mvg_ivfpq_autotune(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                       const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                       float *distances, int64_t *labels, int *choice, float *best_ms);
void mvg_graph_destroy(mvg_graph *g);

/* ---- continuous-batching request loop (no reference counterpart) ------- */
/* For serving callers of an HBM-resident IVF-PQ or Flat SQ8 index. A dispatcher thread
 * packs waiting requests with equal k_factor into batches of at most
 * max_batch_queries queries (requests of different k share a batch, each
 * query with its own k and nprobe; over a Flat SQ8 index only when their
 * k * k_factor fall on the same side of 64, where its MFMA form changes),
 * admits a batch when one of `slots` (1..16) streams is free and its
 * device-memory estimate fits mem_budget_bytes, and overlaps the batches on
 * the device. Each request's result equals the same request searched alone.
 * While a server holds an IVF-PQ index, its mutations and search settings
 * return MVG_ERR_STATE. */
typedef struct mvg_server mvg_server;
int mvg_server_create(mvg_ivfpq *index, int slots, int max_batch_queries, int64_t mem_budget_bytes,
                      mvg_server **out);
/* The same loop over a Flat SQ8 index (mvg_sq8_search per batch; a request's
 * nprobe is ignored). */
int mvg_server_create_sq8(mvg_sq8 *index, int slots, int max_batch_queries,
                          int64_t mem_budget_bytes, mvg_server **out);
/* One loop over both kinds of index (either pointer may be NULL): the same
 * dispatcher, slots and memory budget serve requests of the IVF-PQ index and
 * of the Flat SQ8 index; a batch holds requests of one index. */
int mvg_server_create_multi(mvg_ivfpq *ivfpq, mvg_sq8 *sq8, int slots, int max_batch_queries,
                            int64_t mem_budget_bytes, mvg_server **out);
/* Copies q [nq x d]; distances/labels [nq x k] must stay valid until
 * mvg_server_wait(ticket) returns. The request goes to the server's only
 * index (the IVF-PQ one when it holds both). */
int mvg_server_submit(mvg_server *s, const float *q, int64_t nq, int k, int nprobe, int k_factor,
                      float *distances, int64_t *labels, int64_t *ticket);
/* The same for a named index: target 0 = the IVF-PQ index, 1 = the Flat SQ8
 * index (q has that index's d columns). */
int mvg_server_submit_to(mvg_server *s, int target, const float *q, int64_t nq, int k, int nprobe,
                         int k_factor, float *distances, int64_t *labels, int64_t *ticket);
/* Blocks until the request is done; latency_ms = submit to completion. */
int mvg_server_wait(mvg_server *s, int64_t ticket, double *latency_ms);
int mvg_server_stats(mvg_server *s, int64_t *completed_queries, int64_t *batches,
                     int64_t *max_inflight_batches);
void mvg_server_destroy(mvg_server *s);

/* ---- managed collection: CollectionIVFPQ ------------------------------- */
/* (mojovec/api/collection_ivfpq.mojo:52-333). Messages match the reference.
 * CollectionIVFPQ.__init__ (:68-100); nprobe 0 selects min(10, nlist). */
int mvg_coll_create(int d, int nlist, int M, int nprobe, const char *name, const char *metric,
                    mvg_collection **out);
/* CollectionIVFPQ.train (:169-195); `count` values (a multiple of d). */
int mvg_coll_train(mvg_collection *h, const float *x, int64_t count);
/* CollectionIVFPQ.add (:206-242): unique user ids, auto-train on the first
 * batch. */
int mvg_coll_add(mvg_collection *h, const int64_t *ids, int64_t n, const float *x, int64_t count);
/* CollectionIVFPQ.query (:244-333) plus k_factor (1 = reference behaviour):
 * ids/distances [nq x n_results]; cosine and ip distances are 1 - sim.
 * *nq_out receives count / d. */
int mvg_coll_query(mvg_collection *h, const float *q, int64_t count, int n_results, int k_factor,
                   int64_t *ids, float *distances, int64_t *nq_out);
/* set_nprobe (:152-155). */
int mvg_coll_set_nprobe(mvg_collection *h, int nprobe);
/* k-means iteration counts used by train (15 / 15 in the reference). */
int mvg_coll_set_train_params(mvg_collection *h, int niter_coarse, int niter_pq, uint64_t seed);
/* name(), dimension(), metric(), count(), nlist(), pq_subvectors(),
 * nprobe(), is_trained() (:124-150). metric: 0 l2, 1 ip, 2 cosine. */
int mvg_coll_info(mvg_collection *h, int *dimension, int *metric, int64_t *count, int *nlist,
                  int *M, int *nprobe, int *trained);
int mvg_coll_name(mvg_collection *h, char *buf, int buf_len);
mvg_ivfpq *mvg_coll_index(mvg_collection *h);
void mvg_coll_destroy(mvg_collection *h);

/* ---- snapshots in the reference's file format -------------------------- */
/* CollectionIVFPQ.save (collection_ivfpq.mojo:335-370): collection header,
 * user ids, the IndexIVFPQ stream and a checksum trailer, written to a
 * temporary file beside `path` and renamed over it. */
int mvg_coll_save(mvg_collection *h, const char *path);
/* CollectionIVFPQ.load (collection_ivfpq.mojo:372-476): validates the
 * trailer and every header with the reference's messages, then builds the
 * HBM index. Snapshots carry no FP32 rows: a loaded collection answers
 * k_factor 1 queries only. */
int mvg_coll_load(const char *path, mvg_collection **out);
/* write_index_ivf_pq / read_index_ivf_pq (mojovec/io/serialization.mojo:
 * 1029-1090) as a whole file: the bare index stream, no trailer. Internal
 * ids must fit in int32 on the device. */
int mvg_ivfpq_save(mvg_ivfpq *h, const char *path);
int mvg_ivfpq_load(const char *path, mvg_ivfpq **out);
/* _checksum_prefix (mojovec/io/snapshot_file.mojo:19-59) of a payload held
 * in host memory: 8 MiB chunks from offset 0. Host only; no device needed. */
int mvg_snapshot_checksum(const uint8_t *data, int64_t len, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
