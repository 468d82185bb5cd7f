// Host-side launchers of the HIP kernels (one translation unit per family).
#pragma once

#include "common.h"

namespace mvg {

// ---- flat.hip ----
// Canonical top-k of q against base rows (IndexFlat search). Outputs are any
// subset of: distances (IP un-negated), int64 ids, int32 ids, u64 keys.
void flat_topk(const float *base, int64_t n, int d, int metric, const float *q, int64_t nq, int k,
               const uint8_t *excl, int64_t excl_len, DevBuf &part, float pad_dist, float *D,
               int64_t *I64, int32_t *I32, uint64_t *keys, hipStream_t s);
// Canonical argmin (k = 1); kmeans_mode restates the k-means scan rules.
void flat_argmin(const float *base, int64_t n, int d, int metric, bool kmeans_mode, const float *q,
                 int64_t nq, int32_t *out_id, float *out_dist, hipStream_t s);
// Reduce ncand candidate keys per query to the sorted top-k. With seg_off,
// query q owns units [seg_off[q], seg_off[q+1]) of ncand keys each. With kq
// (nullable, [nq]) query q keeps only its kq[q] * kmul (<= k) smallest keys
// in its k output slots (the rest padded): a prefix of the sorted top-k, so
// the set a top-(kq[q] * kmul) selection alone would keep.
void launch_merge(const uint64_t *cand, int64_t nq, int64_t ncand, int k, int negate,
                  float pad_dist, float *D, int64_t *I64, int32_t *I32, uint64_t *keys,
                  hipStream_t s, const int64_t *seg_off = nullptr, const int32_t *kq = nullptr,
                  int kmul = 1);

// ---- kmeans.hip ----
struct KmeansWork {
  DevBuf assign, keys, rows, keys_sorted, rows_sorted, seg, sort_tmp;
};
// Lloyd iterations over S independent subspaces (S = 1: coarse quantizer;
// S = M: PQ codebooks) of dsub columns each, starting from `cent`
// ([S][k][dsub], updated in place). x is [n][S * dsub] row-major.
void kmeans_fit(const float *x, int64_t n, int S, int dsub, int k, int niter,
                int64_t sample_offset, float *cent, KmeansWork &w, hipStream_t s);
// Per (row, subspace) argmin over the 256 PQ centroids. If coarse/labels are
// given, the residual x - coarse[label] is encoded (index_ivf_pq.mojo:181-192).
void pq_assign(const float *x, int64_t n, int d, int M, const float *cb, const float *coarse,
               const int32_t *labels, uint8_t *codes, int32_t *assign, hipStream_t s);
// LUT[q][m][256] (pq.mojo:252-292) for nq query (or residual) vectors.
void pq_lut(const float *q, int64_t nq, int d, int M, const float *cb, int metric, float *lut,
            hipStream_t s);
// ProductQuantizer.decode (pq.mojo:221-250): codes [n][M] -> rows [n][d].
void pq_decode(const uint8_t *codes, int64_t n, int d, int M, const float *cb, float *out,
               hipStream_t s);
// out[i] = base row ids[i] (every id in range; checked by the caller).
void gather_rows(const float *base, int d, const int64_t *ids, int64_t n, float *out, hipStream_t s);
// x - coarse[label] row-wise.
void residuals(const float *x, int64_t n, int d, const float *coarse, const int32_t *labels,
               float *out, hipStream_t s);

// ---- ivfpq.hip ----
// Compute units of the current device (cached after the first call).
int device_cus();
// Candidate id -> row of the stored FP32 rows (and of their SQ8 codes):
// the id itself when row_of is null, else row_of[id], where ids at or above
// nids and entries of -1 have no row.
struct RowMap {
  const int32_t *row_of = nullptr;
  int64_t nids = 0;
};

struct IvfSearchArgs {
  const float *q;
  int64_t nq;
  int d, M, nlist, metric;
  const float *coarse, *pq;
  const int64_t *list_off;
  const uint8_t *codes;
  const int32_t *ids;
  const float *vectors;  // may be null when no rerank is requested
  int k, k_factor, nprobe;
  const int32_t *nprobe_per_query;  // device, nullable
  const uint8_t *excl;
  int64_t excl_len;
  float *D;
  int64_t *I;
  int precision;       // 0 plain FP32 (reference order), 1 mixed
  const float *ptab;   // mixed L2 tables [nlist][M][256]
  int split_scan;      // 1: one workgroup per (query, list) + merge pass
  const uint16_t *coarse_bf16;  // mixed coarse: bf16 centroid rows (padded)
  const float *coarse_norm;     // and their squared norms
  int wq_entries;  // > 0: persistent work queue, items of about this many entries
  int64_t ntotal;  // rows of `vectors`
  int64_t ncsr;    // list entries (CSR length of codes / ids / pre)
  int64_t deep_cap;  // sum of the nprobe largest lists: most entries one query visits
  int deep_min;      // k * k_factor from which the deep-pool scan runs (0: default)
  int64_t tile_min_batch;  // table bits 16: fewest queries for the tile scan (0: 16 per CU)
  const float *pre;  // mixed L2: per-entry sum_m T[l][m][code_m] [ntotal] (nullable)
  int deep_pool;     // stream scan, large pools: 1 keys in HBM, 0 running set in LDS
  // mixed form: SQ8 codes of the stored rows (nullable) for the pre-rerank,
  // which keeps max(64, prefilter * k) of the k * k_factor candidates
  const uint8_t *vq_codes;  // [ntotal][sq8_row_stride(d)]
  const uint32_t *vq_norms;
  float vq_min, vq_scale;
  int prefilter;
  // mixed L2 deep-pool scan followed by the exact rerank: 8 selects 8-bit
  // tables (per-subquantizer offset, one scale per query, exact integer
  // sums); anything else FP32 tables
  int table_bits;
  const float *pre8;  // with table_bits 8: centred entry terms [ncsr] and
  const float *mu;    // the centre [d] (nullable: FP32 tables)
  RowMap rows;  // explicit ids: id -> row of `vectors` / `vq_codes`
  // per-query k (nullable, device [nq]): query i returns kq[i] <= k results
  // in rows of stride k (padded past kq[i]); results equal those of a search
  // with k = kq[i] alone
  const int32_t *kq;
  // table_bits 16 (mixed L2 with a rerank): FP16 codebooks [M][256][dsub] of
  // -2 * pq scaled by 2^e (cb16_inv = 2^-e) for the MFMA tile scan, and each
  // list's rank in the super-cluster order that groups queries into tiles
  const uint16_t *cb16;
  float cb16_inv;
  const int32_t *list_rank;
};
// pre[e] = sum_m ptab[l][m][code_m] for every entry of the CSR lists.
// 8-bit tables of the mixed L2 form: mu = column means of the coarse
// centroids [d], pre8[e] = pre[e] - 2 <mu, PQ reconstruction of entry e>.
void precompute_centered_terms(const float *coarse, int nlist, int d, const uint8_t *codes, int M,
                               const float *pq, const float *pre, int64_t n, float *mu, float *pre8,
                               hipStream_t s);
void precompute_entry_terms(const int64_t *list_off, int nlist, const uint8_t *codes, int M,
                            const float *ptab, int64_t ntotal, float *pre, hipStream_t s);
// T[l][m][k] = |c_mk|^2 + 2 <c_l,m, c_mk> for the mixed L2 scan.
void precompute_ivf_tables(const float *coarse, const float *pq, int d, int M, int nlist,
                           float *ptab, hipStream_t s);
// Opt-in stage timing: events recorded on the search stream around the
// coarse, scan and merge/rerank stages, and the number of list entries the
// scan visited (sum of probed list lengths).
struct IvfProfile {
  bool on = false;
  bool recorded = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  DevBuf counter;
  hipStream_t stream = nullptr;  // stream of the recorded search (the counter is written there)
};
struct IvfWork {
  DevBuf coarse_part, probes, part, keys, keys2, qbf, qnorm;
  DevBuf wq_poff, wq_nitems, wq_rq, wq_off, wq_gthr, wq_counter;
  DevBuf deep_seg, deep_counter;  // deep candidate pools: key segments, query range counters
  DevBuf ord_cnt, ord_off, ord;    // query order of the persistent deep scans
  DevBuf vq_qcodes, vq_qnorms, vq_keys;  // mixed pre-rerank: query codes, narrowed candidates
  // MFMA tile scan: boundary buffers, tile range counters and the list of
  // queries left to the deep-pool scan
  DevBuf tile_bnd, tile_ctr, tile_ovf, tile_pre;
  bool ord_ready = false;          // `ord` holds this search's order
  IvfProfile prof;
  // Every workspace buffer takes its memory from `a` (before any reserve).
  void bind(Arena *a) {
    for (DevBuf *b : {&coarse_part, &probes, &part, &keys, &keys2, &qbf, &qnorm, &wq_poff, &wq_nitems,
                      &wq_rq, &wq_off, &wq_gthr, &wq_counter, &deep_seg, &deep_counter, &ord_cnt,
                      &ord_off, &ord, &vq_qcodes, &vq_qnorms, &vq_keys, &tile_bnd, &tile_ctr,
                      &tile_ovf, &tile_pre, &prof.counter})
      b->arena = a;
  }
};
void ivfpq_search(const IvfSearchArgs &a, IvfWork &w, hipStream_t s);

// ---- ivfpq_tile.hip ----
// FP16 codebooks for the tile scan: out[m][c][j] = -2 * pq[m][c][j] * 2^e,
// 2^e putting the largest magnitude in [2^14, 2^15); returns 2^-e.
float tile_codebook(const float *pq, int M, int dsub, float max_abs, uint16_t *out, hipStream_t s);
// Whether the MFMA tile scan takes this search (mixed L2, dsub 8, M in
// {8, 12, 16}, at most 128 probes, tables prepared).
bool tile_scan_supported(const IvfSearchArgs &a, int maxp);
// Mixed L2 candidates (kk per query, unordered (key << 32 | id) words) by the
// MFMA tile scan; sets w.ord to the tile order. False when unsupported.
bool launch_tile_scan(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                      uint64_t *out_keys, IvfWork &w, hipStream_t s);
// Queries bucketed by rank[nearest list] (rank nullable: the list itself);
// the order within a bucket is unspecified. Sets w.ord / w.ord_ready.
const int32_t *build_tile_order(const IvfSearchArgs &a, int maxp, const int32_t *probes,
                                const int32_t *rank, IvfWork &w, hipStream_t s);
// The FP32-table deep-pool stream scan (mixed L2) of the queries list[0..*count)
// only (count is read on the device; the launch is always made).
void launch_deep_subset(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                        uint64_t *out_keys, const int32_t *list, const unsigned *count, IvfWork &w,
                        hipStream_t s);

// Exact FP32 rerank of kk candidate keys per query (any order) to the top-k
// (index_hnsw.mojo:449-496 generalised to k_factor).
// Candidates without a row of `vectors` (see RowMap) are dropped.
// order (nullable, [nq]): a query order whose neighbours share candidate
// rows; each XCD then reranks a contiguous stretch of it.
// kq (nullable, [nq]): per-query result counts <= k; the slots of query i
// past kq[i] (stride k) are padded.
void rerank(const float *q, int64_t nq, int d, int metric, const float *vectors, int64_t nrows,
            const uint64_t *cand, int kk, int k, float pad_dist, float *D, int64_t *I,
            hipStream_t s, const int32_t *order = nullptr, RowMap rm = RowMap{},
            const int32_t *kq = nullptr);

// ---- ivfflat.hip ----
struct IvfFlatArgs {
  const float *q;
  int64_t nq;
  int d, nlist, metric, coarse_metric;
  const float *coarse;      // [nlist][d]
  const int64_t *list_off;  // [nlist + 1], CSR over `rows`
  const float *rows;        // [ntotal][d], each list in id order
  const int32_t *ids;       // [ntotal]
  int64_t ntotal, max_list;  // rows, longest list
  int k, nprobe;
  const uint8_t *excl;
  int64_t excl_len;
  float *D;
  int64_t *I;
};
struct IvfFlatWork {
  DevBuf coarse_part, probes, cnt, cursor, pair_off, tile_off, range_ctr, pairs, part;
};
// IndexIVFFlat.search (index_ivf_flat.mojo:139-244) for a trained, non-empty
// index; every pointer is device memory. Pads (1e38, -1).
void ivfflat_search(const IvfFlatArgs &a, IvfFlatWork &w, hipStream_t s);

// ---- gemm.hip ----
// Kinds of the MFMA GEMM-form distance: 0 bf16 L2, 1 bf16 IP, 2 int8 IP
// (symmetric SQ8), 3 uint8 L2 rebased to int8 (affine SQ8), 4 FP8 E4M3 L2,
// 5 FP8 E4M3 IP (rows scaled per row by a power of two).
enum { kGemmBf16L2 = 0, kGemmBf16IP = 1, kGemmI8IP = 2, kGemmU8L2 = 3, kGemmFp8L2 = 4, kGemmFp8IP = 5 };
// Bytes per padded operand row (multiple of 64).
int gemm_row_bytes(int kind, int d);
// FP32 -> bf16 rows padded to gemm_row_bytes; norms (nullable) of the
// rounded rows.
void to_bf16(const float *x, int64_t n, int d, uint16_t *out, float *norms, hipStream_t s);
// FP32 -> FP8 E4M3 rows padded to gemm_row_bytes, each scaled by 2^e so that
// its largest magnitude lies in [224, 448]; inv[row] = 2^-e; norms (nullable)
// of the dequantized rows v~, or (|v~|^2 - |v|^2) / 2 with half_diff (the
// per-row terms of the FP8 inner-product estimate).
void to_fp8(const float *x, int64_t n, int d, uint8_t *out, float *inv, float *norms,
            hipStream_t s, bool half_diff = false);
// How gemm_topk splits a search: the 64-query form (wide; reg: the register
// tile) on S = C x P splits (P > 1: strided, each split's top kw = 64 under a
// per-query bound from a sample pass, with a per-query check and the 16-query
// form for the queries it rejects), or the 16-query form on S contiguous
// splits; ncand candidates per query in a workspace of `bytes`.
struct GemmPlan {
  bool wide, reg, strided;
  int kw, P;
  int64_t S, split_len, ncand;
  size_t bytes;
  // strided: the sample pass that bounds each query's keys (Ps = 0: none),
  // rows 0, Ps, 2 Ps, ... in Ss chunks, top kks per query
  int Ps, kks;
  int64_t Ss, split_len_s;
};
GemmPlan gemm_plan(int kind, int64_t nq, int64_t n, int d, int kk);
// Candidate keys of the GEMM-form distance, part[nq][ncand]: per base split
// the split's sorted top list (sentinel-padded), and the union holds each
// query's exact top kk; reduce with launch_merge over ncand candidates.
void gemm_topk(int kind, const uint8_t *Q, const uint8_t *X, int64_t nq, int64_t n, int d,
               const float *qnorm, const float *xnorm, const uint32_t *qn_u, const uint32_t *qs_u,
               const uint32_t *xn_u, const uint32_t *xs_u, float scale_sq, int kk,
               const uint8_t *excl, int64_t excl_len, DevBuf &part, int64_t &ncand, hipStream_t s,
               const float *qinv = nullptr, const float *xinv = nullptr);

// ---- sq8.hip ----
void sq8_minmax(const float *x, int64_t count, float *out2, DevBuf &tmp, hipStream_t s);
// Device code rows are zero-padded to this stride (a multiple of 64 bytes).
int sq8_row_stride(int d);
// Codes [n][stride] (padding zeroed), UInt32 norms and (nullable) UInt32 code
// sums of n rows.
void sq8_encode(const float *x, int64_t n, int d, int symmetric, float minimum, float inv_scale,
                uint8_t *codes, uint32_t *norms, uint32_t *sums, hipStream_t s);
struct Sq8SearchArgs {
  const float *q;
  int64_t nq;
  int d, metric;
  int64_t n;
  const uint8_t *codes;   // [n][sq8_row_stride(d)]
  const uint32_t *norms;
  const uint32_t *sums;
  const float *vectors;
  float gmin, scale;
  int k, k_factor;
  const uint8_t *excl;
  int64_t excl_len;
  float *D;
  int64_t *I;
  int use_mfma;  // 1: int8 MFMA contraction; 0: dot4 VALU kernel (same results)
  // per-query k (nullable, device [nq]): query i returns kq[i] <= k results
  // in rows of stride k (padded past kq[i])
  const int32_t *kq;
};
struct Sq8Work {
  DevBuf qcodes, qnorms, qsums, part, keys;
};
void sq8_search(const Sq8SearchArgs &a, Sq8Work &w, hipStream_t s);

// ---- pairdist.hip ----
// out[p] = distance of query row (qid ? qid[p] : p / per_query) of qbase to
// row xid[p] of xbase, as FlatDistanceComputer.distance (order 0) or
// distance_batch4 (order 1) computes it; IP distances are negated.
void flat_pair_distances(const float *qbase, const int64_t *qid, int64_t per_query,
                         const float *xbase, const int64_t *xid, int64_t npairs, int d, int metric,
                         int order, float *out, hipStream_t s);
// The SQ8DistanceComputer code distance of encoded rows (stride ld).
void sq8_pair_distances(const uint8_t *qcodes, const uint32_t *qnorms, const int64_t *qid,
                        int64_t per_query, const uint8_t *codes, const uint32_t *norms,
                        const int64_t *xid, int64_t npairs, int ld, int metric, float scale_sq,
                        float *out, hipStream_t s);

}  // namespace mvg
