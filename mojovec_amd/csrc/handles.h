// Internal layout of the IVF-PQ and SQ8 handles, shared by the C ABI (api.cpp)
// and the request loop (server.cpp).
#pragma once

#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/mojovec_gpu.h"
#include "kernels.h"

struct mvg_ivfpq {
  int d = 0, nlist = 0, M = 0, metric = 0, nprobe = 1;
  bool trained = false, keep_vectors = true, dirty = true, tables_dirty = true;
  int precision = 0, split_scan = 0, wq_entries = 0, deep_min = 0, deep_pool = 0;
  // table bits 16: fewest queries of a batch for the tile scan (0: 16 per CU)
  int64_t tile_min_batch = 0;
  std::vector<int64_t> top_sum;  // top_sum[p]: entries in the p largest lists
  mvg::DevBuf ptab, coarse_bf16, coarse_norm, entry_pre;
  bool pre_dirty = true;  // entry_pre out of date with the lists or tables
  mvg::DevBuf entry_pre8, table_mu;  // 8-bit tables: centred entry terms, centre
  bool pre8_ready = false;
  int64_t ntotal = 0;
  std::vector<std::vector<int32_t>> lids;   // inverted lists, append order
  std::vector<std::vector<uint8_t>> lcodes;
  mvg::DevBuf coarse, pq, vectors, list_off, codes, ids;
  mvg::DevBuf tx, tlab, tcodes, tres, q, D, I, excl, npq;
  mvg::IvfWork work;
  mvg::KmeansWork kw;
  // mixed form: SQ8 codes of the stored rows for the pre-rerank (global range
  // as IndexFlatSQ8 computes it); rebuilt when rows were added
  int table_bits = 32;  // mixed L2 scan with a rerank: 8, 16 (FP16 MFMA tiles) or 32
  // table_bits 16: FP16 codebooks of the tile scan (-2 * pq * 2^e, cb16_inv =
  // 2^-e) and each list's rank in the super-cluster order of the centroids
  mvg::DevBuf cb16, list_rank, tile_cent;
  float cb16_inv = 1.0f;
  bool tile_ready = false;
  int prefilter = 8;  // pre-rerank keeps max(64, prefilter * k) candidates; 0: off
  bool vq_dirty = true;
  float vq_min = 0.0f, vq_scale = 1.0f;
  mvg::DevBuf vq_codes, vq_norms, vq_tmp;
  // Explicit ids (add_with_ids): the lists hold the caller's ids while the
  // FP32 rows and their SQ8 codes stay in add order, so the rerank finds a
  // row through row_of[id] (-1: no row). Built once an id differs from its
  // row number while rows are kept.
  bool custom_ids = false, row_of_dirty = false;
  std::vector<int32_t> row_of_host;
  mvg::DevBuf row_of;
  int64_t ncsr = 0;  // entries in the device CSR lists
  // Mutation generation (captured graphs refuse to replay after it moves)
  // and the number of servers holding the handle (mutations and search
  // settings are refused while one does).
  std::atomic<uint64_t> generation{0};
  std::atomic<int> servers{0};
  hipStream_t s = nullptr;
  std::mutex search_mu;  // serialises searches that share `work`
};

// Flat SQ8 handle: FP32 rows (exact rerank), codes [n][sq8_row_stride(d)],
// integer code norms and sums, the global range and scale.
struct mvg_sq8 {
  int d = 0, metric = 0;
  int64_t ntotal = 0;
  float gmin = 3.40282347e38f, gmax = -3.40282347e38f, scale = 1.0f;  // index_flat_sq8.mojo:234-236
  mvg::DevBuf f32, codes, norms, sums, tmp, q, D, I, excl;
  mvg::DevBuf pair_a, pair_b, pair_out;  // batched distance service
  int use_mfma = 1;
  mvg::Sq8Work work;
  hipStream_t s = nullptr;
};

// IVF-Flat handle: host lists in append order (what list accessors and
// snapshots see) and a device CSR copy with each list in id order.
struct mvg_ivfflat {
  int d = 0, nlist = 0, metric = 0, nprobe = 1;
  int quantizer_metric = 0;  // the coarse IndexFlat's own metric
  bool trained = false, dirty = true;
  int64_t ntotal = 0, max_list = 0;
  // Reference storage capacities a snapshot records: the quantizer's rows
  // and each list's (grown by the reference's rules at save time).
  int64_t coarse_capacity = 1024;
  std::vector<int64_t> list_capacity;
  std::vector<std::vector<int32_t>> lids;
  std::vector<std::vector<float>> lrows;
  mvg::DevBuf coarse, list_off, rows, ids;
  mvg::DevBuf tx, tlab, q, D, I, excl;
  mvg::IvfFlatWork work;
  mvg::KmeansWork kw;
  hipStream_t s = nullptr;
  std::mutex search_mu;  // serialises searches that share `work`
};

namespace mvg {

// Brings the device CSR lists and the mixed-form tables up to date (host
// synchronising; call before capture or concurrent searches).
void ivf_prepare(mvg_ivfpq *h);

// Enqueues one search on `s` with the caller's workspace; every pointer is
// device memory. Throws mvg::Error.
// k_per_query (nullable, device [nq]): per-query result counts <= k, in
// output rows of stride k.
void ivf_search_with(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                     const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                     const uint8_t *exclude, int64_t exclude_len, float *distances,
                     int64_t *labels, IvfWork &work, hipStream_t s,
                     const int32_t *k_per_query = nullptr);

// Throws MVG_ERR_STATE while a server holds the handle (mutations and
// search settings would change what its batches read).
void ivf_check_unserved(const mvg_ivfpq *h);

}  // namespace mvg
