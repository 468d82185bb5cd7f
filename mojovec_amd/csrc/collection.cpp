// CollectionIVFPQ (mojovec/api/collection_ivfpq.mojo:37-333): the managed
// IVF-PQ collection above the index handle. Validation order and messages
// follow the reference so that its callers see the same errors.
#include <cmath>
#include <memory>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/mojovec_gpu.h"
#include "collection.h"
#include "common.h"

namespace {

using mvg::check_index;
using mvg::kMetricCosine;
constexpr int kMaxQueryResults = 2048;
constexpr int kPqCentroids = 256;

[[noreturn]] void raise(const char *msg) { mvg::coll_raise(msg); }

// _validate_finite_vectors (collection_vector_query.mojo:17-21).
void validate_finite(const float *x, int64_t count) {
  for (int64_t i = 0; i < count; i++)
    if (!std::isfinite(x[i])) raise("Vector components must be finite.");
}

// _normalize_vectors (collection_vector_query.mojo:24-46): FP64 norm.
std::vector<float> normalize(const float *x, int64_t count, int d) {
  std::vector<float> out((size_t)count);
  const int64_t n = count / d;
  for (int64_t v = 0; v < n; v++) {
    const float *row = x + v * d;
    double sq = 0.0;
    for (int j = 0; j < d; j++) {
      if (!std::isfinite(row[j])) raise("Vector components must be finite.");
      const double w = (double)row[j];
      sq += w * w;
    }
    if (!(sq > 0.0)) raise("Cosine embeddings must be non-zero vectors.");
    const double inv = 1.0 / std::sqrt(sq);
    for (int j = 0; j < d; j++) out[(size_t)(v * d + j)] = (float)((double)row[j] * inv);
  }
  return out;
}

}  // namespace

void mvg::validate_ivfpq_parameters(int64_t d, int64_t nlist, int64_t M, int64_t nprobe) {
  if (d <= 0 || d > 65536) raise("dimension must be between 1 and 65536.");
  if (nlist <= 0 || nlist > kCollMaxNlist) raise("nlist must be between 1 and 1000000.");
  if (M <= 0 || M > d || d % M != 0) raise("M must divide dimension and be between 1 and dimension.");
  if (nprobe <= 0 || nprobe > nlist || nprobe > kCollMaxNprobe)
    raise("nprobe must be between 1 and nlist.");
}

int mvg::parse_metric(const std::string &m) {
  if (m == "l2") return MVG_METRIC_L2;
  if (m == "cosine") return kMetricCosine;
  if (m == "ip") return MVG_METRIC_IP;
  raise("metric must be \"l2\", \"cosine\", or \"ip\".");
}

const char *mvg::metric_name(int metric) {
  return metric == MVG_METRIC_L2 ? "l2" : metric == kMetricCosine ? "cosine" : "ip";
}

namespace {

int train_impl(mvg_collection *h, const float *x, int64_t count) {
  if (count % h->d != 0) raise("Embeddings length must be a multiple of collection dimension.");
  const int64_t n = count / h->d;
  const int64_t minimum = std::max(h->nlist, kPqCentroids);
  if (n < minimum) raise("IVF-PQ training requires at least max(nlist, 256) vectors.");
  int trained = 0;
  check_index(mvg_ivfpq_info(h->ix, nullptr, &trained, nullptr));
  if (trained) return MVG_OK;
  if (h->metric == kMetricCosine) {
    std::vector<float> nv = normalize(x, count, h->d);
    check_index(mvg_ivfpq_train(h->ix, nv.data(), n, h->niter_coarse, h->niter_pq, h->seed));
    return MVG_OK;
  }
  validate_finite(x, count);
  check_index(mvg_ivfpq_train(h->ix, x, n, h->niter_coarse, h->niter_pq, h->seed));
  return MVG_OK;
}

}  // namespace

extern "C" {

int mvg_coll_create(int d, int nlist, int M, int nprobe, const char *name, const char *metric,
                    mvg_collection **out) {
  COLL_BEGIN
  const int eff_nprobe = nprobe == 0 ? std::min(10, nlist) : nprobe;  // :77
  mvg::validate_ivfpq_parameters(d, nlist, M, eff_nprobe);
  const int mt = mvg::parse_metric(metric ? metric : "l2");
  auto h = std::make_unique<mvg_collection>();
  h->name = name ? name : "";
  h->d = d;
  h->metric = mt;
  h->nlist = nlist;
  h->M = M;
  h->list_capacity.assign((size_t)nlist, 0);
  check_index(mvg_ivfpq_create(d, nlist, M, mt == MVG_METRIC_L2 ? MVG_METRIC_L2 : MVG_METRIC_IP, &h->ix));
  check_index(mvg_ivfpq_set_nprobe(h->ix, eff_nprobe));
  *out = h.release();
  COLL_END
}

int mvg_coll_train(mvg_collection *h, const float *x, int64_t count) {
  COLL_BEGIN
  return train_impl(h, x, count);
  COLL_END
}

int mvg_coll_add(mvg_collection *h, const int64_t *ids, int64_t n, const float *x, int64_t count) {
  COLL_BEGIN
  if (count != n * h->d) raise("Embeddings length must equal len(ids) * dimension.");
  if (n == 0) return MVG_OK;
  std::unordered_set<int64_t> batch;  // _validate_new_ids (:197-204)
  for (int64_t i = 0; i < n; i++) {
    if (batch.count(ids[i])) raise("Duplicate IDs are not allowed within an add batch.");
    if (h->id_to_internal.count(ids[i])) raise("Cannot add an ID that already exists.");
    batch.insert(ids[i]);
  }
  int trained = 0;
  check_index(mvg_ivfpq_info(h->ix, nullptr, &trained, nullptr));
  if (!trained) train_impl(h, x, count);  // :228-229
  if (h->metric == kMetricCosine) {
    std::vector<float> nv = normalize(x, count, h->d);
    check_index(mvg_ivfpq_add(h->ix, nv.data(), n));
  } else {
    validate_finite(x, count);
    check_index(mvg_ivfpq_add(h->ix, x, n));
  }
  for (int64_t i = 0; i < n; i++) {
    h->id_to_internal[ids[i]] = (int64_t)h->user_ids.size();
    h->user_ids.push_back(ids[i]);
  }
  COLL_END
}

int mvg_coll_query(mvg_collection *h, const float *q, int64_t count, int n_results, int k_factor,
                   int64_t *ids, float *distances, int64_t *nq_out) {
  COLL_BEGIN
  if (count % h->d != 0) raise("Query embeddings length must be a multiple of dimension.");
  if (n_results <= 0 || n_results > kMaxQueryResults) raise("n_results must be between 1 and 2048.");
  if (k_factor < 1) raise("k_factor must be at least 1.");
  const int64_t nq = count / h->d;
  if (nq_out) *nq_out = nq;
  if (nq == 0) return MVG_OK;
  std::vector<float> nv;
  const float *src = q;
  if (h->metric == kMetricCosine) {
    nv = normalize(q, count, h->d);
    src = nv.data();
  } else {
    validate_finite(q, count);
  }
  h->D.resize((size_t)(nq * n_results));
  h->I.resize((size_t)(nq * n_results));
  check_index(mvg_ivfpq_search(h->ix, src, nq, n_results, nullptr, k_factor, nullptr, 0,
                               h->D.data(), h->I.data()));
  const int64_t count_ids = (int64_t)h->user_ids.size();
  for (int64_t i = 0; i < nq * n_results; i++) {  // _query_prepared (:310-323)
    const int64_t internal = h->I[(size_t)i];
    float dist = h->D[(size_t)i];
    if (internal >= 0 && internal < count_ids) {
      ids[i] = h->user_ids[(size_t)internal];
      if (h->metric != MVG_METRIC_L2) dist = 1.0f - dist;
    } else {
      ids[i] = -1;
    }
    distances[i] = dist;
  }
  COLL_END
}

int mvg_coll_set_nprobe(mvg_collection *h, int nprobe) {
  COLL_BEGIN
  if (nprobe <= 0 || nprobe > h->nlist) raise("nprobe must be between 1 and nlist.");  // :152-155
  check_index(mvg_ivfpq_set_nprobe(h->ix, nprobe));
  COLL_END
}

int mvg_coll_set_train_params(mvg_collection *h, int niter_coarse, int niter_pq, uint64_t seed) {
  COLL_BEGIN
  if (niter_coarse < 0 || niter_pq < 0) raise("k-means iteration counts must be non-negative.");
  h->niter_coarse = niter_coarse;
  h->niter_pq = niter_pq;
  h->seed = seed;
  COLL_END
}

int mvg_coll_info(mvg_collection *h, int *dimension, int *metric, int64_t *count, int *nlist,
                  int *M, int *nprobe, int *trained) {
  COLL_BEGIN
  int np = 0, tr = 0;
  check_index(mvg_ivfpq_info(h->ix, nullptr, &tr, &np));
  if (dimension) *dimension = h->d;
  if (metric) *metric = h->metric;
  if (count) *count = (int64_t)h->user_ids.size();
  if (nlist) *nlist = h->nlist;
  if (M) *M = h->M;
  if (nprobe) *nprobe = np;
  if (trained) *trained = tr;
  COLL_END
}

int mvg_coll_name(mvg_collection *h, char *buf, int buf_len) {
  COLL_BEGIN
  if (buf && buf_len > 0) {
    std::strncpy(buf, h->name.c_str(), (size_t)buf_len - 1);
    buf[buf_len - 1] = 0;
  }
  COLL_END
}

mvg_ivfpq *mvg_coll_index(mvg_collection *h) { return h ? h->ix : nullptr; }

void mvg_coll_destroy(mvg_collection *h) {
  delete h;
}

}  // extern "C"
