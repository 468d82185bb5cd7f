// C ABI (include/mojovec_gpu.h): handles, validation, host<->device staging.
#include "../../include/mojovec_gpu.h"

#include <algorithm>
#include <functional>
#include <cmath>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <unordered_set>

#include "handles.h"
#include "kernels.h"

using mvg::DevBuf;
using mvg::Error;

namespace {

thread_local std::string g_last_error;

int record(int code, const char *msg) {
  g_last_error = msg;
  return code;
}

#define API_BEGIN try {
#define API_END                                                 \
  }                                                             \
  catch (const mvg::Error &e) {                                 \
    return record(e.code, e.what());                            \
  }                                                             \
  catch (const std::bad_alloc &) {                              \
    return record(mvg::kAlloc, "host allocation failed");       \
  }                                                             \
  catch (const std::exception &e) {                             \
    return record(mvg::kDevice, e.what());                      \
  }                                                             \
  return MVG_OK;

void require(bool ok, const char *msg) {
  if (!ok) throw Error(mvg::kInvalid, msg);
}

hipStream_t new_stream() {
  hipStream_t s;
  MVG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

template <typename T>
T *upload(DevBuf &b, const T *src, size_t count, hipStream_t s) {
  b.reserve(sizeof(T) * std::max<size_t>(count, 1));
  if (count) MVG_HIP(hipMemcpyAsync(b.ptr, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
  return b.as<T>();
}

template <typename T>
void download(T *dst, const void *src, size_t count, hipStream_t s) {
  if (count) MVG_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, s));
}

// Philox-4x32-10 (Salmon et al., SC'11) keyed by the 64-bit seed, counter
// (lo, hi, 0, 0): the generator behind the seeded initial centroid picks of
// kmeans.mojo:67-82.
void philox(uint64_t seed, uint64_t counter, uint32_t out[4]) {
  uint32_t c0 = (uint32_t)counter, c1 = (uint32_t)(counter >> 32), c2 = 0, c3 = 0;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Row indices of the initial centroids and the reseed offset
// (kmeans.mojo:70-82).
int64_t kmeans_pick(int64_t n, int k, uint64_t seed, std::vector<int64_t> &rows) {
  uint32_t w[4];
  uint64_t ctr = 0;
  philox(seed, ctr++, w);
  const int64_t offset = (int64_t)(w[0] % (uint32_t)n);
  rows.resize(k);
  for (int i = 0; i < k; i++) {
    if (i != 0 && i % 4 == 0) philox(seed, ctr++, w);
    rows[i] = (int64_t)(w[i % 4] % (uint32_t)n);
  }
  return offset;
}

constexpr uint64_t kDefaultSeed = 0x4D4F4A4F564543ull;  // kmeans.mojo:70

// Device-side PQ training on rows already in HBM ([n x d]).
void pq_train_device(const float *dx, int64_t n, int d, int M, int niter, uint64_t seed,
                     float *dcb, mvg::KmeansWork &kw, hipStream_t s) {
  const int dsub = d / M;
  std::vector<int64_t> rows;
  const int64_t off = kmeans_pick(n, mvg::kKsub, seed, rows);
  std::vector<float> picked((size_t)mvg::kKsub * d), cb((size_t)mvg::kKsub * d);
  for (int c = 0; c < mvg::kKsub; c++)
    MVG_HIP(hipMemcpyAsync(picked.data() + (size_t)c * d, dx + rows[c] * d, sizeof(float) * d,
                           hipMemcpyDeviceToHost, s));
  MVG_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < M; m++)
    for (int c = 0; c < mvg::kKsub; c++)
      for (int j = 0; j < dsub; j++)
        cb[((size_t)m * mvg::kKsub + c) * dsub + j] = picked[(size_t)c * d + m * dsub + j];
  MVG_HIP(hipMemcpyAsync(dcb, cb.data(), sizeof(float) * cb.size(), hipMemcpyHostToDevice, s));
  mvg::kmeans_fit(dx, n, M, dsub, mvg::kKsub, niter, off, dcb, kw, s);
}

}  // namespace

void mvg::set_last_error(const std::string &msg) { g_last_error = msg; }

// ===========================================================================
// Handles

struct mvg_flat {
  int d = 0, metric = 0;
  int64_t ntotal = 0;
  int precision = 0, k_factor = 1;  // precision 0 FP32, 1 BF16 candidates, 2 FP8 E4M3 candidates
  bool bf16_dirty = true, fp8_dirty = true;
  DevBuf base, part, q, D, I, excl, base_bf16, base_norm, qbf, qnorm, keys;
  DevBuf base_fp8, base_fp8_inv, base_fp8_norm, qinv;
  DevBuf pair_a, pair_b, pair_out;  // batched distance service
  hipStream_t s = nullptr;
};

namespace {

// Rejects ids outside [0, n): the distance service reads rows by id.
void check_ids(const int64_t *ids, int64_t count, int64_t n, const char *what) {
  for (int64_t i = 0; i < count; i++)
    if (ids[i] < 0 || ids[i] >= n)
      throw Error(mvg::kInvalid, std::string(what) + " must lie in [0, ntotal)");
}

void ivf_sync_lists(mvg_ivfpq *h) {
  if (!h->dirty) return;
  std::vector<int64_t> off(h->nlist + 1, 0);
  for (int l = 0; l < h->nlist; l++) off[l + 1] = off[l] + (int64_t)h->lids[l].size();
  const int64_t total = off[h->nlist];
  std::vector<int32_t> ids((size_t)std::max<int64_t>(total, 1));
  std::vector<uint8_t> codes((size_t)std::max<int64_t>(total * h->M, 1));
  for (int l = 0; l < h->nlist; l++) {
    std::copy(h->lids[l].begin(), h->lids[l].end(), ids.begin() + off[l]);
    std::copy(h->lcodes[l].begin(), h->lcodes[l].end(), codes.begin() + off[l] * h->M);
  }
  std::vector<int64_t> sizes(h->nlist);
  for (int l = 0; l < h->nlist; l++) sizes[l] = off[l + 1] - off[l];
  std::sort(sizes.begin(), sizes.end(), std::greater<int64_t>());
  h->top_sum.assign(h->nlist + 1, 0);
  for (int l = 0; l < h->nlist; l++) h->top_sum[l + 1] = h->top_sum[l] + sizes[l];
  upload(h->list_off, off.data(), off.size(), h->s);
  h->ncsr = total;
  // +16 bytes so 16-byte code loads of the last entry stay inside the buffer.
  h->codes.reserve(codes.size() + 16);
  MVG_HIP(hipMemcpyAsync(h->codes.ptr, codes.data(), codes.size(), hipMemcpyHostToDevice, h->s));
  upload(h->ids, ids.data(), ids.size(), h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  h->dirty = false;
  h->pre_dirty = true;
}

}  // namespace

extern "C" {

const char *mvg_last_error(void) { return g_last_error.c_str(); }

const char *mvg_version(void) { return "mojovec-gpu 0.1.0 (gfx950)"; }

int mvg_device_info(char *name, int name_len, int *compute_units) {
  API_BEGIN
  int dev = 0;
  MVG_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  MVG_HIP(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    std::strncpy(name, p.gcnArchName, (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (compute_units) *compute_units = p.multiProcessorCount;
  API_END
}

// ---- Flat ------------------------------------------------------------------

int mvg_flat_create(int d, int metric, mvg_flat **out) {
  API_BEGIN
  require(d > 0 && d <= 65536, "dimension must be between 1 and 65536.");
  require(metric == MVG_METRIC_L2 || metric == MVG_METRIC_IP, "metric must be 0 (L2) or 1 (IP).");
  auto h = std::make_unique<mvg_flat>();
  h->d = d;
  h->metric = metric;
  h->s = new_stream();
  *out = h.release();
  API_END
}

int mvg_flat_add(mvg_flat *h, const float *x, int64_t n) {
  API_BEGIN
  require(h && n >= 0, "invalid arguments");
  if (n == 0) return MVG_OK;
  const size_t keep = sizeof(float) * (size_t)h->ntotal * h->d;
  const size_t need = keep + sizeof(float) * (size_t)n * h->d;
  h->base.grow_keep(std::max(need, h->base.bytes * 2), keep, h->s);
  MVG_HIP(hipMemcpyAsync((char *)h->base.ptr + keep, x, need - keep, hipMemcpyHostToDevice, h->s));
  MVG_HIP(hipStreamSynchronize(h->s));
  h->ntotal += n;
  h->bf16_dirty = h->fp8_dirty = true;
  API_END
}

int mvg_flat_search(mvg_flat *h, const float *q, int64_t nq, int k, const uint8_t *exclude,
                    int64_t exclude_len, float *distances, int64_t *labels) {
  API_BEGIN
  require(h && nq >= 0 && k >= 1, "invalid arguments");
  if (nq == 0) return MVG_OK;
  const float *dq = upload(h->q, q, (size_t)nq * h->d, h->s);
  const uint8_t *dx = exclude && exclude_len > 0 ? upload(h->excl, exclude, (size_t)exclude_len, h->s) : nullptr;
  h->D.reserve(sizeof(float) * nq * k);
  h->I.reserve(sizeof(int64_t) * nq * k);
  if (h->precision == 0 || h->ntotal == 0) {
    mvg::flat_topk(h->base.as<float>(), h->ntotal, h->d, h->metric, dq, nq, k, dx,
                   dx ? exclude_len : 0, h->part, 0.0f, h->D.as<float>(), h->I.as<int64_t>(),
                   nullptr, nullptr, h->s);
  } else {
    // mixed: BF16 or FP8 E4M3 GEMM-form candidates on MFMA (FP32 accumulate),
    // optional exact FP32 rerank
    const bool ip = h->metric == MVG_METRIC_IP;
    const bool fp8 = h->precision == 2;
    const int kind = fp8 ? (ip ? mvg::kGemmFp8IP : mvg::kGemmFp8L2) : (ip ? mvg::kGemmBf16IP : mvg::kGemmBf16L2);
    const int rb = mvg::gemm_row_bytes(kind, h->d);
    const uint8_t *xrows, *qrows;
    const float *xnorm, *xinv = nullptr, *qinv = nullptr;
    h->qnorm.reserve(sizeof(float) * nq);
    h->qbf.reserve((size_t)nq * rb);
    if (fp8) {
      if (h->fp8_dirty) {
        h->base_fp8.reserve((size_t)h->ntotal * rb);
        h->base_fp8_inv.reserve(sizeof(float) * h->ntotal);
        h->base_fp8_norm.reserve(sizeof(float) * h->ntotal);
        mvg::to_fp8(h->base.as<float>(), h->ntotal, h->d, h->base_fp8.as<uint8_t>(),
                    h->base_fp8_inv.as<float>(), h->base_fp8_norm.as<float>(), h->s, ip);
        h->fp8_dirty = false;
      }
      h->qinv.reserve(sizeof(float) * nq);
      mvg::to_fp8(dq, nq, h->d, h->qbf.as<uint8_t>(), h->qinv.as<float>(), h->qnorm.as<float>(), h->s,
                  ip);
      xrows = h->base_fp8.as<uint8_t>();
      xnorm = h->base_fp8_norm.as<float>();
      xinv = h->base_fp8_inv.as<float>();
      qinv = h->qinv.as<float>();
    } else {
      if (h->bf16_dirty) {
        h->base_bf16.reserve((size_t)h->ntotal * rb);
        h->base_norm.reserve(sizeof(float) * h->ntotal);
        mvg::to_bf16(h->base.as<float>(), h->ntotal, h->d, h->base_bf16.as<uint16_t>(),
                     h->base_norm.as<float>(), h->s);
        h->bf16_dirty = false;
      }
      mvg::to_bf16(dq, nq, h->d, h->qbf.as<uint16_t>(), h->qnorm.as<float>(), h->s);
      xrows = h->base_bf16.as<uint8_t>();
      xnorm = h->base_norm.as<float>();
    }
    qrows = h->qbf.as<uint8_t>();
    const int kf = h->k_factor;
    const int kk = k * kf;
    if (kk > mvg::kMaxCandidates) throw Error(mvg::kInvalid, "k * k_factor exceeds the device limit of 4096");
    int64_t ncand = 0;
    mvg::gemm_topk(kind, qrows, xrows, nq, h->ntotal, h->d, h->qnorm.as<float>(), xnorm, nullptr,
                   nullptr, nullptr, nullptr, 0.0f, kk, dx, dx ? exclude_len : 0, h->part, ncand,
                   h->s, qinv, xinv);
    if (kf == 1) {
      mvg::launch_merge(h->part.as<uint64_t>(), nq, ncand, k, ip ? 1 : 0, 0.0f,
                        h->D.as<float>(), h->I.as<int64_t>(), nullptr, nullptr, h->s);
    } else {
      h->keys.reserve(sizeof(uint64_t) * nq * kk);
      mvg::launch_merge(h->part.as<uint64_t>(), nq, ncand, kk, 0, 0.0f, nullptr,
                        nullptr, nullptr, h->keys.as<uint64_t>(), h->s);
      mvg::rerank(dq, nq, h->d, h->metric, h->base.as<float>(), h->ntotal, h->keys.as<uint64_t>(), kk, k, 0.0f,
                  h->D.as<float>(), h->I.as<int64_t>(), h->s);
    }
  }
  download(distances, h->D.ptr, (size_t)nq * k, h->s);
  download(labels, h->I.ptr, (size_t)nq * k, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_flat_set_search_mode(mvg_flat *h, int precision, int k_factor) {
  API_BEGIN
  require(h && precision >= 0 && precision <= 2 && k_factor >= 1, "invalid search mode");
  h->precision = precision;
  h->k_factor = k_factor;
  API_END
}

int mvg_flat_distances(mvg_flat *h, const float *q, int64_t nq, const int64_t *ids,
                       int64_t per_query, int order, float *out) {
  API_BEGIN
  require(h && nq >= 0 && per_query >= 0, "invalid arguments");
  require(order == 0 || order == 1, "order must be 0 (distance) or 1 (distance_batch4).");
  const int64_t np = nq * per_query;
  if (np == 0) return MVG_OK;
  require(q && ids && out, "invalid arguments");
  check_ids(ids, np, h->ntotal, "ids");
  const float *dq = upload(h->q, q, (size_t)nq * h->d, h->s);
  const int64_t *di = upload(h->pair_a, ids, (size_t)np, h->s);
  h->pair_out.reserve(sizeof(float) * np);
  mvg::flat_pair_distances(dq, nullptr, per_query, h->base.as<float>(), di, np, h->d, h->metric,
                           order, h->pair_out.as<float>(), h->s);
  download(out, h->pair_out.ptr, (size_t)np, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_flat_pair_distances(mvg_flat *h, const int64_t *a, const int64_t *b, int64_t npairs,
                            float *out) {
  API_BEGIN
  require(h && npairs >= 0, "invalid arguments");
  if (npairs == 0) return MVG_OK;
  require(a && b && out, "invalid arguments");
  check_ids(a, npairs, h->ntotal, "ids");
  check_ids(b, npairs, h->ntotal, "ids");
  const int64_t *da = upload(h->pair_a, a, (size_t)npairs, h->s);
  const int64_t *db = upload(h->pair_b, b, (size_t)npairs, h->s);
  h->pair_out.reserve(sizeof(float) * npairs);
  mvg::flat_pair_distances(h->base.as<float>(), da, 0, h->base.as<float>(), db, npairs, h->d,
                           h->metric, 0, h->pair_out.as<float>(), h->s);
  download(out, h->pair_out.ptr, (size_t)npairs, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int64_t mvg_flat_ntotal(const mvg_flat *h) { return h ? h->ntotal : 0; }

}  // extern "C"

namespace {

// Stored FP32 rows ids[0..n) of a Flat or SQ8 handle into out [n][d].
void read_rows(const float *base, int d, int64_t ntotal, const int64_t *ids, int64_t n, float *out,
               DevBuf &dids, DevBuf &dout, hipStream_t s) {
  require(n >= 0 && (n == 0 || (ids && out)), "invalid arguments");
  check_ids(ids, n, ntotal, "ids");
  if (n == 0) return;
  const int64_t *di = upload(dids, ids, (size_t)n, s);
  dout.reserve(sizeof(float) * (size_t)n * d);
  mvg::gather_rows(base, d, di, n, dout.as<float>(), s);
  download(out, dout.ptr, (size_t)n * d, s);
  MVG_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

int mvg_flat_get_vectors(mvg_flat *h, const int64_t *ids, int64_t n, float *out) {
  API_BEGIN
  require(h, "invalid handle");
  read_rows(h->base.as<float>(), h->d, h->ntotal, ids, n, out, h->pair_a, h->pair_out, h->s);
  API_END
}

int mvg_flat_get_vector(mvg_flat *h, int64_t id, float *out) {
  return mvg_flat_get_vectors(h, &id, 1, out);
}

void mvg_flat_destroy(mvg_flat *h) {
  if (!h) return;
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

// ---- SQ8 ---------------------------------------------------------------------

int mvg_sq8_create(int d, int metric, mvg_sq8 **out) {
  API_BEGIN
  require(d > 0 && d <= 65536, "dimension must be between 1 and 65536.");
  require(metric == MVG_METRIC_L2 || metric == MVG_METRIC_IP, "metric must be 0 (L2) or 1 (IP).");
  auto h = std::make_unique<mvg_sq8>();
  h->d = d;
  h->metric = metric;
  h->s = new_stream();
  *out = h.release();
  API_END
}

int mvg_sq8_add(mvg_sq8 *h, const float *x, int64_t n) {
  API_BEGIN
  require(h && n >= 0, "invalid arguments");
  if (n == 0) return MVG_OK;
  const int d = h->d;
  const size_t keep = sizeof(float) * (size_t)h->ntotal * d;
  const size_t need = keep + sizeof(float) * (size_t)n * d;
  const size_t cap_rows = std::max<size_t>(need / (sizeof(float) * d), h->f32.bytes / (sizeof(float) * d) * 2);
  const int ld = mvg::sq8_row_stride(d);
  h->f32.grow_keep(cap_rows * d * sizeof(float), keep, h->s);
  h->codes.grow_keep(cap_rows * ld + 16, (size_t)h->ntotal * ld, h->s);
  h->norms.grow_keep(cap_rows * sizeof(uint32_t), (size_t)h->ntotal * sizeof(uint32_t), h->s);
  h->sums.grow_keep(cap_rows * sizeof(uint32_t), (size_t)h->ntotal * sizeof(uint32_t), h->s);
  float *dst = (float *)((char *)h->f32.ptr + keep);
  MVG_HIP(hipMemcpyAsync(dst, x, need - keep, hipMemcpyHostToDevice, h->s));
  float mm[2];  // index_flat_sq8.mojo:317-331
  mvg::sq8_minmax(dst, n * d, mm, h->tmp, h->s);
  bool requant = false;
  if (mm[0] < h->gmin) { h->gmin = mm[0]; requant = true; }
  if (mm[1] > h->gmax) { h->gmax = mm[1]; requant = true; }
  if (h->metric == MVG_METRIC_L2) {  // :333-342
    h->scale = h->gmax == h->gmin ? 1.0f : (h->gmax - h->gmin) / 255.0f;
  } else {
    const float mx = std::max(std::fabs(h->gmin), std::fabs(h->gmax));
    h->scale = mx > 0.0f ? mx / 127.0f : 1.0f;
  }
  const float inv = 1.0f / h->scale;
  const int sym = h->metric != MVG_METRIC_L2;
  const float mn = sym ? 0.0f : h->gmin;
  if (requant && h->ntotal > 0)  // :346-367
    mvg::sq8_encode(h->f32.as<float>(), h->ntotal, d, sym, mn, inv, h->codes.as<uint8_t>(),
                    h->norms.as<uint32_t>(), h->sums.as<uint32_t>(), h->s);
  mvg::sq8_encode(dst, n, d, sym, mn, inv, h->codes.as<uint8_t>() + (size_t)h->ntotal * ld,
                  h->norms.as<uint32_t>() + h->ntotal, h->sums.as<uint32_t>() + h->ntotal, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  h->ntotal += n;
  API_END
}

int mvg_sq8_search(mvg_sq8 *h, const float *q, int64_t nq, int k, int k_factor,
                   const uint8_t *exclude, int64_t exclude_len, float *distances,
                   int64_t *labels) {
  API_BEGIN
  require(h && nq >= 0 && k >= 1 && k_factor >= 1, "invalid arguments");
  if (nq == 0) return MVG_OK;
  mvg::Sq8SearchArgs a{};
  a.q = upload(h->q, q, (size_t)nq * h->d, h->s);
  a.nq = nq;
  a.d = h->d;
  a.metric = h->metric;
  a.n = h->ntotal;
  a.codes = h->codes.as<uint8_t>();
  a.norms = h->norms.as<uint32_t>();
  a.sums = h->sums.as<uint32_t>();
  a.use_mfma = h->use_mfma;
  a.vectors = h->f32.as<float>();
  a.gmin = h->gmin;
  a.scale = h->scale;
  a.k = k;
  a.k_factor = k_factor;
  a.excl = exclude && exclude_len > 0 ? upload(h->excl, exclude, (size_t)exclude_len, h->s) : nullptr;
  a.excl_len = a.excl ? exclude_len : 0;
  h->D.reserve(sizeof(float) * nq * k);
  h->I.reserve(sizeof(int64_t) * nq * k);
  a.D = h->D.as<float>();
  a.I = h->I.as<int64_t>();
  mvg::sq8_search(a, h->work, h->s);
  download(distances, h->D.ptr, (size_t)nq * k, h->s);
  download(labels, h->I.ptr, (size_t)nq * k, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_sq8_distances(mvg_sq8 *h, const float *q, int64_t nq, const int64_t *ids,
                      int64_t per_query, float *out) {
  API_BEGIN
  require(h && nq >= 0 && per_query >= 0, "invalid arguments");
  const int64_t np = nq * per_query;
  if (np == 0) return MVG_OK;
  require(q && ids && out, "invalid arguments");
  check_ids(ids, np, h->ntotal, "ids");
  const bool ip = h->metric == MVG_METRIC_IP;
  const int ld = mvg::sq8_row_stride(h->d);
  const float *dq = upload(h->q, q, (size_t)nq * h->d, h->s);
  h->work.qcodes.reserve((size_t)nq * ld);
  h->work.qnorms.reserve(sizeof(uint32_t) * nq);
  h->work.qsums.reserve(sizeof(uint32_t) * nq);
  // the query codec of get_distance_computer (index_flat_sq8.mojo:536-555)
  mvg::sq8_encode(dq, nq, h->d, ip ? 1 : 0, ip ? 0.0f : h->gmin, 1.0f / h->scale,
                  h->work.qcodes.as<uint8_t>(), h->work.qnorms.as<uint32_t>(),
                  h->work.qsums.as<uint32_t>(), h->s);
  const int64_t *di = upload(h->pair_a, ids, (size_t)np, h->s);
  h->pair_out.reserve(sizeof(float) * np);
  mvg::sq8_pair_distances(h->work.qcodes.as<uint8_t>(), h->work.qnorms.as<uint32_t>(), nullptr,
                          per_query, h->codes.as<uint8_t>(), h->norms.as<uint32_t>(), di, np, ld,
                          h->metric, h->scale * h->scale, h->pair_out.as<float>(), h->s);
  download(out, h->pair_out.ptr, (size_t)np, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_sq8_pair_distances(mvg_sq8 *h, const int64_t *a, const int64_t *b, int64_t npairs,
                           float *out) {
  API_BEGIN
  require(h && npairs >= 0, "invalid arguments");
  if (npairs == 0) return MVG_OK;
  require(a && b && out, "invalid arguments");
  check_ids(a, npairs, h->ntotal, "ids");
  check_ids(b, npairs, h->ntotal, "ids");
  const int64_t *da = upload(h->pair_a, a, (size_t)npairs, h->s);
  const int64_t *db = upload(h->pair_b, b, (size_t)npairs, h->s);
  h->pair_out.reserve(sizeof(float) * npairs);
  mvg::sq8_pair_distances(h->codes.as<uint8_t>(), h->norms.as<uint32_t>(), da, 0,
                          h->codes.as<uint8_t>(), h->norms.as<uint32_t>(), db, npairs,
                          mvg::sq8_row_stride(h->d), h->metric, h->scale * h->scale,
                          h->pair_out.as<float>(), h->s);
  download(out, h->pair_out.ptr, (size_t)npairs, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_sq8_params(const mvg_sq8 *h, float *gmin, float *gmax, float *scale, int64_t *ntotal) {
  API_BEGIN
  require(h, "invalid handle");
  if (gmin) *gmin = h->gmin;
  if (gmax) *gmax = h->gmax;
  if (scale) *scale = h->scale;
  if (ntotal) *ntotal = h->ntotal;
  API_END
}

int mvg_sq8_get_vectors(mvg_sq8 *h, const int64_t *ids, int64_t n, float *out) {
  API_BEGIN
  require(h, "invalid handle");
  read_rows(h->f32.as<float>(), h->d, h->ntotal, ids, n, out, h->pair_a, h->pair_out, h->s);
  API_END
}

int mvg_sq8_get_vector(mvg_sq8 *h, int64_t id, float *out) {
  return mvg_sq8_get_vectors(h, &id, 1, out);
}

int mvg_sq8_get_codes(mvg_sq8 *h, uint8_t *codes, uint32_t *norms) {
  API_BEGIN
  require(h, "invalid handle");
  const int ld = mvg::sq8_row_stride(h->d);
  if (codes && h->ntotal)
    MVG_HIP(hipMemcpy2DAsync(codes, h->d, h->codes.ptr, ld, h->d, h->ntotal, hipMemcpyDeviceToHost, h->s));
  if (norms) download(norms, h->norms.ptr, (size_t)h->ntotal, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_sq8_set_search_mode(mvg_sq8 *h, int use_mfma) {
  API_BEGIN
  require(h, "invalid handle");
  h->use_mfma = use_mfma != 0;
  API_END
}

void mvg_sq8_destroy(mvg_sq8 *h) {
  if (!h) return;
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

// ---- k-means / PQ ------------------------------------------------------------

int mvg_kmeans_train(const float *x, int64_t n, int d, int k, int niter, const float *init,
                     int64_t sample_offset, uint64_t seed, float *centroids, int32_t *assign) {
  API_BEGIN
  require(n > 0 && d > 0 && k > 0 && niter >= 0, "invalid k-means arguments");
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } g{s};
  DevBuf bx, bc;
  mvg::KmeansWork kw;
  const float *dx = upload(bx, x, (size_t)n * d, s);
  std::vector<float> c0((size_t)k * d);
  int64_t off = sample_offset;
  if (init) {
    std::copy(init, init + (size_t)k * d, c0.begin());
  } else {
    std::vector<int64_t> rows;
    off = kmeans_pick(n, k, seed, rows);
    for (int c = 0; c < k; c++) std::copy(x + rows[c] * d, x + rows[c] * d + d, c0.begin() + (size_t)c * d);
  }
  float *dc = upload(bc, c0.data(), c0.size(), s);
  mvg::kmeans_fit(dx, n, 1, d, k, niter, off, dc, kw, s);
  download(centroids, dc, (size_t)k * d, s);
  if (assign && niter > 0) download(assign, kw.assign.ptr, (size_t)n, s);
  MVG_HIP(hipStreamSynchronize(s));
  API_END
}

int mvg_pq_train(const float *x, int64_t n, int d, int M, int niter, uint64_t seed,
                 float *codebooks) {
  API_BEGIN
  require(d > 0, "PQ dimension must be positive.");
  require(M > 0 && M <= d && d % M == 0, "PQ M must divide dimension and be between 1 and dimension.");
  require(n > 0, "PQ training needs at least one vector");
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } g{s};
  DevBuf bx, bcb;
  mvg::KmeansWork kw;
  const float *dx = upload(bx, x, (size_t)n * d, s);
  bcb.reserve(sizeof(float) * (size_t)d * mvg::kKsub);
  pq_train_device(dx, n, d, M, niter, seed, bcb.as<float>(), kw, s);
  download(codebooks, bcb.ptr, (size_t)d * mvg::kKsub, s);
  MVG_HIP(hipStreamSynchronize(s));
  API_END
}

int mvg_fp8_encode(const float *x, int64_t n, int d, uint8_t *codes, float *inv_scale,
                   float *norms) {
  API_BEGIN
  require(n >= 0 && d > 0 && codes && inv_scale, "invalid arguments");
  if (n == 0) return MVG_OK;
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } g{s};
  const int rb = mvg::gemm_row_bytes(mvg::kGemmFp8L2, d);
  DevBuf bx, bc, bi, bn;
  const float *dx = upload(bx, x, (size_t)n * d, s);
  bc.reserve((size_t)n * rb);
  bi.reserve(sizeof(float) * n);
  bn.reserve(sizeof(float) * n);
  mvg::to_fp8(dx, n, d, bc.as<uint8_t>(), bi.as<float>(), bn.as<float>(), s);
  download(codes, bc.ptr, (size_t)n * rb, s);
  download(inv_scale, bi.ptr, (size_t)n, s);
  if (norms) download(norms, bn.ptr, (size_t)n, s);
  MVG_HIP(hipStreamSynchronize(s));
  API_END
}

int mvg_pq_encode(const float *x, int64_t n, int d, int M, const float *codebooks, uint8_t *codes) {
  API_BEGIN
  require(M > 0 && M <= d && d % M == 0, "PQ M must divide dimension and be between 1 and dimension.");
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } g{s};
  DevBuf bx, bcb, bcodes;
  const float *dx = upload(bx, x, (size_t)n * d, s);
  const float *dcb = upload(bcb, codebooks, (size_t)d * mvg::kKsub, s);
  bcodes.reserve((size_t)std::max<int64_t>(n * M, 1));
  mvg::pq_assign(dx, n, d, M, dcb, nullptr, nullptr, bcodes.as<uint8_t>(), nullptr, s);
  download(codes, bcodes.ptr, (size_t)n * M, s);
  MVG_HIP(hipStreamSynchronize(s));
  API_END
}

int mvg_pq_decode(const uint8_t *codes, int64_t n, int d, int M, const float *codebooks, float *out) {
  API_BEGIN
  require(M > 0 && M <= d && d % M == 0, "PQ M must divide dimension and be between 1 and dimension.");
  require(n >= 0 && (n == 0 || (codes && codebooks && out)), "invalid arguments");
  if (n == 0) return MVG_OK;
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } g{s};
  DevBuf bc, bcb, bo;
  const uint8_t *dc = upload(bc, codes, (size_t)n * M, s);
  const float *dcb = upload(bcb, codebooks, (size_t)d * mvg::kKsub, s);
  bo.reserve(sizeof(float) * (size_t)n * d);
  mvg::pq_decode(dc, n, d, M, dcb, bo.as<float>(), s);
  download(out, bo.ptr, (size_t)n * d, s);
  MVG_HIP(hipStreamSynchronize(s));
  API_END
}

int mvg_pq_lut(const float *q, int64_t nq, int d, int M, const float *codebooks, int metric,
               float *lut) {
  API_BEGIN
  require(M > 0 && M <= d && d % M == 0, "PQ M must divide dimension and be between 1 and dimension.");
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } g{s};
  DevBuf bq, bcb, bl;
  const float *dq = upload(bq, q, (size_t)nq * d, s);
  const float *dcb = upload(bcb, codebooks, (size_t)d * mvg::kKsub, s);
  bl.reserve(sizeof(float) * (size_t)std::max<int64_t>(nq, 1) * M * mvg::kKsub);
  mvg::pq_lut(dq, nq, d, M, dcb, metric, bl.as<float>(), s);
  download(lut, bl.ptr, (size_t)nq * M * mvg::kKsub, s);
  MVG_HIP(hipStreamSynchronize(s));
  API_END
}

// ---- IVF-PQ ------------------------------------------------------------------

int mvg_ivfpq_create(int d, int nlist, int M, int metric, mvg_ivfpq **out) {
  API_BEGIN
  require(d > 0, "IVF-PQ dimension must be positive.");  // index_ivf_pq.mojo:52-59
  require(nlist > 0 && nlist <= 1000000, "IVF-PQ nlist must be between 1 and 1000000.");
  require(M > 0 && M <= d && d % M == 0,
          "IVF-PQ M must divide dimension and be between 1 and dimension.");
  require(metric == MVG_METRIC_L2 || metric == MVG_METRIC_IP, "metric must be 0 (L2) or 1 (IP).");
  require(d / M <= 64, "PQ subvector dimension above 64 is not supported");
  auto h = std::make_unique<mvg_ivfpq>();
  h->d = d;
  h->nlist = nlist;
  h->M = M;
  h->metric = metric;
  h->lids.resize(nlist);
  h->lcodes.resize(nlist);
  h->s = new_stream();
  h->coarse.reserve(sizeof(float) * (size_t)nlist * d);
  h->pq.reserve(sizeof(float) * (size_t)d * mvg::kKsub);
  *out = h.release();
  API_END
}

int mvg_ivfpq_set_keep_vectors(mvg_ivfpq *h, int keep) {
  API_BEGIN
  require(h, "invalid handle");
  mvg::ivf_check_unserved(h);
  if (h->ntotal > 0 && (keep != 0) != h->keep_vectors)
    throw Error(mvg::kState, "keep_vectors can only change before the first add");
  h->keep_vectors = keep != 0;
  API_END
}

int mvg_ivfpq_train(mvg_ivfpq *h, const float *x, int64_t n, int niter_coarse, int niter_pq,
                    uint64_t seed) {
  API_BEGIN
  require(h && n >= 0, "invalid arguments");
  mvg::ivf_check_unserved(h);
  if (h->trained || n == 0) return MVG_OK;  // index_ivf_pq.mojo:93
  const int d = h->d;
  float *dx = upload(h->tx, x, (size_t)n * d, h->s);
  // 1. coarse k-means (:97-103)
  std::vector<int64_t> rows;
  const int64_t off = kmeans_pick(n, h->nlist, seed, rows);
  std::vector<float> c0((size_t)h->nlist * d);
  for (int c = 0; c < h->nlist; c++) std::copy(x + rows[c] * d, x + rows[c] * d + d, c0.begin() + (size_t)c * d);
  float *dc = h->coarse.as<float>();
  MVG_HIP(hipMemcpyAsync(dc, c0.data(), sizeof(float) * c0.size(), hipMemcpyHostToDevice, h->s));
  mvg::kmeans_fit(dx, n, 1, d, h->nlist, niter_coarse, off, dc, h->kw, h->s);
  // 2. residuals against quantizer.search(x, 1) (:106-128), PQ (:130-132)
  h->tlab.reserve(sizeof(int32_t) * n);
  mvg::flat_argmin(dc, h->nlist, d, h->metric, false, dx, n, h->tlab.as<int32_t>(), nullptr, h->s);
  h->tres.reserve(sizeof(float) * (size_t)n * d);
  mvg::residuals(dx, n, d, dc, h->tlab.as<int32_t>(), h->tres.as<float>(), h->s);
  pq_train_device(h->tres.as<float>(), n, d, h->M, niter_pq, seed, h->pq.as<float>(), h->kw, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  h->trained = true;
  h->tables_dirty = true;
  h->generation++;
  API_END
}

int mvg_ivfpq_set_codebooks(mvg_ivfpq *h, const float *coarse, const float *pq) {
  API_BEGIN
  require(h && coarse && pq, "invalid arguments");
  mvg::ivf_check_unserved(h);
  MVG_HIP(hipMemcpyAsync(h->coarse.ptr, coarse, sizeof(float) * (size_t)h->nlist * h->d,
                         hipMemcpyHostToDevice, h->s));
  MVG_HIP(hipMemcpyAsync(h->pq.ptr, pq, sizeof(float) * (size_t)h->d * mvg::kKsub,
                         hipMemcpyHostToDevice, h->s));
  MVG_HIP(hipStreamSynchronize(h->s));
  h->trained = true;
  h->tables_dirty = true;
  h->generation++;
  API_END
}

int mvg_ivfpq_get_codebooks(mvg_ivfpq *h, float *coarse, float *pq) {
  API_BEGIN
  require(h, "invalid handle");
  if (coarse) download(coarse, h->coarse.ptr, (size_t)h->nlist * h->d, h->s);
  if (pq) download(pq, h->pq.ptr, (size_t)h->d * mvg::kKsub, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_ivfpq_add(mvg_ivfpq *h, const float *x, int64_t n) {
  return mvg_ivfpq_add_with_ids(h, x, n, nullptr);
}

int mvg_ivfpq_add_with_ids(mvg_ivfpq *h, const float *x, int64_t n, const int64_t *ids) {
  API_BEGIN
  require(h && n >= 0, "invalid arguments");
  mvg::ivf_check_unserved(h);
  if (!h->trained || n == 0) return MVG_OK;  // index_ivf_pq.mojo:161
  if (h->ntotal + n > INT32_MAX) throw Error(mvg::kInvalid, "IVF-PQ holds at most 2^31-1 vectors");
  bool explicit_rows = false;  // some id differs from its row number
  if (ids) {
    for (int64_t i = 0; i < n; i++) {
      require(ids[i] >= 0 && ids[i] <= INT32_MAX, "IVF-PQ ids must be in [0, 2^31 - 1]");
      explicit_rows |= ids[i] != h->ntotal + i;
    }
  }
  // FP32 rows stay in add order; once an id differs from its row the rerank
  // needs id -> row, which must then be a function: duplicate ids are refused
  // while rows are kept (checked before anything changes).
  const bool map_rows = h->keep_vectors && (h->custom_ids || explicit_rows);
  std::vector<int32_t> next_map;
  if (map_rows) {
    next_map = h->row_of_host;
    if (!h->custom_ids) {  // every earlier row was added under its own number
      next_map.resize((size_t)h->ntotal);
      for (int64_t r = 0; r < h->ntotal; r++) next_map[(size_t)r] = (int32_t)r;
    }
    for (int64_t i = 0; i < n; i++) {
      const int64_t id = ids ? ids[i] : h->ntotal + i;
      if ((int64_t)next_map.size() <= id) next_map.resize((size_t)id + 1, -1);
      if (next_map[(size_t)id] >= 0)
        throw Error(mvg::kInvalid,
                    "duplicate IVF-PQ id " + std::to_string(id) +
                        ": ids must be unique while FP32 rows are kept for the rerank");
      next_map[(size_t)id] = (int32_t)(h->ntotal + i);
    }
  }
  const int d = h->d, M = h->M;
  const float *dx = upload(h->tx, x, (size_t)n * d, h->s);
  h->tlab.reserve(sizeof(int32_t) * n);
  h->tcodes.reserve((size_t)n * M);
  int32_t *dl = h->tlab.as<int32_t>();
  mvg::flat_argmin(h->coarse.as<float>(), h->nlist, d, h->metric, false, dx, n, dl, nullptr, h->s);
  mvg::pq_assign(dx, n, d, M, h->pq.as<float>(), h->coarse.as<float>(), dl,
                 h->tcodes.as<uint8_t>(), nullptr, h->s);
  std::vector<int32_t> lab(n);
  std::vector<uint8_t> codes((size_t)n * M);
  download(lab.data(), dl, (size_t)n, h->s);
  download(codes.data(), h->tcodes.ptr, (size_t)n * M, h->s);
  if (h->keep_vectors) {
    const size_t keep = sizeof(float) * (size_t)h->ntotal * d;
    const size_t need = keep + sizeof(float) * (size_t)n * d;
    h->vectors.grow_keep(std::max(need, h->vectors.bytes + h->vectors.bytes / 2), keep, h->s);
    MVG_HIP(hipMemcpyAsync((char *)h->vectors.ptr + keep, dx, need - keep,
                           hipMemcpyDeviceToDevice, h->s));
  }
  MVG_HIP(hipStreamSynchronize(h->s));
  for (int64_t i = 0; i < n; i++) {  // :199-211, append order
    const int l = lab[i];
    if (l < 0 || l >= h->nlist) continue;
    h->lids[l].push_back((int32_t)(ids ? ids[i] : h->ntotal + i));
    h->lcodes[l].insert(h->lcodes[l].end(), codes.begin() + i * M, codes.begin() + (i + 1) * M);
  }
  if (map_rows) {
    h->row_of_host.swap(next_map);
    h->custom_ids = true;
    h->row_of_dirty = true;
  }
  h->ntotal += n;
  h->dirty = true;
  h->vq_dirty = true;
  h->generation++;
  API_END
}

int mvg_ivfpq_set_nprobe(mvg_ivfpq *h, int nprobe) {
  API_BEGIN
  require(h && nprobe >= 1, "nprobe must be at least 1");
  h->nprobe = nprobe;
  API_END
}

int mvg_ivfpq_search_device(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                            const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                            const uint8_t *exclude, int64_t exclude_len, float *distances,
                            int64_t *labels, void *stream) {
  return mvg_ivfpq_search_device_k(h, q, nq, nullptr, k, nprobe_per_query, max_nprobe, k_factor,
                                   exclude, exclude_len, distances, labels, stream);
}

int mvg_ivfpq_search_device_k(mvg_ivfpq *h, const float *q, int64_t nq, const int32_t *k_per_query,
                              int max_k, const int32_t *nprobe_per_query, int max_nprobe,
                              int k_factor, const uint8_t *exclude, int64_t exclude_len,
                              float *distances, int64_t *labels, void *stream) {
  API_BEGIN
  require(h && nq >= 0 && max_k >= 1 && k_factor >= 1, "invalid arguments");
  require(h->trained && h->ntotal > 0, "search_device needs a trained, non-empty index");
  hipStream_t s = stream ? (hipStream_t)stream : h->s;
  std::lock_guard<std::mutex> lk(h->search_mu);
  mvg::ivf_prepare(h);
  mvg::ivf_search_with(h, q, nq, max_k, nprobe_per_query, max_nprobe, k_factor, exclude, exclude_len,
                       distances, labels, h->work, s, k_per_query);
  API_END
}

}  // extern "C"

// ---- allocation-free search ---------------------------------------------------

namespace {

// Workspace of mvg_ivfpq_search_device_noalloc: one fixed array per GPU,
// held until the process ends, shared by every call on that GPU.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<char *> g_scratch_ptr;  // device address of `scratch`, per device

// The current device's address of `scratch`, looked up once per device with
// that device current (the host symbol is only a stand-in, not an address).
char *scratch_base() {
  int dev = 0;
  MVG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[(size_t)dev]) {
    void *p = nullptr;
    MVG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(mvg::kDevice, "scratch array has no device address");
    g_scratch_ptr[(size_t)dev] = static_cast<char *>(p);
  }
  return g_scratch_ptr[(size_t)dev];
}

}  // namespace

extern "C" int mvg_ivfpq_search_device_noalloc(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                                               const int32_t *nprobe_per_query, int max_nprobe,
                                               int k_factor, const uint8_t *exclude,
                                               int64_t exclude_len, float *distances,
                                               int64_t *labels, void *stream) {
  API_BEGIN
  require(h && nq >= 0 && k >= 1 && k_factor >= 1, "invalid arguments");
  require(h->trained && h->ntotal > 0, "search_device needs a trained, non-empty index");
  hipStream_t s = stream ? (hipStream_t)stream : h->s;
  std::lock_guard<std::mutex> lk(h->search_mu);
  // Preparing the index (CSR upload, mixed-form tables) allocates: it must
  // have happened already, through any other search of the current state.
  const bool mixed = h->precision != 0;
  if (h->dirty || h->row_of_dirty || (mixed && h->tables_dirty) ||
      (mixed && h->metric == MVG_METRIC_L2 && h->pre_dirty) ||
      (mixed && h->metric == MVG_METRIC_L2 && h->table_bits != 32 && !h->pre8_ready) ||
      (mixed && h->metric == MVG_METRIC_L2 && h->table_bits == 16 && !h->tile_ready) ||
      (mixed && h->prefilter > 0 && h->keep_vectors && h->vq_dirty))
    throw Error(mvg::kState,
                "the index is not prepared: run one search (mvg_ivfpq_search or "
                "mvg_ivfpq_search_device) after the last change before the allocation-free one");
  // The first pieces every search takes (probe lists, candidate keys) checked
  // before anything is enqueued; the rest is checked piece by piece.
  const int maxp = std::min(nprobe_per_query ? max_nprobe : h->nprobe, h->nlist);
  const size_t first = sizeof(int32_t) * (size_t)nq * (size_t)std::max(maxp, 1) +
                       (k_factor > 1 ? sizeof(uint64_t) * (size_t)nq * k * k_factor : 0);
  if (first > kScratchBytes)
    throw Error(mvg::kInvalid, "the request needs more than the " + std::to_string(kScratchBytes >> 20) +
                                   " MiB scratch of the allocation-free entry point");
  if (nq == 0) return MVG_OK;
  mvg::Arena arena;
  arena.base = scratch_base();
  arena.size = kScratchBytes;
  mvg::IvfWork work;
  work.bind(&arena);
  mvg::ivf_search_with(h, q, nq, k, nprobe_per_query, max_nprobe, k_factor, exclude, exclude_len,
                       distances, labels, work, s);
  API_END
}

void mvg::ivf_prepare(mvg_ivfpq *h) {
  ivf_sync_lists(h);
  if (h->row_of_dirty) {
    upload(h->row_of, h->row_of_host.data(), h->row_of_host.size(), h->s);
    MVG_HIP(hipStreamSynchronize(h->s));
    h->row_of_dirty = false;
  }
  if (h->precision != 0 && h->tables_dirty) {
    if (h->metric == MVG_METRIC_L2) {
      h->ptab.reserve(sizeof(float) * (size_t)h->nlist * h->M * mvg::kKsub);
      mvg::precompute_ivf_tables(h->coarse.as<float>(), h->pq.as<float>(), h->d, h->M, h->nlist,
                                 h->ptab.as<float>(), h->s);
    }
    h->coarse_bf16.reserve((size_t)h->nlist * mvg::gemm_row_bytes(mvg::kGemmBf16L2, h->d));
    h->coarse_norm.reserve(sizeof(float) * h->nlist);
    mvg::to_bf16(h->coarse.as<float>(), h->nlist, h->d, h->coarse_bf16.as<uint16_t>(),
                 h->coarse_norm.as<float>(), h->s);
    MVG_HIP(hipStreamSynchronize(h->s));
    h->tables_dirty = false;
    h->pre_dirty = true;
    h->tile_ready = false;
  }
  if (h->precision != 0 && h->prefilter > 0 && h->keep_vectors && h->vq_dirty && h->ntotal > 0) {
    // SQ8 codes of the stored rows (index_flat_sq8.mojo:317-367 range and scale)
    const int64_t n = h->ntotal;
    const int d = h->d, ld = mvg::sq8_row_stride(d);
    float mm[2];
    mvg::sq8_minmax(h->vectors.as<float>(), n * d, mm, h->vq_tmp, h->s);
    const bool sym = h->metric != MVG_METRIC_L2;
    if (!sym) {
      h->vq_scale = mm[1] == mm[0] ? 1.0f : (mm[1] - mm[0]) / 255.0f;
    } else {
      const float mx = std::max(std::fabs(mm[0]), std::fabs(mm[1]));
      h->vq_scale = mx > 0.0f ? mx / 127.0f : 1.0f;
    }
    h->vq_min = sym ? 0.0f : mm[0];
    h->vq_codes.reserve((size_t)n * ld + 16);
    h->vq_norms.reserve(sizeof(uint32_t) * (size_t)n);
    mvg::sq8_encode(h->vectors.as<float>(), n, d, sym ? 1 : 0, h->vq_min, 1.0f / h->vq_scale,
                    h->vq_codes.as<uint8_t>(), h->vq_norms.as<uint32_t>(), nullptr, h->s);
    MVG_HIP(hipStreamSynchronize(h->s));
    h->vq_dirty = false;
  }
  if (h->precision != 0 && h->metric == MVG_METRIC_L2 && h->pre_dirty) {
    // one term per CSR entry (the lists' own total, whatever ntotal says)
    h->entry_pre.reserve(sizeof(float) * (size_t)std::max<int64_t>(h->ncsr, 1));
    mvg::precompute_entry_terms(h->list_off.as<int64_t>(), h->nlist, h->codes.as<uint8_t>(), h->M,
                                h->ptab.as<float>(), h->ncsr, h->entry_pre.as<float>(), h->s);
    MVG_HIP(hipStreamSynchronize(h->s));
    h->pre_dirty = false;
    h->pre8_ready = false;
  }
  // 8-bit terms: table bits 8, and 16 (its batches below the tile scan's
  // minimum take the 8-bit deep-pool scan)
  if (h->precision != 0 && h->metric == MVG_METRIC_L2 && h->table_bits != 32 && !h->pre8_ready) {
    h->entry_pre8.reserve(sizeof(float) * (size_t)std::max<int64_t>(h->ncsr, 1));
    h->table_mu.reserve(sizeof(float) * (size_t)h->d);
    mvg::precompute_centered_terms(h->coarse.as<float>(), h->nlist, h->d, h->codes.as<uint8_t>(), h->M,
                                   h->pq.as<float>(), h->entry_pre.as<float>(), h->ncsr,
                                   h->table_mu.as<float>(), h->entry_pre8.as<float>(), h->s);
    MVG_HIP(hipStreamSynchronize(h->s));
    h->pre8_ready = true;
  }
  if (h->precision != 0 && h->metric == MVG_METRIC_L2 && h->table_bits == 16 && !h->tile_ready) {
    // FP16 codebooks and the list order of the MFMA tile scan (PQ with
    // 8-dimensional subspaces and M in {8, 12, 16}; other shapes keep the
    // FP32 tables)
    const int dsub = h->d / h->M;
    h->cb16.release();
    if (dsub == 8 && (h->M == 8 || h->M == 12 || h->M == 16)) {
      std::vector<float> pqh((size_t)h->d * mvg::kKsub);
      download(pqh.data(), h->pq.ptr, pqh.size(), h->s);
      MVG_HIP(hipStreamSynchronize(h->s));
      float mx = 0.0f;
      for (float v : pqh) mx = std::max(mx, std::fabs(v));
      h->cb16.reserve(sizeof(uint16_t) * pqh.size());
      h->cb16_inv = mvg::tile_codebook(h->pq.as<float>(), h->M, dsub, mx, h->cb16.as<uint16_t>(), h->s);
      // Super-clusters of the coarse centroids (Lloyd from evenly spaced
      // centroids); lists are ranked by (super-cluster, list id), and queries
      // sorted by the rank of their nearest list form tiles whose probe sets
      // overlap.
      const int nl = h->nlist, d = h->d;
      const int nsup = std::max(1, std::min(nl / 16, 4096));
      h->tile_cent.reserve(sizeof(float) * (size_t)nsup * d);
      for (int c = 0; c < nsup; c++)
        MVG_HIP(hipMemcpyAsync(h->tile_cent.as<float>() + (size_t)c * d,
                               h->coarse.as<float>() + (size_t)((int64_t)c * nl / nsup) * d, sizeof(float) * d,
                               hipMemcpyDeviceToDevice, h->s));
      if (nsup > 1)
        mvg::kmeans_fit(h->coarse.as<float>(), nl, 1, d, nsup, 10, 0, h->tile_cent.as<float>(), h->kw, h->s);
      mvg::DevBuf dsup;
      dsup.reserve(sizeof(int32_t) * (size_t)nl);
      mvg::flat_argmin(h->tile_cent.as<float>(), nsup, d, mvg::kMetricL2, true, h->coarse.as<float>(), nl,
                       dsup.as<int32_t>(), nullptr, h->s);
      std::vector<int32_t> sup((size_t)nl), rank((size_t)nl), byrank((size_t)nl);
      download(sup.data(), dsup.ptr, sup.size(), h->s);
      MVG_HIP(hipStreamSynchronize(h->s));
      for (int l = 0; l < nl; l++) byrank[l] = l;
      std::stable_sort(byrank.begin(), byrank.end(), [&](int x, int y) { return sup[x] < sup[y]; });
      for (int r = 0; r < nl; r++) rank[byrank[r]] = r;
      upload(h->list_rank, rank.data(), rank.size(), h->s);
    }
    MVG_HIP(hipStreamSynchronize(h->s));
    h->tile_ready = true;
  }
}

void mvg::ivf_check_unserved(const mvg_ivfpq *h) {
  if (h->servers.load() > 0)
    throw Error(mvg::kState, "the index is held by a server: destroy the server before changing it");
}

void mvg::ivf_search_with(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                          const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                          const uint8_t *exclude, int64_t exclude_len, float *distances,
                          int64_t *labels, mvg::IvfWork &work, hipStream_t s,
                          const int32_t *k_per_query) {
  mvg::IvfSearchArgs a{};
  a.q = q;
  a.nq = nq;
  a.d = h->d;
  a.M = h->M;
  a.nlist = h->nlist;
  a.metric = h->metric;
  a.coarse = h->coarse.as<float>();
  a.pq = h->pq.as<float>();
  a.list_off = h->list_off.as<int64_t>();
  a.codes = h->codes.as<uint8_t>();
  a.ids = h->ids.as<int32_t>();
  a.vectors = h->keep_vectors ? h->vectors.as<float>() : nullptr;
  a.k = k;
  a.k_factor = k_factor;
  a.nprobe = std::min(nprobe_per_query ? max_nprobe : h->nprobe, h->nlist);
  require(a.nprobe >= 1, "max_nprobe must be at least 1");
  a.nprobe_per_query = nprobe_per_query;
  a.excl = exclude_len > 0 ? exclude : nullptr;
  a.excl_len = exclude_len;
  a.D = distances;
  a.I = labels;
  a.precision = h->precision;
  a.split_scan = h->split_scan;
  a.wq_entries = h->wq_entries;
  a.ntotal = h->ntotal;
  a.ncsr = h->ncsr;
  a.deep_cap = h->top_sum.empty() ? 0 : h->top_sum[(size_t)a.nprobe];
  a.deep_min = h->deep_min;
  a.tile_min_batch = h->tile_min_batch;
  a.deep_pool = h->deep_pool == 0 ? 1 : 0;  // ABI pool 0: keys in HBM; 1: running set in LDS
  a.kq = k_per_query;
  if (h->custom_ids && a.vectors) {
    if (h->row_of_dirty) throw Error(mvg::kState, "explicit-id row map is not prepared");
    a.rows.row_of = h->row_of.as<int32_t>();
    a.rows.nids = (int64_t)h->row_of_host.size();
  }
  if (h->precision != 0) {
    if (h->tables_dirty) throw Error(mvg::kState, "mixed-precision tables are not prepared");
    if (h->metric == MVG_METRIC_L2) {
      a.ptab = h->ptab.as<float>();
      if (h->pre_dirty) throw Error(mvg::kState, "mixed-precision entry terms are not prepared");
      a.pre = h->entry_pre.as<float>();
    }
    a.coarse_bf16 = h->coarse_bf16.as<uint16_t>();
    a.coarse_norm = h->coarse_norm.as<float>();
    if (h->prefilter > 0 && a.vectors && h->ntotal > 0) {
      if (h->vq_dirty) throw Error(mvg::kState, "pre-rerank row codes are not prepared");
      a.vq_codes = h->vq_codes.as<uint8_t>();
      a.vq_norms = h->vq_norms.as<uint32_t>();
      a.vq_min = h->vq_min;
      a.vq_scale = h->vq_scale;
      a.prefilter = h->prefilter;
    }
    a.table_bits = h->table_bits;
    if (h->metric == MVG_METRIC_L2 && h->table_bits == 16) {
      if (!h->tile_ready) throw Error(mvg::kState, "tile-scan codebooks are not prepared");
      if (h->cb16.ptr) {
        a.cb16 = h->cb16.as<uint16_t>();
        a.cb16_inv = h->cb16_inv;
        a.list_rank = h->list_rank.as<int32_t>();
      }
    }
    if (h->metric == MVG_METRIC_L2 && h->table_bits != 32) {
      if (!h->pre8_ready) throw Error(mvg::kState, "8-bit table terms are not prepared");
      a.pre8 = h->entry_pre8.as<float>();
      a.mu = h->table_mu.as<float>();
    }
  }
  mvg::ivfpq_search(a, work, s);
}

extern "C" {

int mvg_ivfpq_search(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                     const int32_t *nprobe_per_query, int k_factor, const uint8_t *exclude,
                     int64_t exclude_len, float *distances, int64_t *labels) {
  API_BEGIN
  require(h && nq >= 0 && k >= 1 && k_factor >= 1, "invalid arguments");
  if (nq == 0) return MVG_OK;
  if (!h->trained || h->ntotal == 0) {  // index_ivf_pq.mojo:247-251
    for (int64_t i = 0; i < nq * k; i++) {
      distances[i] = mvg::kIvfPadDist;
      labels[i] = -1;
    }
    return MVG_OK;
  }
  std::lock_guard<std::mutex> lk(h->search_mu);
  const float *dq = upload(h->q, q, (size_t)nq * h->d, h->s);
  const int32_t *dnp = nullptr;
  int maxp = h->nprobe;
  if (nprobe_per_query) {
    std::vector<int32_t> npq(nprobe_per_query, nprobe_per_query + nq);
    maxp = 1;
    for (auto &v : npq) {
      require(v >= 1, "nprobe must be at least 1");
      v = std::min(v, h->nlist);
      maxp = std::max(maxp, (int)v);
    }
    dnp = upload(h->npq, npq.data(), (size_t)nq, h->s);
  }
  const uint8_t *dx = exclude && exclude_len > 0 ? upload(h->excl, exclude, (size_t)exclude_len, h->s) : nullptr;
  h->D.reserve(sizeof(float) * nq * k);
  h->I.reserve(sizeof(int64_t) * nq * k);
  mvg::ivf_prepare(h);
  mvg::ivf_search_with(h, dq, nq, k, dnp, maxp, k_factor, dx, dx ? exclude_len : 0, h->D.as<float>(),
                       h->I.as<int64_t>(), h->work, h->s);
  download(distances, h->D.ptr, (size_t)nq * k, h->s);
  download(labels, h->I.ptr, (size_t)nq * k, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_ivfpq_set_search_mode(mvg_ivfpq *h, int precision, int split_scan) {
  API_BEGIN
  require(h && (precision == 0 || precision == 1), "precision must be 0 (plain) or 1 (mixed)");
  mvg::ivf_check_unserved(h);
  h->precision = precision;
  h->split_scan = split_scan != 0;
  API_END
}

int mvg_ivfpq_set_prefilter(mvg_ivfpq *h, int per_k) {
  API_BEGIN
  require(h && per_k >= 0, "per_k must be >= 0");
  mvg::ivf_check_unserved(h);
  h->prefilter = per_k;
  API_END
}

int mvg_ivfpq_set_table_bits(mvg_ivfpq *h, int bits) {
  API_BEGIN
  require(h && (bits == 8 || bits == 16 || bits == 32), "bits must be 8, 16 or 32");
  mvg::ivf_check_unserved(h);
  h->table_bits = bits;
  API_END
}

int mvg_ivfpq_set_work_queue(mvg_ivfpq *h, int entries_per_item) {
  API_BEGIN
  require(h && entries_per_item >= 0, "entries_per_item must be >= 0");
  mvg::ivf_check_unserved(h);
  h->wq_entries = entries_per_item;
  API_END
}

int mvg_ivfpq_set_deep_select(mvg_ivfpq *h, int min_candidates, int pool) {
  API_BEGIN
  require(h && min_candidates >= 0 && (pool == 0 || pool == 1), "invalid arguments");
  mvg::ivf_check_unserved(h);
  h->deep_min = min_candidates;
  h->deep_pool = pool;
  API_END
}

int mvg_ivfpq_set_tile_min_batch(mvg_ivfpq *h, int64_t min_queries) {
  API_BEGIN
  require(h && min_queries >= 0, "min_queries must be >= 0");
  mvg::ivf_check_unserved(h);
  h->tile_min_batch = min_queries;
  API_END
}

int mvg_ivfpq_set_profiling(mvg_ivfpq *h, int on) {
  API_BEGIN
  require(h, "invalid handle");
  h->work.prof.on = on != 0;
  h->work.prof.recorded = false;
  API_END
}

int mvg_ivfpq_read_profile(mvg_ivfpq *h, float *stage_ms, int64_t *scanned_entries,
                           int64_t *pool_writes) {
  API_BEGIN
  require(h, "invalid handle");
  auto &pf = h->work.prof;
  if (!pf.recorded) throw Error(mvg::kState, "no profiled search has run");
  MVG_HIP(hipEventSynchronize(pf.ev[3]));
  for (int i = 0; i < 3; i++)
    if (stage_ms) MVG_HIP(hipEventElapsedTime(&stage_ms[i], pf.ev[i], pf.ev[i + 1]));
  if (scanned_entries || pool_writes) {  // written on the search's stream, up to after ev[3]
    unsigned long long v[2] = {0, 0};
    MVG_HIP(hipMemcpyAsync(v, pf.counter.ptr, sizeof(v), hipMemcpyDeviceToHost, pf.stream));
    MVG_HIP(hipStreamSynchronize(pf.stream));
    if (scanned_entries) *scanned_entries = (int64_t)v[0];
    if (pool_writes) *pool_writes = (int64_t)v[1];
  }
  API_END
}

int mvg_ivfpq_info(mvg_ivfpq *h, int64_t *ntotal, int *trained, int *nprobe) {
  API_BEGIN
  require(h, "invalid handle");
  if (ntotal) *ntotal = h->ntotal;
  if (trained) *trained = h->trained ? 1 : 0;
  if (nprobe) *nprobe = h->nprobe;
  API_END
}

int mvg_ivfpq_shape(mvg_ivfpq *h, int *d, int *nlist, int *M, int *metric) {
  API_BEGIN
  require(h, "invalid handle");
  if (d) *d = h->d;
  if (nlist) *nlist = h->nlist;
  if (M) *M = h->M;
  if (metric) *metric = h->metric;
  API_END
}

int mvg_ivfpq_list_sizes(mvg_ivfpq *h, int64_t *sizes) {
  API_BEGIN
  require(h, "invalid handle");
  for (int l = 0; l < h->nlist; l++) sizes[l] = (int64_t)h->lids[l].size();
  API_END
}

int mvg_ivfpq_get_list(mvg_ivfpq *h, int list, int64_t *ids, uint8_t *codes) {
  API_BEGIN
  require(h && list >= 0 && list < h->nlist, "list index out of range");
  const auto &li = h->lids[list];
  for (size_t i = 0; i < li.size(); i++) ids[i] = li[i];
  std::copy(h->lcodes[list].begin(), h->lcodes[list].end(), codes);
  API_END
}

void mvg_ivfpq_destroy(mvg_ivfpq *h) {
  if (!h) return;
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

}  // extern "C"

// ---- captured batch ---------------------------------------------------------

struct mvg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  mvg::IvfWork work;  // the workspace the graph's kernels point into
  mvg_ivfpq *index = nullptr;
  uint64_t generation = 0;  // the index's mutation generation at capture
};

extern "C" {

int mvg_ivfpq_capture(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                      const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                      float *distances, int64_t *labels, mvg_graph **out) {
  return mvg_ivfpq_capture_k(h, q, nq, nullptr, k, nprobe_per_query, max_nprobe, k_factor, distances,
                             labels, out);
}

int mvg_ivfpq_capture_k(mvg_ivfpq *h, const float *q, int64_t nq, const int32_t *k_per_query,
                        int max_k, const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                        float *distances, int64_t *labels, mvg_graph **out) {
  API_BEGIN
  const int k = max_k;
  require(h && nq >= 1 && k >= 1 && k_factor >= 1 && out, "invalid arguments");
  require(h->trained && h->ntotal > 0, "capture needs a trained, non-empty index");
  std::lock_guard<std::mutex> lk(h->search_mu);
  mvg::ivf_prepare(h);
  auto g = std::make_unique<mvg_graph>();
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } guard{s};
  const bool prof = h->work.prof.on;
  g->work.prof.on = false;
  // One eager run sizes every buffer, so the capture allocates nothing.
  mvg::ivf_search_with(h, q, nq, k, nprobe_per_query, max_nprobe, k_factor, nullptr, 0, distances,
                       labels, g->work, s, k_per_query);
  MVG_HIP(hipStreamSynchronize(s));
  MVG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  try {
    mvg::ivf_search_with(h, q, nq, k, nprobe_per_query, max_nprobe, k_factor, nullptr, 0, distances,
                         labels, g->work, s, k_per_query);
  } catch (...) {
    hipGraph_t dummy = nullptr;
    (void)hipStreamEndCapture(s, &dummy);
    if (dummy) (void)hipGraphDestroy(dummy);
    throw;
  }
  MVG_HIP(hipStreamEndCapture(s, &g->graph));
  MVG_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  (void)prof;
  g->index = h;
  g->generation = h->generation.load();
  *out = g.release();
  API_END
}

int mvg_ivfpq_autotune(mvg_ivfpq *h, const float *q, int64_t nq, int k,
                       const int32_t *nprobe_per_query, int max_nprobe, int k_factor,
                       float *distances, int64_t *labels, int *choice, float *best_ms) {
  API_BEGIN
  require(h && nq >= 1 && k >= 1 && k_factor >= 1, "invalid arguments");
  require(h->trained && h->ntotal > 0, "autotune needs a trained, non-empty index");
  std::lock_guard<std::mutex> lk(h->search_mu);
  mvg::ivf_prepare(h);
  // Candidate scan schedules: {split_scan, entries per work-queue item}.
  const int cands[][2] = {{0, 0}, {0, 1024}, {0, 4096}, {0, 16384}, {1, 0}};
  const int ncand = sizeof(cands) / sizeof(cands[0]);
  hipStream_t s = new_stream();
  struct Guard { hipStream_t s; ~Guard() { (void)hipStreamDestroy(s); } } guard{s};
  hipEvent_t e0, e1;
  MVG_HIP(hipEventCreate(&e0));
  MVG_HIP(hipEventCreate(&e1));
  struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
  const int keep_split = h->split_scan, keep_wq = h->wq_entries;
  int best = -1;
  float best_t = 0.0f;
  for (int c = 0; c < ncand; c++) {
    h->split_scan = cands[c][0];
    h->wq_entries = cands[c][1];
    try {
      mvg::ivf_search_with(h, q, nq, k, nprobe_per_query, max_nprobe, k_factor, nullptr, 0,
                           distances, labels, h->work, s);  // warm-up sizes the workspace
      MVG_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < 3; r++)
        mvg::ivf_search_with(h, q, nq, k, nprobe_per_query, max_nprobe, k_factor, nullptr, 0,
                             distances, labels, h->work, s);
      MVG_HIP(hipEventRecord(e1, s));
      MVG_HIP(hipEventSynchronize(e1));
    } catch (const Error &) {
      continue;  // a schedule this shape cannot run (workspace limits)
    }
    float ms = 0.0f;
    MVG_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms /= 3.0f;
    if (best < 0 || ms < best_t) {
      best = c;
      best_t = ms;
    }
  }
  if (best < 0) {
    h->split_scan = keep_split;
    h->wq_entries = keep_wq;
    throw Error(mvg::kState, "no scan schedule ran for this batch shape");
  }
  h->split_scan = cands[best][0];
  h->wq_entries = cands[best][1];
  if (choice) *choice = best;
  if (best_ms) *best_ms = best_t;
  API_END
}

int mvg_graph_launch(mvg_graph *g, void *stream) {
  API_BEGIN
  require(g && g->exec, "invalid graph");
  // the graph's kernels hold the index buffers of capture time
  if (g->index && g->index->generation.load() != g->generation)
    throw Error(mvg::kState, "the index changed after this graph was captured: capture it again");
  MVG_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  API_END
}

void mvg_graph_destroy(mvg_graph *g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

}  // extern "C"

// ---- IVF-Flat (mojovec/index/index_ivf_flat.mojo) ------------------------------

namespace {

// Device CSR of the host lists; each list is written in id order (stable).
void ivfflat_sync(mvg_ivfflat *h) {
  if (!h->dirty) return;
  const int d = h->d;
  std::vector<int64_t> off(h->nlist + 1, 0);
  for (int l = 0; l < h->nlist; l++) off[l + 1] = off[l] + (int64_t)h->lids[l].size();
  const int64_t total = off[h->nlist];
  std::vector<int32_t> ids((size_t)std::max<int64_t>(total, 1));
  std::vector<float> rows((size_t)std::max<int64_t>(total * d, 1));
  std::vector<int32_t> order;
  h->max_list = 0;
  for (int l = 0; l < h->nlist; l++) {
    const auto &li = h->lids[l];
    const int64_t L = (int64_t)li.size();
    h->max_list = std::max(h->max_list, L);
    order.resize(L);
    for (int64_t j = 0; j < L; j++) order[j] = (int32_t)j;
    if (!std::is_sorted(li.begin(), li.end()))
      std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return li[a] < li[b]; });
    for (int64_t j = 0; j < L; j++) {
      ids[off[l] + j] = li[order[j]];
      std::copy(h->lrows[l].begin() + (size_t)order[j] * d, h->lrows[l].begin() + (size_t)(order[j] + 1) * d,
                rows.begin() + (size_t)(off[l] + j) * d);
    }
  }
  upload(h->list_off, off.data(), off.size(), h->s);
  upload(h->ids, ids.data(), ids.size(), h->s);
  upload(h->rows, rows.data(), rows.size(), h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  h->dirty = false;
}

void ivfflat_search_on(mvg_ivfflat *h, const float *q, int64_t nq, int k, const uint8_t *exclude,
                       int64_t exclude_len, float *distances, int64_t *labels, hipStream_t s) {
  mvg::IvfFlatArgs a{};
  a.q = q;
  a.nq = nq;
  a.d = h->d;
  a.nlist = h->nlist;
  a.metric = h->metric;
  a.coarse_metric = h->quantizer_metric;
  a.coarse = h->coarse.as<float>();
  a.list_off = h->list_off.as<int64_t>();
  a.rows = h->rows.as<float>();
  a.ids = h->ids.as<int32_t>();
  a.ntotal = h->ntotal;
  a.max_list = h->max_list;
  a.k = k;
  a.nprobe = h->nprobe;
  a.excl = exclude_len > 0 ? exclude : nullptr;
  a.excl_len = exclude_len;
  a.D = distances;
  a.I = labels;
  mvg::ivfflat_search(a, h->work, s);
}

}  // namespace

extern "C" {

int mvg_ivfflat_create(int d, int nlist, int metric, mvg_ivfflat **out) {
  return mvg_ivfflat_create_with_quantizer(d, nlist, metric, metric, out);
}

int mvg_ivfflat_create_with_quantizer(int d, int nlist, int metric, int quantizer_metric,
                                      mvg_ivfflat **out) {
  API_BEGIN
  require(out != nullptr, "invalid arguments");
  require(d > 0, "IVF-Flat dimension must be positive.");
  require(nlist > 0 && nlist <= 1000000, "IVF-Flat nlist must be between 1 and 1000000.");
  require((metric == MVG_METRIC_L2 || metric == MVG_METRIC_IP) &&
              (quantizer_metric == MVG_METRIC_L2 || quantizer_metric == MVG_METRIC_IP),
          "metric must be 0 (L2) or 1 (IP).");
  auto h = std::make_unique<mvg_ivfflat>();
  h->d = d;
  h->nlist = nlist;
  h->metric = metric;
  h->quantizer_metric = quantizer_metric;
  h->list_capacity.assign((size_t)nlist, 0);
  h->lids.resize(nlist);
  h->lrows.resize(nlist);
  h->s = new_stream();
  h->coarse.reserve(sizeof(float) * (size_t)nlist * d);
  *out = h.release();
  API_END
}

int mvg_ivfflat_train(mvg_ivfflat *h, const float *x, int64_t n, int niter, uint64_t seed) {
  API_BEGIN
  require(h && n >= 0 && niter >= 0, "invalid arguments");
  if (h->trained || n == 0) return MVG_OK;  // index_ivf_flat.mojo:71
  const int d = h->d;
  float *dx = upload(h->tx, x, (size_t)n * d, h->s);
  std::vector<int64_t> rows;  // KMeans(d, nlist, 15).train(x) (:73-76)
  const int64_t off = kmeans_pick(n, h->nlist, seed, rows);
  std::vector<float> c0((size_t)h->nlist * d);
  for (int c = 0; c < h->nlist; c++) std::copy(x + rows[c] * d, x + rows[c] * d + d, c0.begin() + (size_t)c * d);
  float *dc = h->coarse.as<float>();
  MVG_HIP(hipMemcpyAsync(dc, c0.data(), sizeof(float) * c0.size(), hipMemcpyHostToDevice, h->s));
  mvg::kmeans_fit(dx, n, 1, d, h->nlist, niter, off, dc, h->kw, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  h->trained = true;
  API_END
}

int mvg_ivfflat_set_centroids(mvg_ivfflat *h, const float *coarse) {
  API_BEGIN
  require(h && coarse, "invalid arguments");
  MVG_HIP(hipMemcpyAsync(h->coarse.ptr, coarse, sizeof(float) * (size_t)h->nlist * h->d,
                         hipMemcpyHostToDevice, h->s));
  MVG_HIP(hipStreamSynchronize(h->s));
  h->trained = true;
  API_END
}

int mvg_ivfflat_get_centroids(mvg_ivfflat *h, float *coarse) {
  API_BEGIN
  require(h && coarse, "invalid arguments");
  download(coarse, h->coarse.ptr, (size_t)h->nlist * h->d, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_ivfflat_add(mvg_ivfflat *h, const float *x, int64_t n, const int64_t *ids) {
  API_BEGIN
  require(h && n >= 0, "invalid arguments");
  if (!h->trained || n == 0) return MVG_OK;  // index_ivf_flat.mojo:102-104
  if (ids) {
    for (int64_t i = 0; i < n; i++)
      require(ids[i] >= 0 && ids[i] <= INT32_MAX, "IVF-Flat ids must be in [0, 2^31 - 1]");
  } else if (h->ntotal + n - 1 > INT32_MAX) {
    throw Error(mvg::kInvalid, "IVF-Flat sequential ids must stay below 2^31");
  }
  const int d = h->d;
  const float *dx = upload(h->tx, x, (size_t)n * d, h->s);
  h->tlab.reserve(sizeof(int32_t) * n);
  int32_t *dl = h->tlab.as<int32_t>();
  // quantizer.search(x, 1) (:110-116)
  mvg::flat_argmin(h->coarse.as<float>(), h->nlist, d, h->quantizer_metric, false, dx, n, dl, nullptr,
                   h->s);
  std::vector<int32_t> lab(n);
  download(lab.data(), dl, (size_t)n, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  for (int64_t i = 0; i < n; i++) {  // :123-135, append order
    const int l = lab[i];
    if (l < 0 || l >= h->nlist) continue;
    h->lids[l].push_back((int32_t)(ids ? ids[i] : h->ntotal + i));
    h->lrows[l].insert(h->lrows[l].end(), x + i * d, x + (i + 1) * d);
  }
  h->ntotal += n;  // :137
  h->dirty = true;
  API_END
}

int mvg_ivfflat_set_nprobe(mvg_ivfflat *h, int nprobe) {
  API_BEGIN
  require(h && nprobe >= 1, "nprobe must be at least 1");
  h->nprobe = nprobe;
  API_END
}

int mvg_ivfflat_search(mvg_ivfflat *h, const float *q, int64_t nq, int k, const uint8_t *exclude,
                       int64_t exclude_len, float *distances, int64_t *labels) {
  API_BEGIN
  require(h && nq >= 0 && k >= 1, "invalid arguments");
  if (nq == 0) return MVG_OK;
  if (!h->trained || h->ntotal == 0) {  // index_ivf_flat.mojo:171-175
    for (int64_t i = 0; i < nq * k; i++) {
      distances[i] = mvg::kIvfPadDist;
      labels[i] = -1;
    }
    return MVG_OK;
  }
  std::lock_guard<std::mutex> lk(h->search_mu);
  ivfflat_sync(h);
  const float *dq = upload(h->q, q, (size_t)nq * h->d, h->s);
  const uint8_t *dx = exclude && exclude_len > 0 ? upload(h->excl, exclude, (size_t)exclude_len, h->s) : nullptr;
  h->D.reserve(sizeof(float) * nq * k);
  h->I.reserve(sizeof(int64_t) * nq * k);
  ivfflat_search_on(h, dq, nq, k, dx, dx ? exclude_len : 0, h->D.as<float>(), h->I.as<int64_t>(), h->s);
  download(distances, h->D.ptr, (size_t)nq * k, h->s);
  download(labels, h->I.ptr, (size_t)nq * k, h->s);
  MVG_HIP(hipStreamSynchronize(h->s));
  API_END
}

int mvg_ivfflat_search_device(mvg_ivfflat *h, const float *q, int64_t nq, int k,
                              const uint8_t *exclude, int64_t exclude_len, float *distances,
                              int64_t *labels, void *stream) {
  API_BEGIN
  require(h && nq >= 0 && k >= 1, "invalid arguments");
  require(h->trained && h->ntotal > 0, "search_device needs a trained, non-empty index");
  hipStream_t s = stream ? (hipStream_t)stream : h->s;
  std::lock_guard<std::mutex> lk(h->search_mu);
  ivfflat_sync(h);
  ivfflat_search_on(h, q, nq, k, exclude, exclude_len, distances, labels, s);
  API_END
}

int mvg_ivfflat_info(mvg_ivfflat *h, int64_t *ntotal, int *trained, int *nprobe) {
  API_BEGIN
  require(h, "invalid handle");
  if (ntotal) *ntotal = h->ntotal;
  if (trained) *trained = h->trained ? 1 : 0;
  if (nprobe) *nprobe = h->nprobe;
  API_END
}

int mvg_ivfflat_shape(mvg_ivfflat *h, int *d, int *nlist, int *metric, int *quantizer_metric) {
  API_BEGIN
  require(h, "invalid handle");
  if (d) *d = h->d;
  if (nlist) *nlist = h->nlist;
  if (metric) *metric = h->metric;
  if (quantizer_metric) *quantizer_metric = h->quantizer_metric;
  API_END
}

int mvg_ivfflat_list_sizes(mvg_ivfflat *h, int64_t *sizes) {
  API_BEGIN
  require(h && sizes, "invalid arguments");
  for (int l = 0; l < h->nlist; l++) sizes[l] = (int64_t)h->lids[l].size();
  API_END
}

int mvg_ivfflat_get_list(mvg_ivfflat *h, int list, int64_t *ids, float *rows) {
  API_BEGIN
  require(h && list >= 0 && list < h->nlist, "list index out of range");
  const auto &li = h->lids[list];
  if (ids)
    for (size_t i = 0; i < li.size(); i++) ids[i] = li[i];
  if (rows) std::copy(h->lrows[list].begin(), h->lrows[list].end(), rows);
  API_END
}

void mvg_ivfflat_destroy(mvg_ivfflat *h) {
  if (!h) return;
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

}  // extern "C"
