// Continuous-batching request loop over HBM-resident indexes: an IVF-PQ
// index, a Flat SQ8 index, or both served by the same loop (one dispatcher,
// one set of slots, one memory budget; each request names its index and a
// batch holds requests of one index).
//
// Requests (a block of queries with its own k, nprobe and k_factor; nprobe is
// ignored by a Flat SQ8 index) queue on the host. A dispatcher thread packs
// waiting requests that share k_factor into a batch (per-query k and nprobe
// travel with each query; a Flat SQ8 batch keeps k * k_factor on one side of
// 64, where its MFMA form changes), admits the batch when a slot is free and
// the batch's device-memory estimate fits the budget, and enqueues it on the
// slot's stream with the slot's own workspace, so batches overlap on the
// device. A batch is searched for its largest k and each query cut to its
// own k (the merge keeps a query's own k * k_factor candidates, which is the
// candidate set of a search with that k alone). Completed slots are retired as
// soon as their event fires; a request's result depends only on its own
// queries (every search kernel works per query, and a split of the base rows
// only changes which partial lists hold a candidate, not the canonical
// (distance, id) top-k they merge to), so it is the same alone and in any mix.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "handles.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t ticket = 0;
  int target = 0;  // 0: the IVF-PQ index, 1: the Flat SQ8 index
  std::vector<float> q;
  int64_t nq = 0;
  int k = 10, nprobe = 1, k_factor = 1;
  float *D = nullptr;
  int64_t *I = nullptr;
  Clock::time_point t_submit, t_done;
  bool done = false;
  int status = MVG_OK;
  std::string error;
};

template <typename T>
struct PinnedBuf {
  T *ptr = nullptr;
  size_t count = 0;
  ~PinnedBuf() {
    if (ptr) (void)hipHostFree(ptr);
  }
  void reserve(size_t n) {
    if (n <= count) return;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    MVG_HIP(hipHostMalloc((void **)&ptr, sizeof(T) * n, hipHostMallocDefault));
    count = n;
  }
};

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  mvg::IvfWork work;
  mvg::Sq8Work swork;
  mvg::DevBuf q, np, kq, D, I;
  PinnedBuf<float> hq, hD;
  PinnedBuf<int32_t> hnp, hkq;
  PinnedBuf<int64_t> hI;
  std::vector<Request *> reqs;
  int64_t nq = 0, bytes = 0;
  int k = 0;  // the batch's largest k: row stride of hD / hI
  bool busy = false;
};

}  // namespace

struct mvg_server {
  mvg_ivfpq *ix = nullptr;  // at least one of ix, sq is set
  mvg_sq8 *sq = nullptr;
  int device = 0;
  int max_batch = 4096;
  int64_t budget = 0, in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> pending;
  std::unordered_map<int64_t, std::unique_ptr<Request>> all;
  int64_t next_ticket = 1;
  std::vector<std::unique_ptr<Slot>> slots;
  std::thread th;
  bool stop = false;
  int64_t done_queries = 0, batches = 0, max_inflight = 0, inflight = 0;
  std::string fatal;
};

namespace {

// Row width of a request's index.
int target_d(const mvg_server *sv, int target) { return target == 0 ? sv->ix->d : sv->sq->d; }

int64_t batch_bytes(const mvg_server *sv, int target, int64_t nq, int k, int kf, int maxp) {
  const int64_t kk = (int64_t)k * kf;
  const int64_t d = target_d(sv, target);
  if (target == 1) {
    // queries, query codes and norms, outputs, merged candidates, and the
    // per-split partial lists: at most 64 splits per query in the 64-query
    // form up to 64 candidates, at most 2048 workgroups of 16 queries in the
    // 16-query form, and above 64 candidates the MFMA forms' own plan
    const int64_t ld = mvg::sq8_row_stride((int)d);
    int64_t lists = nq * kk * 8 * 64 + (int64_t)2048 * 16 * kk * 8;
    if (sv->sq->use_mfma && sv->sq->ntotal > 0) {
      try {
        const int kind = sv->sq->metric == mvg::kMetricIP ? mvg::kGemmI8IP : mvg::kGemmU8L2;
        lists = std::max<int64_t>(lists, (int64_t)mvg::gemm_plan(kind, nq, sv->sq->ntotal, (int)ld, (int)kk).bytes);
      } catch (const mvg::Error &) {
        // a shape the MFMA forms reject fails in the search itself
      }
    }
    return nq * (d * 4 + ld + 8 + (int64_t)k * 12 + kk * 8) + lists + (1 << 20);
  }
  // queries, probes, per-query candidates (work queue: up to 8 items), outputs
  return nq * (d * 6 + (int64_t)maxp * 12 + kk * 8 * 9 + (int64_t)k * 12) + (1 << 20);
}

void retire(mvg_server *sv, Slot &sl) {
  // results: pinned staging -> each request's buffers
  int64_t off = 0;
  const auto now = Clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  for (Request *r : sl.reqs) {
    for (int64_t i = 0; i < r->nq; i++) {
      std::copy(sl.hD.ptr + (off + i) * sl.k, sl.hD.ptr + (off + i) * sl.k + r->k, r->D + i * r->k);
      std::copy(sl.hI.ptr + (off + i) * sl.k, sl.hI.ptr + (off + i) * sl.k + r->k, r->I + i * r->k);
    }
    off += r->nq;
    r->t_done = now;
    r->done = true;
  }
  sv->done_queries += sl.nq;
  sv->in_use -= sl.bytes;
  sv->inflight--;
  sl.reqs.clear();
  sl.busy = false;
  sv->cv_done.notify_all();
}

void fail_slot(mvg_server *sv, Slot &sl, const std::string &msg) {
  std::lock_guard<std::mutex> lk(sv->mu);
  for (Request *r : sl.reqs) {
    r->status = MVG_ERR_DEVICE;
    r->error = msg;
    r->t_done = Clock::now();
    r->done = true;
  }
  sv->in_use -= sl.bytes;
  sv->inflight--;
  sl.reqs.clear();
  sl.busy = false;
  sv->cv_done.notify_all();
}

// Whether a request may join a batch whose head is `head`: equal k_factor;
// over a Flat SQ8 index also k * k_factor on the head's side of 64 (up to 64
// candidates per query the MFMA form keeps one top list per base split,
// larger pools take strided splits with a per-query check, so mixing them
// would slow the small ones).
bool joins(const mvg_server *sv, const Request *head, const Request *r) {
  (void)sv;
  if (r->target != head->target || r->k_factor != head->k_factor) return false;
  if (head->target == 1) {
    const bool small_head = (int64_t)head->k * head->k_factor <= 64;
    const bool small_r = (int64_t)r->k * r->k_factor <= 64;
    if (small_head != small_r) return false;
  }
  return true;
}

// Takes a batch from the queue front under the lock; empty if none fits.
std::vector<Request *> take_batch(mvg_server *sv, int64_t &bytes) {
  std::vector<Request *> batch;
  if (sv->pending.empty()) return batch;
  Request *head = sv->pending.front();
  const int kf = head->k_factor;
  int64_t nq = 0;
  int maxp = 1, maxk = 1;
  for (auto it = sv->pending.begin(); it != sv->pending.end() && batch.size() < 4096;) {
    Request *r = *it;
    if (!joins(sv, head, r)) {
      ++it;
      continue;
    }
    if (!batch.empty() && nq + r->nq > sv->max_batch) break;
    const int mp = head->target == 0 ? std::max(maxp, std::min(r->nprobe, sv->ix->nlist)) : 1;
    const int mk = std::max(maxk, r->k);
    const int64_t b = batch_bytes(sv, head->target, nq + r->nq, mk, kf, mp);
    // admission by memory: the first request of an idle server always runs
    if (sv->in_use + b > sv->budget && !(batch.empty() && sv->inflight == 0)) break;
    batch.push_back(r);
    nq += r->nq;
    maxp = mp;
    maxk = mk;
    bytes = b;
    it = sv->pending.erase(it);
  }
  return batch;
}

void launch(mvg_server *sv, Slot &sl, std::vector<Request *> &batch, int64_t bytes) {
  const int target = batch[0]->target;
  const int d = target_d(sv, target);
  mvg_ivfpq *ix = target == 0 ? sv->ix : nullptr;
  int64_t nq = 0;
  int maxp = 1, k = 1;
  bool mixed_k = false;
  for (Request *r : batch) {
    nq += r->nq;
    k = std::max(k, r->k);
    mixed_k |= r->k != batch[0]->k;
  }
  const int kf = batch[0]->k_factor;
  sl.hq.reserve((size_t)nq * d);
  sl.hnp.reserve((size_t)nq);
  sl.hkq.reserve((size_t)nq);
  sl.hD.reserve((size_t)nq * k);
  sl.hI.reserve((size_t)nq * k);
  int64_t off = 0;
  for (Request *r : batch) {
    std::copy(r->q.begin(), r->q.end(), sl.hq.ptr + off * d);
    const int np = ix ? std::min(r->nprobe, ix->nlist) : 1;
    maxp = std::max(maxp, np);
    std::fill(sl.hnp.ptr + off, sl.hnp.ptr + off + r->nq, np);
    std::fill(sl.hkq.ptr + off, sl.hkq.ptr + off + r->nq, r->k);
    off += r->nq;
  }
  sl.q.reserve(sizeof(float) * nq * d);
  sl.np.reserve(sizeof(int32_t) * nq);
  sl.kq.reserve(sizeof(int32_t) * nq);
  sl.D.reserve(sizeof(float) * nq * k);
  sl.I.reserve(sizeof(int64_t) * nq * k);
  MVG_HIP(hipMemcpyAsync(sl.q.ptr, sl.hq.ptr, sizeof(float) * nq * d, hipMemcpyHostToDevice, sl.s));
  if (ix)
    MVG_HIP(hipMemcpyAsync(sl.np.ptr, sl.hnp.ptr, sizeof(int32_t) * nq, hipMemcpyHostToDevice, sl.s));
  const int32_t *kq = nullptr;
  if (mixed_k) {
    MVG_HIP(hipMemcpyAsync(sl.kq.ptr, sl.hkq.ptr, sizeof(int32_t) * nq, hipMemcpyHostToDevice, sl.s));
    kq = sl.kq.as<int32_t>();
  }
  if (ix) {
    // the handle's search lock: a concurrent eager search cannot prepare
    // (re-upload) the index buffers while this batch is being enqueued
    std::lock_guard<std::mutex> lk(ix->search_mu);
    mvg::ivf_prepare(ix);
    mvg::ivf_search_with(ix, sl.q.as<float>(), nq, k, sl.np.as<int32_t>(), maxp, kf, nullptr, 0,
                         sl.D.as<float>(), sl.I.as<int64_t>(), sl.work, sl.s, kq);
  } else {
    const mvg_sq8 *h = sv->sq;
    mvg::Sq8SearchArgs a{};
    a.q = sl.q.as<float>();
    a.nq = nq;
    a.d = h->d;
    a.metric = h->metric;
    a.n = h->ntotal;
    a.codes = h->codes.as<uint8_t>();
    a.norms = h->norms.as<uint32_t>();
    a.sums = h->sums.as<uint32_t>();
    a.vectors = h->f32.as<float>();
    a.gmin = h->gmin;
    a.scale = h->scale;
    a.k = k;
    a.k_factor = kf;
    a.D = sl.D.as<float>();
    a.I = sl.I.as<int64_t>();
    a.use_mfma = h->use_mfma;
    a.kq = kq;
    mvg::sq8_search(a, sl.swork, sl.s);
  }
  MVG_HIP(hipMemcpyAsync(sl.hD.ptr, sl.D.ptr, sizeof(float) * nq * k, hipMemcpyDeviceToHost, sl.s));
  MVG_HIP(hipMemcpyAsync(sl.hI.ptr, sl.I.ptr, sizeof(int64_t) * nq * k, hipMemcpyDeviceToHost, sl.s));
  MVG_HIP(hipEventRecord(sl.done, sl.s));
  sl.reqs = batch;
  sl.nq = nq;
  sl.k = k;
  sl.bytes = bytes;
}

void dispatcher(mvg_server *sv) {
  (void)hipSetDevice(sv->device);
  for (;;) {
    bool progressed = false;
    for (auto &up : sv->slots) {  // retire finished batches
      Slot &sl = *up;
      if (!sl.busy) continue;
      const hipError_t e = hipEventQuery(sl.done);
      if (e == hipSuccess) {
        retire(sv, sl);
        progressed = true;
      } else if (e != hipErrorNotReady) {
        fail_slot(sv, sl, hipGetErrorString(e));
        progressed = true;
      }
    }
    for (auto &up : sv->slots) {  // admit waiting requests into free slots
      Slot &sl = *up;
      if (sl.busy) continue;
      std::vector<Request *> batch;
      int64_t bytes = 0;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        batch = take_batch(sv, bytes);
        if (batch.empty()) break;
        sl.busy = true;
        sl.reqs = batch;
        sv->in_use += bytes;
        sv->inflight++;
        sv->batches++;
        sv->max_inflight = std::max(sv->max_inflight, sv->inflight);
      }
      sl.bytes = bytes;
      try {
        launch(sv, sl, batch, bytes);
      } catch (const std::exception &ex) {
        // work enqueued before the failure still uses the slot's buffers
        (void)hipStreamSynchronize(sl.s);
        fail_slot(sv, sl, ex.what());
      }
      progressed = true;
    }
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->stop && sv->pending.empty() && sv->inflight == 0) return;
    if (!progressed) {
      if (sv->inflight > 0)
        sv->cv_work.wait_for(lk, std::chrono::microseconds(20));
      else
        sv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

// Validates the settings, creates the slots' streams and starts the dispatcher.
int start_server(mvg_ivfpq *ix, mvg_sq8 *sq, int slots, int max_batch_queries,
                 int64_t mem_budget_bytes, mvg_server **out) {
  if (slots < 1 || slots > 16 || max_batch_queries < 1 || mem_budget_bytes <= 0 || !out)
    throw mvg::Error(MVG_ERR_INVALID,
                     "server needs 1..16 slots, max_batch_queries >= 1 and a positive memory "
                     "budget");
  auto sv = std::make_unique<mvg_server>();
  sv->ix = ix;
  sv->sq = sq;
  MVG_HIP(hipGetDevice(&sv->device));
  sv->max_batch = max_batch_queries;
  sv->budget = mem_budget_bytes;
  for (int i = 0; i < slots; i++) {
    auto sl = std::make_unique<Slot>();
    MVG_HIP(hipStreamCreateWithFlags(&sl->s, hipStreamNonBlocking));
    MVG_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
    sv->slots.push_back(std::move(sl));
  }
  mvg_server *raw = sv.get();
  sv->th = std::thread(dispatcher, raw);
  *out = sv.release();
  return MVG_OK;
}

}  // namespace

extern "C" {

int mvg_server_create_multi(mvg_ivfpq *ix, mvg_sq8 *sq, int slots, int max_batch_queries,
                            int64_t mem_budget_bytes, mvg_server **out) {
  try {
    if (!ix && !sq) throw mvg::Error(MVG_ERR_INVALID, "server needs an index");
    if (ix && (!ix->trained || ix->ntotal == 0))
      throw mvg::Error(MVG_ERR_STATE, "server needs a trained, non-empty index");
    if (sq && sq->ntotal == 0) throw mvg::Error(MVG_ERR_STATE, "server needs a non-empty index");
    if (ix) {
      std::lock_guard<std::mutex> lk(ix->search_mu);
      mvg::ivf_prepare(ix);
    }
    const int rc = start_server(ix, sq, slots, max_batch_queries, mem_budget_bytes, out);
    if (rc == MVG_OK && ix) ix->servers++;
    return rc;
  } catch (const mvg::Error &e) {
    mvg::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    mvg::set_last_error(e.what());
    return MVG_ERR_DEVICE;
  }
}

int mvg_server_create(mvg_ivfpq *ix, int slots, int max_batch_queries, int64_t mem_budget_bytes,
                      mvg_server **out) {
  if (!ix) {
    mvg::set_last_error("server needs an index");
    return MVG_ERR_INVALID;
  }
  return mvg_server_create_multi(ix, nullptr, slots, max_batch_queries, mem_budget_bytes, out);
}

int mvg_server_create_sq8(mvg_sq8 *sq, int slots, int max_batch_queries, int64_t mem_budget_bytes,
                          mvg_server **out) {
  if (!sq) {
    mvg::set_last_error("server needs an index");
    return MVG_ERR_INVALID;
  }
  return mvg_server_create_multi(nullptr, sq, slots, max_batch_queries, mem_budget_bytes, out);
}

int mvg_server_submit_to(mvg_server *sv, int target, const float *q, int64_t nq, int k, int nprobe,
                         int k_factor, float *distances, int64_t *labels, int64_t *ticket) {
  if (!sv || nq < 1 || k < 1 || nprobe < 1 || k_factor < 1 || !q || !distances || !labels || !ticket) {
    mvg::set_last_error("invalid request");
    return MVG_ERR_INVALID;
  }
  if ((target == 0 && !sv->ix) || (target == 1 && !sv->sq) || target < 0 || target > 1) {
    mvg::set_last_error("the server holds no index for this target");
    return MVG_ERR_INVALID;
  }
  if ((int64_t)k * k_factor > mvg::kMaxCandidates) {
    mvg::set_last_error("k * k_factor exceeds the device limit of 4096");
    return MVG_ERR_INVALID;
  }
  auto r = std::make_unique<Request>();
  r->target = target;
  r->q.assign(q, q + nq * target_d(sv, target));
  r->nq = nq;
  r->k = k;
  r->nprobe = nprobe;
  r->k_factor = k_factor;
  r->D = distances;
  r->I = labels;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    r->ticket = sv->next_ticket++;
    *ticket = r->ticket;
    sv->pending.push_back(r.get());
    sv->all[r->ticket] = std::move(r);
  }
  sv->cv_work.notify_one();
  return MVG_OK;
}

int mvg_server_submit(mvg_server *sv, const float *q, int64_t nq, int k, int nprobe, int k_factor,
                      float *distances, int64_t *labels, int64_t *ticket) {
  if (!sv) {
    mvg::set_last_error("invalid request");
    return MVG_ERR_INVALID;
  }
  return mvg_server_submit_to(sv, sv->ix ? 0 : 1, q, nq, k, nprobe, k_factor, distances, labels, ticket);
}

int mvg_server_wait(mvg_server *sv, int64_t ticket, double *latency_ms) {
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->all.find(ticket);
  if (it == sv->all.end()) {
    mvg::set_last_error("unknown ticket");
    return MVG_ERR_INVALID;
  }
  Request *r = it->second.get();
  sv->cv_done.wait(lk, [&] { return r->done; });
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  const int st = r->status;
  if (st != MVG_OK) mvg::set_last_error(r->error);
  sv->all.erase(it);
  return st;
}

int mvg_server_stats(mvg_server *sv, int64_t *completed_queries, int64_t *batches,
                     int64_t *max_inflight) {
  std::lock_guard<std::mutex> lk(sv->mu);
  if (completed_queries) *completed_queries = sv->done_queries;
  if (batches) *batches = sv->batches;
  if (max_inflight) *max_inflight = sv->max_inflight;
  return MVG_OK;
}

void mvg_server_destroy(mvg_server *sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->th.joinable()) sv->th.join();
  for (auto &sl : sv->slots) {
    (void)hipStreamSynchronize(sl->s);
    (void)hipEventDestroy(sl->done);
    (void)hipStreamDestroy(sl->s);
  }
  if (sv->ix) sv->ix->servers--;
  delete sv;
}

}  // extern "C"
