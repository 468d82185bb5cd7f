// Internal layout of the collection handle, shared by the collection API
// (collection.cpp) and its snapshot reader/writer (snapshot.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mojovec_gpu.h"
#include "common.h"

namespace mvg {

constexpr int kMetricCosine = 2;  // collection_codec.mojo:21
constexpr int kCollMaxNlist = 1000000;
constexpr int kCollMaxNprobe = 65536;

// Raised inside the collection layer; the C ABI turns it into a status code
// and mvg_last_error().
struct CollFail {
  int code;
  std::string msg;
};

[[noreturn]] inline void coll_raise(const std::string &msg, int code = MVG_ERR_INVALID) {
  throw CollFail{code, msg};
}

inline void check_index(int rc) {
  if (rc != MVG_OK) throw CollFail{rc, mvg_last_error()};
}

// _validate_ivfpq_parameters (collection_ivfpq.mojo:37-49) and
// validation.mojo:11-14, in the reference's order.
void validate_ivfpq_parameters(int64_t d, int64_t nlist, int64_t M, int64_t nprobe);

// "l2" / "cosine" / "ip" (collection_codec.mojo:24-31); raises otherwise.
int parse_metric(const std::string &m);
const char *metric_name(int metric);

}  // namespace mvg

struct mvg_collection {
  std::string name;
  int d = 0, metric = 0, nlist = 0, M = 0;
  mvg_ivfpq *ix = nullptr;
  std::vector<int64_t> user_ids;
  std::unordered_map<int64_t, int64_t> id_to_internal;
  int niter_coarse = 15, niter_pq = 15;  // index_ivf_pq.mojo:98, pq.mojo:64
  uint64_t seed = 0x4D4F4A4F564543ull;
  std::vector<float> D;
  std::vector<int64_t> I;
  // Capacities the reference's storage would report for this content: the
  // coarse IndexFlat starts at 1024 rows (index_flat.mojo:134) and each
  // inverted list grows 0 -> max(16, n) -> doubling (inverted_lists.mojo:119-150).
  // Kept so that a saved snapshot carries the same header fields.
  int64_t coarse_capacity = 1024;
  std::vector<int64_t> list_capacity;

  ~mvg_collection() { mvg_ivfpq_destroy(ix); }
};

#define COLL_BEGIN try {
#define COLL_END                                              \
  }                                                           \
  catch (const mvg::CollFail &f) {                            \
    mvg::set_last_error(f.msg);                               \
    return f.code;                                            \
  }                                                           \
  catch (const mvg::Error &e) {                               \
    mvg::set_last_error(e.what());                            \
    return e.code;                                            \
  }                                                           \
  catch (const std::bad_alloc &) {                            \
    mvg::set_last_error("host allocation failed");            \
    return MVG_ERR_ALLOC;                                     \
  }                                                           \
  catch (const std::exception &e) {                           \
    mvg::set_last_error(e.what());                            \
    return MVG_ERR_STATE;                                     \
  }                                                           \
  return MVG_OK;
