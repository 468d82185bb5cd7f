// IVF-PQ and IVF-Flat snapshots in the reference's on-disk format, so that
// collections and indexes saved by either side load on the other:
//   collection file  mojovec/api/collection_ivfpq.mojo:335-476
//   index stream     mojovec/io/serialization.mojo:1029-1090 (IndexIVFPQ),
//                    :980-1020 (IndexIVFFlat),
//                    :229-262 (IndexFlat), :903-944 (ArrayInvertedLists),
//                    :946-975 (ProductQuantizer), :66-80,136-147 (Int, Bool)
//   checksum trailer mojovec/io/snapshot_file.mojo:8-108
// Integers are 8-byte little-endian, Bool is one byte, floats are raw FP32.
// Validation order and messages follow the reference readers; the file is
// read through one read-only mapping, and the lists are uploaded to HBM as
// one CSR on the first search.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "collection.h"
#include "common.h"
#include "handles.h"

namespace {

using mvg::coll_raise;

constexpr int64_t kMagicFlat = 0x4d4a4f46;      // serialization.mojo:21
constexpr int64_t kMagicIvfFlat = 0x4d4a4f49;   // :23
constexpr int64_t kMagicIvfPq = 0x4d4a4f50;     // :24
constexpr int64_t kMagicInvlists = 0x4d4a4f4c;  // :25
constexpr int64_t kMagicPq = 0x4d4a4f51;        // :26
constexpr int64_t kCollMagic = 0x43495650;      // collection_ivfpq.mojo:29
constexpr int64_t kCollVersion = 1;             // :30
constexpr uint64_t kSumMagic = 0x4D4F4A4F534E4150ull;  // snapshot_file.mojo:8
constexpr uint64_t kSumSeed = 0xD6E8FEB86659FD93ull;   // :9
constexpr size_t kSumTrailer = 16;                     // :10
constexpr size_t kSumChunk = (size_t)8 << 20;          // :11
constexpr int64_t kMaxSerialized = INT64_MAX;          // serialization.mojo:35
constexpr int kPqKsub = 256;

inline uint64_t rotl(uint64_t v, int a) { return (v << a) | (v >> (64 - a)); }

// _checksum_update (snapshot_file.mojo:19-42): one streaming chunk.
uint64_t checksum_chunk(uint64_t sum, const uint8_t *p, size_t n) {
  const size_t words = n / 8;
  for (size_t i = 0; i < words; i++) {
    uint64_t w;
    std::memcpy(&w, p + 8 * i, 8);
    sum ^= w;
    sum = rotl(sum, 27) ^ rotl(w, 17);
    sum ^= sum >> 23;
  }
  uint64_t tail = 0;
  for (size_t i = words * 8; i < n; i++) tail |= (uint64_t)p[i] << ((i - words * 8) * 8);
  sum ^= tail;
  sum ^= (uint64_t)n << 32;
  return rotl(sum, 11);
}

// _checksum_prefix (:45-59): the payload in 8 MiB chunks from offset 0.
uint64_t checksum_payload(const uint8_t *p, size_t n) {
  uint64_t sum = kSumSeed;
  for (size_t off = 0; off < n; off += kSumChunk)
    sum = checksum_chunk(sum, p + off, std::min(kSumChunk, n - off));
  return sum;
}

// check_size_limit / checked_byte_count (serialization.mojo:40-54).
void check_size_limit(int64_t size, int64_t max_allowed) {
  if (size < 0 || size > max_allowed)
    coll_raise("Security Error: Deserialized size is out of valid range or exceeds safety limits");
}

int64_t checked_byte_count(int64_t count, int64_t item) {
  if (count < 0 || item <= 0 || count > kMaxSerialized / item)
    coll_raise("Serialized byte count overflows the supported range.");
  return count * item;
}

void validate_dimension(int64_t d) {  // validation.mojo:11-13
  if (d <= 0 || d > 65536) coll_raise("dimension must be between 1 and 65536.");
}

// ArrayInvertedLists.resize (inverted_lists.mojo:119-130): capacity after one
// resize of a list of capacity `cap` to `size` entries.
int64_t resized_capacity(int64_t cap, int64_t size) {
  if (size <= cap) return cap;
  return cap == 0 ? std::max<int64_t>(16, size) : std::max(2 * cap, size);
}

// Capacity after appending one entry at a time up to `size` entries, as
// IndexIVFPQ.add does (index_ivf_pq.mojo:199-211).
int64_t appended_capacity(int64_t cap, int64_t size) {
  while (cap < size) cap = resized_capacity(cap, cap + 1);
  return cap;
}

// Coarse IndexFlat capacity after adding `n` rows to one of capacity `cap`
// (index_flat.mojo:199-206).
int64_t grown_flat_capacity(int64_t cap, int64_t n) { return n > cap ? std::max(2 * cap, n) : cap; }

// ---- writing ----------------------------------------------------------------

// Buffered writer that checksums exactly the 8 MiB chunks a reader of the
// finished file sees.
class Writer {
 public:
  explicit Writer(const std::string &path) : path_(path) {
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) coll_raise("Cannot open snapshot file for writing: " + path, MVG_ERR_STATE);
    buf_.reserve(kSumChunk);
  }
  ~Writer() {
    if (f_) std::fclose(f_);
  }
  void bytes(const void *src, size_t n) {
    const uint8_t *s = static_cast<const uint8_t *>(src);
    while (n > 0) {
      const size_t take = std::min(n, kSumChunk - buf_.size());
      buf_.insert(buf_.end(), s, s + take);
      s += take;
      n -= take;
      if (buf_.size() == kSumChunk) flush();
    }
  }
  void zeros(size_t n) {
    static const uint8_t z[4096] = {};
    while (n > 0) {
      const size_t take = std::min(n, sizeof(z));
      bytes(z, take);
      n -= take;
    }
  }
  void i64(int64_t v) { bytes(&v, 8); }  // write_int (serialization.mojo:66-71)
  void boolean(bool b) {                 // write_bool (:136-140)
    const uint8_t v = b ? 1 : 0;
    bytes(&v, 1);
  }
  void string(const std::string &s) {  // _write_string (collection_codec.mojo:49-53)
    i64((int64_t)s.size());
    bytes(s.data(), s.size());
  }
  // Ends the payload; with `trailer` appends {magic, checksum}
  // (append_snapshot_checksum, snapshot_file.mojo:74-84), then syncs.
  void close(bool trailer) {
    flush();
    if (trailer) {
      const uint64_t t[2] = {kSumMagic, sum_};
      put(t, sizeof(t));
    }
    if (std::fflush(f_) != 0 || ::fsync(::fileno(f_)) != 0)
      coll_raise("Failed to write snapshot file: " + path_, MVG_ERR_STATE);
    const int rc = std::fclose(f_);
    f_ = nullptr;
    if (rc != 0) coll_raise("Failed to write snapshot file: " + path_, MVG_ERR_STATE);
  }

 private:
  void put(const void *p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_) != n)
      coll_raise("Failed to write snapshot file: " + path_, MVG_ERR_STATE);
  }
  void flush() {
    if (buf_.empty()) return;
    sum_ = checksum_chunk(sum_, buf_.data(), buf_.size());
    put(buf_.data(), buf_.size());
    buf_.clear();
  }
  std::string path_;
  std::FILE *f_ = nullptr;
  std::vector<uint8_t> buf_;
  uint64_t sum_ = kSumSeed;
};

// write_index_flat (serialization.mojo:229-238): `rows` real rows, then the
// unused capacity as zeros.
void write_flat(Writer &w, int64_t d, int64_t rows, int64_t capacity, int metric,
                const std::vector<float> &data) {
  w.i64(kMagicFlat);
  w.i64(d);
  w.i64(rows);
  w.i64(capacity);
  w.i64(metric == MVG_METRIC_IP ? 1 : 0);
  w.bytes(data.data(), sizeof(float) * (size_t)(rows * d));
  w.zeros(sizeof(float) * (size_t)((capacity - rows) * d));
}

// Host copy of everything an IndexIVFPQ stream holds.
struct IvfImage {
  int64_t d = 0, nlist = 0, M = 0, nprobe = 1, ntotal = 0;
  bool trained = false;
  int metric = MVG_METRIC_L2;  // index metric: L2 or IP
  int64_t coarse_capacity = 1024, coarse_rows = 0;
  std::vector<float> coarse;  // coarse_rows x d
  std::vector<float> pq;      // M x 256 x dsub
  std::vector<int64_t> list_capacity;
  std::vector<std::vector<int64_t>> ids;
  std::vector<std::vector<uint8_t>> codes;
};

// write_index_ivf_pq (serialization.mojo:1029-1044) with its IndexFlat
// (:229-238), ArrayInvertedLists (:903-914) and ProductQuantizer (:946-952).
// The coarse rows past ntotal are the reference's unused capacity, written as
// zeros.
void write_ivf(Writer &w, const mvg_ivfpq *h, const std::vector<float> &coarse,
               const std::vector<float> &pq, int64_t coarse_capacity,
               const std::vector<int64_t> &list_capacity) {
  const int64_t d = h->d, M = h->M, nlist = h->nlist;
  const int64_t qrows = h->trained ? nlist : 0;
  w.i64(kMagicIvfPq);
  w.i64(d);
  w.i64(nlist);
  w.i64(M);
  w.i64(h->nprobe);
  w.i64(h->ntotal);
  w.boolean(h->trained);
  w.i64(h->metric == MVG_METRIC_IP ? 1 : 0);
  write_flat(w, d, qrows, coarse_capacity, h->metric, coarse);
  // inverted lists: ids widened to the reference's 8-byte Int
  w.i64(kMagicInvlists);
  w.i64(nlist);
  w.i64(M);
  std::vector<int64_t> wide;
  for (int64_t l = 0; l < nlist; l++) {
    const auto &li = h->lids[(size_t)l];
    const int64_t size = (int64_t)li.size();
    w.i64(size);
    w.i64(list_capacity[(size_t)l]);
    wide.assign(li.begin(), li.end());
    w.bytes(wide.data(), sizeof(int64_t) * wide.size());
    w.bytes(h->lcodes[(size_t)l].data(), h->lcodes[(size_t)l].size());
  }
  // product quantizer (untrained centroids are unset in the reference; zeros)
  w.i64(kMagicPq);
  w.i64(d);
  w.i64(M);
  w.i64(kPqKsub);
  w.boolean(h->trained);
  if (h->trained) w.bytes(pq.data(), sizeof(float) * pq.size());
  else w.zeros(sizeof(float) * (size_t)(d * kPqKsub));
}

void fetch_codebooks(mvg_ivfpq *h, std::vector<float> &coarse, std::vector<float> &pq) {
  coarse.assign((size_t)h->nlist * h->d, 0.0f);
  pq.assign((size_t)h->d * kPqKsub, 0.0f);
  if (h->trained) mvg::check_index(mvg_ivfpq_get_codebooks(h, coarse.data(), pq.data()));
}

std::vector<int64_t> default_list_capacity(const mvg_ivfpq *h) {
  std::vector<int64_t> cap((size_t)h->nlist);
  for (int l = 0; l < h->nlist; l++) cap[(size_t)l] = appended_capacity(0, (int64_t)h->lids[(size_t)l].size());
  return cap;
}

// atomic_temporary_path / atomic_replace (atomic_file.mojo): write beside the
// destination, then rename over it and sync the directory.
std::string temporary_path(const std::string &path) {
  const auto ns = std::chrono::steady_clock::now().time_since_epoch().count();
  return path + ".tmp." + std::to_string((long long)::getpid()) + "." + std::to_string((long long)ns);
}

void publish(const std::string &tmp, const std::string &path) {
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    coll_raise("Failed to replace snapshot file: " + path, MVG_ERR_STATE);
  }
  std::string dir = path;
  const size_t slash = dir.find_last_of('/');
  dir = slash == std::string::npos ? "." : (slash == 0 ? "/" : dir.substr(0, slash));
  const int fd = ::open(dir.c_str(), O_RDONLY | O_DIRECTORY);
  if (fd >= 0) {
    (void)::fsync(fd);
    ::close(fd);
  }
}

template <class Body>
void write_atomically(const std::string &path, Body body) {
  const std::string tmp = temporary_path(path);
  try {
    body(tmp);
  } catch (...) {
    std::remove(tmp.c_str());
    throw;
  }
  publish(tmp, path);
}

// ---- reading ----------------------------------------------------------------

class MappedFile {
 public:
  explicit MappedFile(const std::string &path) {
    fd_ = ::open(path.c_str(), O_RDONLY);
    if (fd_ < 0) coll_raise("Cannot open snapshot file: " + path, MVG_ERR_STATE);
    struct stat st;
    if (::fstat(fd_, &st) != 0) coll_raise("Cannot stat snapshot file: " + path, MVG_ERR_STATE);
    size_ = (size_t)st.st_size;
    if (size_ > 0) {
      void *p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p == MAP_FAILED) coll_raise("Cannot map snapshot file: " + path, MVG_ERR_STATE);
      data_ = static_cast<const uint8_t *>(p);
    }
  }
  ~MappedFile() {
    if (data_) ::munmap(const_cast<uint8_t *>(data_), size_);
    if (fd_ >= 0) ::close(fd_);
  }
  const uint8_t *data() const { return data_; }
  size_t size() const { return size_; }

 private:
  int fd_ = -1;
  const uint8_t *data_ = nullptr;
  size_t size_ = 0;
};

class Reader {
 public:
  Reader(const uint8_t *p, size_t n) : p_(p), n_(n) {}
  size_t pos() const { return pos_; }
  const uint8_t *take(size_t k, const char *short_msg = "Unexpected end of serialized data.") {
    if (k > n_ - pos_) coll_raise(short_msg);
    const uint8_t *at = p_ + pos_;
    pos_ += k;
    return at;
  }
  int64_t i64() {  // read_int (serialization.mojo:74-79)
    int64_t v;
    std::memcpy(&v, take(8), 8);
    return v;
  }
  bool boolean() {  // read_bool (:142-147)
    const uint8_t v = *take(1);
    if (v > 1) coll_raise("Invalid serialized Bool value.");
    return v == 1;
  }
  std::string string() {  // _read_string (collection_codec.mojo:56-64)
    const int64_t size = i64();
    check_size_limit(size, 1048576);
    const uint8_t *s = take((size_t)size, "Unexpected end of serialized String.");
    return std::string(reinterpret_cast<const char *>(s), (size_t)size);
  }
  template <class T>
  void array(T *out, int64_t count) {
    if (count == 0) return;
    const int64_t bytes = checked_byte_count(count, (int64_t)sizeof(T));
    std::memcpy(out, take((size_t)bytes), (size_t)bytes);
  }

 private:
  const uint8_t *p_;
  size_t n_, pos_ = 0;
};

// read_index_flat (serialization.mojo:240-262): header fields, capacity and
// the first qrows rows.
void read_flat(Reader &r, int64_t &qd, int64_t &qrows, int &qmetric, int64_t &capacity,
               std::vector<float> &rows) {
  if (r.i64() != kMagicFlat) coll_raise("Invalid magic for IndexFlat");
  qd = r.i64();
  validate_dimension(qd);
  qrows = r.i64();
  check_size_limit(qrows, 1000000000);
  const int64_t cap = r.i64();
  check_size_limit(cap, 1000000000);
  qmetric = r.i64() == 1 ? MVG_METRIC_IP : MVG_METRIC_L2;
  // The reference reads capacity rows and keeps ntotal of them; a header
  // whose ntotal exceeds its capacity would address rows that are not there.
  if (qrows > cap) coll_raise("IndexFlat size exceeds its capacity.");
  const int64_t floats = cap * qd;  // cap, qd bounded above: no overflow
  checked_byte_count(floats, 4);
  const uint8_t *src = r.take((size_t)(floats * 4));
  capacity = cap;
  rows.resize((size_t)(qrows * qd));
  std::memcpy(rows.data(), src, sizeof(float) * rows.size());
}

// read_invlists (serialization.mojo:916-944) into an index of nlist lists of
// code_size-byte codes.
void read_invlists(Reader &r, int64_t nlist_expect, int64_t code_size_expect,
                   std::vector<int64_t> &capacity, std::vector<std::vector<int64_t>> &ids,
                   std::vector<std::vector<uint8_t>> &codes) {
  if (r.i64() != kMagicInvlists) coll_raise("Invalid magic for ArrayInvertedLists");
  const int64_t nlist = r.i64();
  check_size_limit(nlist, 1000000);
  const int64_t code_size = r.i64();
  check_size_limit(code_size, 65536);
  if (nlist != nlist_expect || code_size != code_size_expect)
    coll_raise("Inverted-list shape does not match its index header.");
  ids.resize((size_t)nlist);
  codes.resize((size_t)nlist);
  capacity.resize((size_t)nlist);
  for (int64_t l = 0; l < nlist; l++) {
    const int64_t size = r.i64();
    const int64_t cap = r.i64();
    check_size_limit(cap, 1000000000);
    check_size_limit(size, cap);
    capacity[(size_t)l] = resized_capacity(0, cap);  // resize(i, capacity)
    ids[(size_t)l].resize((size_t)size);
    r.array(ids[(size_t)l].data(), size);
    codes[(size_t)l].resize((size_t)(size * code_size));
    r.array(codes[(size_t)l].data(), size * code_size);
  }
}

// read_index_ivf_pq (serialization.mojo:1046-1090) with read_invlists
// (:916-944) and read_pq (:954-975).
IvfImage read_ivf(Reader &r) {
  IvfImage img;
  if (r.i64() != kMagicIvfPq) coll_raise("Invalid magic for IndexIVFPQ");
  img.d = r.i64();
  validate_dimension(img.d);
  img.nlist = r.i64();
  check_size_limit(img.nlist, 1000000);
  img.M = r.i64();
  check_size_limit(img.M, img.d);
  if (img.M <= 0 || img.d % img.M != 0) coll_raise("Invalid IndexIVFPQ subvector count.");
  img.nprobe = r.i64();
  check_size_limit(img.nprobe, img.nlist);
  if (img.nprobe <= 0) coll_raise("Invalid IndexIVFPQ nprobe.");
  img.ntotal = r.i64();
  check_size_limit(img.ntotal, 1000000000);
  img.trained = r.boolean();
  const int64_t metric_int = r.i64();
  if (metric_int != 0 && metric_int != 1) coll_raise("Invalid IndexIVFPQ metric.");
  img.metric = metric_int == 1 ? MVG_METRIC_IP : MVG_METRIC_L2;

  int64_t qd = 0, qrows = 0;
  int qmetric = 0;
  read_flat(r, qd, qrows, qmetric, img.coarse_capacity, img.coarse);
  img.coarse_rows = qrows;
  if (qd != img.d || (img.trained && qrows != img.nlist) || (!img.trained && qrows != 0) ||
      qmetric != img.metric)
    coll_raise("IndexIVFPQ coarse quantizer header mismatch.");

  read_invlists(r, img.nlist, img.M, img.list_capacity, img.ids, img.codes);
  const int64_t code_size = img.M;

  if (r.i64() != kMagicPq) coll_raise("Invalid magic for ProductQuantizer");
  const int64_t pd = r.i64();
  check_size_limit(pd, 65536);
  const int64_t pM = r.i64();
  check_size_limit(pM, 65536);
  const int64_t ksub = r.i64();
  check_size_limit(ksub, 256);
  if (pd <= 0 || pM <= 0 || pM > pd || pd % pM != 0) coll_raise("Invalid ProductQuantizer shape.");
  if (ksub <= 0 || ksub > 256) coll_raise("Invalid ProductQuantizer centroid count.");
  if (pd != img.d || pM != img.M || ksub != kPqKsub)
    coll_raise("ProductQuantizer shape does not match its index header.");
  const bool pq_trained = r.boolean();
  img.pq.resize((size_t)(pd * ksub));
  r.array(img.pq.data(), pd * ksub);
  if (code_size != img.M) coll_raise("IndexIVFPQ code size does not match M.");
  if (pq_trained != img.trained) coll_raise("IndexIVFPQ training state mismatch.");
  return img;
}

// Creates the device index from an image: codebooks to HBM, lists into the
// host append lists (uploaded as CSR on first search). Snapshots hold no FP32
// rows, so a loaded index cannot rerank (k_factor 1 only) until rebuilt.
mvg_ivfpq *install(const IvfImage &img) {
  int64_t entries = 0;
  for (int64_t l = 0; l < img.nlist; l++) {
    entries += (int64_t)img.ids[(size_t)l].size();
    for (int64_t id : img.ids[(size_t)l])
      if (id < 0 || id > INT32_MAX) coll_raise("IVF-PQ snapshot contains an invalid internal ID.");
  }
  if (img.ntotal > INT32_MAX) coll_raise("IVF-PQ holds at most 2^31-1 vectors");
  // Every add appends each vector to one list (index_ivf_pq.mojo:199-212), so
  // the lists hold ntotal entries; a file that disagrees is not an index the
  // reference wrote, and the per-entry device terms are sized by the lists.
  if (entries != img.ntotal) coll_raise("IVF-PQ inverted lists do not hold ntotal entries.");
  mvg_ivfpq *raw = nullptr;
  mvg::check_index(mvg_ivfpq_create((int)img.d, (int)img.nlist, (int)img.M, img.metric, &raw));
  std::unique_ptr<mvg_ivfpq, void (*)(mvg_ivfpq *)> h(raw, mvg_ivfpq_destroy);
  mvg::check_index(mvg_ivfpq_set_keep_vectors(h.get(), 0));
  mvg::check_index(mvg_ivfpq_set_nprobe(h.get(), (int)img.nprobe));
  if (img.trained) mvg::check_index(mvg_ivfpq_set_codebooks(h.get(), img.coarse.data(), img.pq.data()));
  for (int64_t l = 0; l < img.nlist; l++) {
    h->lids[(size_t)l].assign(img.ids[(size_t)l].begin(), img.ids[(size_t)l].end());
    h->lcodes[(size_t)l] = img.codes[(size_t)l];
  }
  h->ntotal = img.ntotal;
  h->dirty = true;
  return h.release();
}

// ---- IVF-Flat ------------------------------------------------------------------

// write_index_ivf_flat (serialization.mojo:980-993): header, the quantizer
// IndexFlat with its own metric, inverted lists of code_size 4d whose codes
// are the raw FP32 rows.
void write_ivfflat(Writer &w, const mvg_ivfflat *h, const std::vector<float> &coarse,
                   int64_t coarse_capacity, const std::vector<int64_t> &list_capacity) {
  const int64_t d = h->d, nlist = h->nlist;
  w.i64(kMagicIvfFlat);
  w.i64(d);
  w.i64(nlist);
  w.i64(h->nprobe);
  w.i64(h->ntotal);
  w.boolean(h->trained);
  w.i64(h->metric == MVG_METRIC_IP ? 1 : 0);
  write_flat(w, d, h->trained ? nlist : 0, coarse_capacity, h->quantizer_metric, coarse);
  w.i64(kMagicInvlists);
  w.i64(nlist);
  w.i64(4 * d);
  std::vector<int64_t> wide;
  for (int64_t l = 0; l < nlist; l++) {
    const auto &li = h->lids[(size_t)l];
    w.i64((int64_t)li.size());
    w.i64(list_capacity[(size_t)l]);
    wide.assign(li.begin(), li.end());
    w.bytes(wide.data(), sizeof(int64_t) * wide.size());
    w.bytes(h->lrows[(size_t)l].data(), sizeof(float) * h->lrows[(size_t)l].size());
  }
}

struct IvfFlatImage {
  int64_t d = 0, nlist = 0, nprobe = 1, ntotal = 0;
  bool trained = false;
  int metric = MVG_METRIC_L2, quantizer_metric = MVG_METRIC_L2;
  int64_t coarse_capacity = 1024, coarse_rows = 0;
  std::vector<float> coarse;
  std::vector<int64_t> list_capacity;
  std::vector<std::vector<int64_t>> ids;
  std::vector<std::vector<uint8_t>> rows;  // 4d bytes per entry
};

// read_index_ivf_flat (serialization.mojo:995-1020). The reference takes any
// metric value other than 1 as L2 and does not compare the quantizer with the
// header; here a quantizer whose dimension or row count contradicts the
// header, a non-positive nprobe and ids outside [0, 2^31 - 1] are rejected,
// since the index could not search them.
IvfFlatImage read_ivfflat(Reader &r) {
  IvfFlatImage img;
  if (r.i64() != kMagicIvfFlat) coll_raise("Invalid magic for IndexIVFFlat");
  img.d = r.i64();
  validate_dimension(img.d);
  img.nlist = r.i64();
  check_size_limit(img.nlist, 1000000);
  img.nprobe = r.i64();
  check_size_limit(img.nprobe, 1000000);
  img.ntotal = r.i64();
  check_size_limit(img.ntotal, 1000000000);
  img.trained = r.boolean();
  img.metric = r.i64() == 1 ? MVG_METRIC_IP : MVG_METRIC_L2;
  int64_t qd = 0;
  read_flat(r, qd, img.coarse_rows, img.quantizer_metric, img.coarse_capacity, img.coarse);
  read_invlists(r, img.nlist, 4 * img.d, img.list_capacity, img.ids, img.rows);
  if (img.nlist <= 0) coll_raise("IVF-Flat nlist must be between 1 and 1000000.");
  if (img.nprobe <= 0) coll_raise("Invalid IndexIVFFlat nprobe.");
  if (qd != img.d || (img.trained && img.coarse_rows != img.nlist) ||
      (!img.trained && img.coarse_rows != 0))
    coll_raise("IndexIVFFlat coarse quantizer header mismatch.");
  for (const auto &li : img.ids)
    for (int64_t id : li)
      if (id < 0 || id > INT32_MAX) coll_raise("IVF-Flat snapshot contains an id outside [0, 2^31 - 1].");
  return img;
}

mvg_ivfflat *install_ivfflat(const IvfFlatImage &img) {
  mvg_ivfflat *raw = nullptr;
  mvg::check_index(mvg_ivfflat_create_with_quantizer((int)img.d, (int)img.nlist, img.metric,
                                                     img.quantizer_metric, &raw));
  std::unique_ptr<mvg_ivfflat, void (*)(mvg_ivfflat *)> h(raw, mvg_ivfflat_destroy);
  mvg::check_index(mvg_ivfflat_set_nprobe(h.get(), (int)std::min<int64_t>(img.nprobe, INT32_MAX)));
  if (img.trained) mvg::check_index(mvg_ivfflat_set_centroids(h.get(), img.coarse.data()));
  for (int64_t l = 0; l < img.nlist; l++) {
    h->lids[(size_t)l].assign(img.ids[(size_t)l].begin(), img.ids[(size_t)l].end());
    const auto &b = img.rows[(size_t)l];
    h->lrows[(size_t)l].resize(b.size() / sizeof(float));
    if (!b.empty()) std::memcpy(h->lrows[(size_t)l].data(), b.data(), b.size());
  }
  h->ntotal = img.ntotal;
  h->coarse_capacity = img.coarse_capacity;
  h->list_capacity = img.list_capacity;
  h->dirty = true;
  return h.release();
}

// validate_snapshot_checksum (snapshot_file.mojo:87-108); returns the payload size.
size_t validate_checksum(const MappedFile &f) {
  if (f.size() < kSumTrailer) coll_raise("IVF-PQ snapshot checksum trailer is missing.");
  const size_t payload = f.size() - kSumTrailer;
  uint64_t t[2];
  std::memcpy(t, f.data() + payload, sizeof(t));
  if (t[0] != kSumMagic) coll_raise("Invalid snapshot checksum trailer.");
  if (checksum_payload(f.data(), payload) != t[1]) coll_raise("Snapshot checksum mismatch.");
  return payload;
}

}  // namespace

extern "C" {

int mvg_snapshot_checksum(const uint8_t *data, int64_t len, uint64_t *out) {
  COLL_BEGIN
  if (len < 0 || (len > 0 && !data) || !out) coll_raise("invalid arguments");
  *out = checksum_payload(data, (size_t)len);
  COLL_END
}

int mvg_ivfpq_save(mvg_ivfpq *h, const char *path) {
  COLL_BEGIN
  if (!h || !path) coll_raise("invalid arguments");
  std::vector<float> coarse, pq;
  fetch_codebooks(h, coarse, pq);
  const int64_t qcap = grown_flat_capacity(1024, h->trained ? h->nlist : 0);
  const std::vector<int64_t> lcap = default_list_capacity(h);
  write_atomically(path, [&](const std::string &tmp) {
    Writer w(tmp);
    write_ivf(w, h, coarse, pq, qcap, lcap);
    w.close(false);
  });
  COLL_END
}

int mvg_ivfpq_load(const char *path, mvg_ivfpq **out) {
  COLL_BEGIN
  if (!path || !out) coll_raise("invalid arguments");
  MappedFile f(path);
  Reader r(f.data(), f.size());
  const IvfImage img = read_ivf(r);
  *out = install(img);
  COLL_END
}

int mvg_ivfflat_save(mvg_ivfflat *h, const char *path) {
  COLL_BEGIN
  if (!h || !path) coll_raise("invalid arguments");
  std::vector<float> coarse((size_t)h->nlist * h->d, 0.0f);
  if (h->trained) mvg::check_index(mvg_ivfflat_get_centroids(h, coarse.data()));
  const int64_t qcap = grown_flat_capacity(h->coarse_capacity, h->trained ? h->nlist : 0);
  std::vector<int64_t> lcap((size_t)h->nlist);
  for (int l = 0; l < h->nlist; l++)
    lcap[(size_t)l] = appended_capacity(h->list_capacity[(size_t)l], (int64_t)h->lids[(size_t)l].size());
  write_atomically(path, [&](const std::string &tmp) {
    Writer w(tmp);
    write_ivfflat(w, h, coarse, qcap, lcap);
    w.close(false);
  });
  h->coarse_capacity = qcap;
  h->list_capacity = lcap;
  COLL_END
}

int mvg_ivfflat_load(const char *path, mvg_ivfflat **out) {
  COLL_BEGIN
  if (!path || !out) coll_raise("invalid arguments");
  MappedFile f(path);
  Reader r(f.data(), f.size());
  const IvfFlatImage img = read_ivfflat(r);
  *out = install_ivfflat(img);
  COLL_END
}

int mvg_coll_save(mvg_collection *h, const char *path) {
  COLL_BEGIN
  if (!h || !path) coll_raise("invalid arguments");
  mvg_ivfpq *ix = h->ix;
  std::vector<float> coarse, pq;
  fetch_codebooks(ix, coarse, pq);
  const int64_t qcap = grown_flat_capacity(h->coarse_capacity, ix->trained ? h->nlist : 0);
  std::vector<int64_t> lcap((size_t)h->nlist);
  for (int l = 0; l < h->nlist; l++)
    lcap[(size_t)l] = appended_capacity(h->list_capacity[(size_t)l], (int64_t)ix->lids[(size_t)l].size());
  write_atomically(path, [&](const std::string &tmp) {
    Writer w(tmp);  // collection_ivfpq.mojo:349-362
    w.i64(kCollMagic);
    w.i64(kCollVersion);
    w.i64(h->d);
    w.string(h->name);
    w.string(mvg::metric_name(h->metric));
    w.i64((int64_t)h->user_ids.size());
    w.bytes(h->user_ids.data(), sizeof(int64_t) * h->user_ids.size());
    write_ivf(w, ix, coarse, pq, qcap, lcap);
    w.close(true);
  });
  h->coarse_capacity = qcap;
  h->list_capacity = lcap;
  COLL_END
}

int mvg_coll_load(const char *path, mvg_collection **out) {
  COLL_BEGIN
  if (!path || !out) coll_raise("invalid arguments");
  MappedFile f(path);  // CollectionIVFPQ.load (collection_ivfpq.mojo:372-476)
  const size_t payload = validate_checksum(f);
  Reader r(f.data(), f.size());
  if (r.i64() != kCollMagic) coll_raise("Invalid CollectionIVFPQ snapshot magic.");
  if (r.i64() != kCollVersion) coll_raise("Unsupported CollectionIVFPQ snapshot version.");
  const int64_t dimension = r.i64();
  validate_dimension(dimension);
  const std::string name = r.string();
  const std::string metric = r.string();
  const int metric_type = mvg::parse_metric(metric);
  const int64_t num_ids = r.i64();
  check_size_limit(num_ids, 1000000000);
  const int64_t ids_bytes = checked_byte_count(num_ids, 8);
  if (ids_bytes > (int64_t)payload - (int64_t)r.pos())
    coll_raise("IVF-PQ ID data exceeds the snapshot payload.");
  std::vector<int64_t> ids((size_t)num_ids);
  r.array(ids.data(), num_ids);
  const IvfImage img = read_ivf(r);
  if (r.pos() != payload) coll_raise("IVF-PQ snapshot contains unexpected payload data.");
  mvg::validate_ivfpq_parameters(dimension, img.nlist, img.M, img.nprobe);
  if (img.d != dimension || img.ntotal != num_ids)
    coll_raise("IVF-PQ snapshot dimensions or record count mismatch.");
  const int index_metric = metric_type == MVG_METRIC_L2 ? MVG_METRIC_L2 : MVG_METRIC_IP;
  if (img.metric != index_metric) coll_raise("IVF-PQ snapshot metric mismatch.");
  if ((img.trained && img.coarse_rows != img.nlist) || (!img.trained && img.coarse_rows != 0))
    coll_raise("IVF-PQ coarse quantizer does not match its index.");
  std::vector<uint8_t> seen((size_t)num_ids, 0);
  int64_t indexed = 0;
  for (int64_t l = 0; l < img.nlist; l++) {
    indexed += (int64_t)img.ids[(size_t)l].size();
    for (int64_t id : img.ids[(size_t)l]) {
      if (id < 0 || id >= num_ids) coll_raise("IVF-PQ snapshot contains an invalid internal ID.");
      if (seen[(size_t)id]) coll_raise("IVF-PQ snapshot contains a duplicate internal ID.");
      seen[(size_t)id] = 1;
    }
  }
  if (indexed != num_ids) coll_raise("IVF-PQ inverted-list count does not match its IDs.");

  auto c = std::make_unique<mvg_collection>();
  c->name = name;
  c->d = (int)dimension;
  c->metric = metric_type;
  c->nlist = (int)img.nlist;
  c->M = (int)img.M;
  c->coarse_capacity = img.coarse_capacity;
  c->list_capacity = img.list_capacity;
  for (int64_t i = 0; i < num_ids; i++) {
    if (c->id_to_internal.count(ids[(size_t)i])) coll_raise("IVF-PQ snapshot contains duplicate IDs.");
    c->id_to_internal[ids[(size_t)i]] = i;
  }
  c->user_ids = std::move(ids);
  c->ix = install(img);
  *out = c.release();
  COLL_END
}

}  // extern "C"
