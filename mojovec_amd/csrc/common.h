// Shared host/device helpers for the MI355X vector-scan library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace mvg {

constexpr int kMetricL2 = 0;
constexpr int kMetricIP = 1;
constexpr int kKsub = 256;          // PQ centroids per subspace (8-bit codes)
constexpr int kMaxCandidates = 4096; // k * k_factor ceiling on the device path
constexpr float kIvfPadDist = 1e38f;  // index_ivf_pq.mojo:284-286
constexpr uint32_t kSentinel = 0xFFFFFFFFu;
constexpr uint64_t kSentinel64 = ~0ull;

// Error carried from a failing call to the C-ABI wrapper, which stores the
// message in the thread-local last-error slot and returns a status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum Status { kOk = 0, kInvalid = 1, kDevice = 2, kAlloc = 3, kState = 4 };

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw Error(e == hipErrorOutOfMemory ? kAlloc : kDevice,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define MVG_HIP(call) ::mvg::hip_check((call), #call)
#define MVG_LAUNCH_CHECK(name) ::mvg::hip_check(hipGetLastError(), name)

// Fixed device region handed out front to back (256-byte aligned pieces);
// nothing is ever freed back. Lets a search run with no device allocation.
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  void *take(size_t n) {
    const size_t at = (used + 255) & ~size_t(255);
    if (at > size || n > size - at)
      throw Error(kInvalid, "the request needs more than the " + std::to_string(size >> 20) +
                                " MiB scratch of the allocation-free entry point");
    used = at + n;
    return base + at;
  }
  size_t left() const { return used < size ? size - used : 0; }
};

// Grow-only device buffer. Contents are not preserved by reserve(). Bound
// to an arena, it takes its memory from there (no hipMalloc / hipFree).
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  Arena *arena = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr && !arena) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  void reserve(size_t n) {
    if (n <= bytes) return;
    release();
    if (n == 0) return;
    if (arena) {
      ptr = arena->take(n);
    } else {
      MVG_HIP(hipMalloc(&ptr, n));
    }
    bytes = n;
  }
  // Grow while keeping the first `keep` bytes.
  void grow_keep(size_t n, size_t keep, hipStream_t s) {
    if (n <= bytes) return;
    if (arena) throw Error(kInvalid, "arena buffers cannot grow in place");
    void *np = nullptr;
    MVG_HIP(hipMalloc(&np, n));
    if (ptr && keep) MVG_HIP(hipMemcpyAsync(np, ptr, keep, hipMemcpyDeviceToDevice, s));
    MVG_HIP(hipStreamSynchronize(s));
    if (ptr) (void)hipFree(ptr);
    ptr = np;
    bytes = n;
  }
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Thread-local message read by mvg_last_error().
void set_last_error(const std::string &msg);

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Orderable keys: a total order on float distances matching `<` on finite
// values; -0.0 is folded onto +0.0 so equal distances compare equal.
__host__ __device__ inline uint32_t f2key(float f) {
  f = f + 0.0f;
  uint32_t u;
  memcpy(&u, &f, 4);
  // negative: all bits flipped; otherwise the sign bit set
  return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}

__host__ __device__ inline float key2f(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__host__ __device__ inline uint64_t pack_key(float d, uint32_t id) {
  return ((uint64_t)f2key(d) << 32) | id;
}

}  // namespace mvg
