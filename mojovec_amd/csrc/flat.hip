// Exact Flat search, IVF coarse assignment and k-means assignment.
//
// Plain FP32 form of IndexFlat._search_impl (mojovec/index/index_flat.mojo:
// 260-325): distance by l2_distance_simd[8] / -inner_product_simd[8] in the
// reference's operation order, canonical top-k. The work is split into
// (query tile x base chunk) workgroups: each thread keeps one base vector in
// registers and streams the tile's queries from LDS (broadcast reads), the
// chunk's keys stay in LDS and a radix select keeps the chunk's top-k; a merge
// kernel reduces the per-chunk winners of each query.
#include "dist.h"
#include "kernels.h"
#include "select.h"

namespace mvg {

namespace {

constexpr int kTQ = 8;      // queries per workgroup
constexpr int kCB = 1024;   // base vectors per chunk
constexpr int kPerThread = kCB / kNT;
constexpr int kMergeCap = 8192;

// Distance of LDS query `qs` to base row `xg` (global) or its register copy.
template <int D, bool IP, bool KMEANS>
__device__ __forceinline__ float pair_dist(const float *xr, const float *xg, const float *qs, int d) {
  if constexpr (D > 0) {
    return IP ? -ip_w8_fixed<D>(xr, qs) : l2_w8_fixed<D>(xr, qs);
  } else {
    if (IP) return -ip_wn<8>(qs, xg, d);
    return KMEANS ? l2_kmeans(qs, xg, d) : l2_w8(qs, xg, d);
  }
}

// One workgroup: kTQ queries x one chunk of kCB base rows -> per-query top-k
// (u64 canonical keys) of the chunk.
template <int D, bool IP>
__global__ __launch_bounds__(kNT) void flat_chunk_topk_kernel(
    const float *__restrict__ base, int64_t n, int d, const float *__restrict__ q, int64_t nq,
    int k, const uint8_t *__restrict__ excl, int64_t excl_len, uint64_t *__restrict__ part,
    int nchunks) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *qs = reinterpret_cast<float *>(smem);                        // kTQ * d
  uint32_t *keys = reinterpret_cast<uint32_t *>(qs + kTQ * d);         // kTQ * kCB
  uint32_t *sel = keys + kTQ * kCB;                                    // k
  SelectScratch *scr = reinterpret_cast<SelectScratch *>(sel + ((k + 3) & ~3));

  const int tid = threadIdx.x;
  const int64_t q0 = (int64_t)blockIdx.x * kTQ;
  const int chunk = blockIdx.y;
  const int64_t c0 = (int64_t)chunk * kCB;
  const int qcount = (int)min<int64_t>(kTQ, nq - q0);

  for (int i = tid; i < kTQ * d; i += kNT) {
    int qi = i / d;
    qs[i] = qi < qcount ? q[(q0 + qi) * d + (i - qi * d)] : 0.0f;
  }
  __syncthreads();

  for (int r = 0; r < kPerThread; r++) {
    const int local = r * kNT + tid;
    const int64_t j = c0 + local;
    const bool live = j < n && !(excl && j < excl_len && excl[j] > 0);
    float xr[D > 0 ? D : 1];
    const float *xg = base + (live ? j : 0) * d;
    if constexpr (D > 0) {
      if (live) {
#pragma unroll
        for (int e = 0; e < D; e += 4) {
          float4 t = *reinterpret_cast<const float4 *>(xg + e);
          xr[e] = t.x; xr[e + 1] = t.y; xr[e + 2] = t.z; xr[e + 3] = t.w;
        }
      }
    }
    for (int qi = 0; qi < kTQ; qi++) {
      uint32_t key = kSentinel;
      if (live && qi < qcount) key = f2key(pair_dist<D, IP, false>(xr, xg, qs + qi * d, d));
      keys[qi * kCB + local] = key;
    }
  }
  __syncthreads();

  const int clen = (int)min<int64_t>(kCB, n - c0);
  if (k <= 64) {
    // wavefront top-k: each wavefront folds its queries' chunk keys into the
    // k smallest (key, row) composites held one per lane, no workgroup barrier
    const int lane = tid & 63, wave = tid >> 6;
    for (int qi = wave; qi < qcount; qi += kNW) {
      const uint64_t best = wave_fold_keys<kCB>(keys + qi * kCB, clen, (uint32_t)c0, kSentinel64, k);
      if (lane < k) part[((q0 + qi) * nchunks + chunk) * (int64_t)k + lane] = best;
    }
    return;
  }
  for (int qi = 0; qi < qcount; qi++) {
    const int cnt = select_topk_u32(keys + qi * kCB, clen, k, *scr, sel);
    uint64_t *dst = part + ((q0 + qi) * nchunks + chunk) * (int64_t)k;
    for (int t = tid; t < k; t += kNT) {
      uint64_t v = kSentinel64;
      if (t < cnt) {
        const uint32_t idx = sel[t];
        v = ((uint64_t)keys[qi * kCB + idx] << 32) | (uint32_t)(c0 + idx);
      }
      dst[t] = v;
    }
    __syncthreads();
  }
}

// Canonical argmin over all base rows for kTQ queries per workgroup. With
// KMEANS the reference's `min_dist = 1e38; if dist < min_dist` scan is
// restated (kmeans.mojo:124-147): no candidate at or above 1e38, and the
// fallback cluster is 0.
template <int D, bool IP, bool KMEANS>
__global__ __launch_bounds__(kNT) void flat_argmin_kernel(
    const float *__restrict__ base, int64_t n, int d, const float *__restrict__ q, int64_t nq,
    int32_t *__restrict__ out_id, float *__restrict__ out_dist) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *qs = reinterpret_cast<float *>(smem);
  uint64_t *red = reinterpret_cast<uint64_t *>(qs + kTQ * d);  // kTQ * kNW
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t q0 = (int64_t)blockIdx.x * kTQ;
  const int qcount = (int)min<int64_t>(kTQ, nq - q0);
  for (int i = tid; i < kTQ * d; i += kNT) {
    int qi = i / d;
    qs[i] = qi < qcount ? q[(q0 + qi) * d + (i - qi * d)] : 0.0f;
  }
  __syncthreads();
  uint64_t best[kTQ];
#pragma unroll
  for (int qi = 0; qi < kTQ; qi++) best[qi] = kSentinel64;
  for (int64_t j = tid; j < n; j += kNT) {
    float xr[D > 0 ? D : 1];
    const float *xg = base + j * d;
    if constexpr (D > 0) {
#pragma unroll
      for (int e = 0; e < D; e += 4) {
        float4 t = *reinterpret_cast<const float4 *>(xg + e);
        xr[e] = t.x; xr[e + 1] = t.y; xr[e + 2] = t.z; xr[e + 3] = t.w;
      }
    }
#pragma unroll
    for (int qi = 0; qi < kTQ; qi++) {
      if (qi < qcount) {
        const float dist = pair_dist<D, IP, KMEANS>(xr, xg, qs + qi * d, d);
        if (!KMEANS || dist < 1e38f) {
          uint64_t key = pack_key(dist, (uint32_t)j);
          best[qi] = key < best[qi] ? key : best[qi];
        }
      }
    }
  }
#pragma unroll
  for (int qi = 0; qi < kTQ; qi++) {
    uint64_t v = best[qi];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      uint64_t w = __shfl_xor(v, o, 64);
      v = w < v ? w : v;
    }
    if (lane == 0) red[qi * kNW + wave] = v;
  }
  __syncthreads();
  if (tid < qcount) {
    uint64_t v = red[tid * kNW];
    for (int w = 1; w < kNW; w++) v = red[tid * kNW + w] < v ? red[tid * kNW + w] : v;
    int32_t id;
    float dist;
    if (v == kSentinel64) {
      id = KMEANS ? 0 : -1;
      dist = 1e38f;
    } else {
      id = (int32_t)(v & 0xFFFFFFFFu);
      dist = key2f((uint32_t)(v >> 32));
    }
    out_id[q0 + tid] = id;
    if (out_dist) out_dist[q0 + tid] = dist;
  }
}

}  // namespace

// Per-query reduction of candidate keys to the final sorted top-k. Also
// used by the IVF-PQ path.
__global__ __launch_bounds__(kNT) void merge_topk_kernel(
    const uint64_t *__restrict__ cand, int64_t ncand, int k, int negate, float pad_dist,
    float *__restrict__ D, int64_t *__restrict__ I64, int32_t *__restrict__ I32,
    uint64_t *__restrict__ out_keys, const int64_t *__restrict__ seg_off, int cap,
    const int32_t *__restrict__ kq, int kmul) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t *buf = reinterpret_cast<uint64_t *>(smem);  // cap
  uint64_t *top = buf + cap;                             // next_pow2(k)
  SelectScratch64 &scr = *reinterpret_cast<SelectScratch64 *>(top + next_pow2(k));
  const int tid = threadIdx.x;
  const int64_t qi = blockIdx.x;
  // Fixed-size candidate rows, or (seg_off) rows of seg_off[q+1]-seg_off[q]
  // units of ncand keys.
  const uint64_t *src = seg_off ? cand + seg_off[qi] * ncand : cand + qi * ncand;
  const int64_t n = seg_off ? (seg_off[qi + 1] - seg_off[qi]) * ncand : ncand;
  int cur = 0;
  const int room = cap - k;
  for (int64_t b = 0; b < n; b += room) {
    const int len = (int)min<int64_t>(room, n - b);
    for (int i = tid; i < cur; i += kNT) buf[i] = top[i];
    for (int i = tid; i < len; i += kNT) buf[cur + i] = src[b + i];
    __syncthreads();
    cur = select_sort_u64(buf, cur + len, k, scr, top);
  }
  if (kq) cur = min(cur, (int)min<int64_t>((int64_t)kq[qi] * kmul, k));
  for (int t = tid; t < k; t += kNT) {
    float dv = pad_dist;
    int64_t iv = -1;
    uint64_t kv = kSentinel64;
    if (t < cur) {
      kv = top[t];
      dv = key2f((uint32_t)(kv >> 32));
      if (negate) dv = -dv;
      iv = (int64_t)(uint32_t)(kv & 0xFFFFFFFFu);
    }
    const int64_t o = qi * k + t;
    if (D) D[o] = dv;
    if (I64) I64[o] = iv;
    if (I32) I32[o] = (int32_t)iv;
    if (out_keys) out_keys[o] = kv;
  }
}

// Results of queries whose candidate rows are already their sorted top k
// (sentinel padded): one thread per slot.
__global__ void sorted_topk_out_kernel(const uint64_t *__restrict__ cand, int64_t total, int negate,
                                       float pad_dist, float *__restrict__ D, int64_t *__restrict__ I64,
                                       int32_t *__restrict__ I32, uint64_t *__restrict__ out_keys) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= total) return;
  const uint64_t kv = cand[o];
  float dv = pad_dist;
  int64_t iv = -1;
  if (kv != kSentinel64) {
    dv = key2f((uint32_t)(kv >> 32));
    if (negate) dv = -dv;
    iv = (int64_t)(uint32_t)(kv & 0xFFFFFFFFu);
  }
  if (D) D[o] = dv;
  if (I64) I64[o] = iv;
  if (I32) I32[o] = (int32_t)iv;
  if (out_keys) out_keys[o] = kv;
}

void launch_merge(const uint64_t *cand, int64_t nq, int64_t ncand, int k, int negate,
                  float pad_dist, float *D, int64_t *I64, int32_t *I32, uint64_t *keys,
                  hipStream_t s, const int64_t *seg_off, const int32_t *kq, int kmul) {
  if (nq == 0) return;
  if (k > kMaxCandidates) throw Error(kInvalid, "k exceeds the device top-k limit of 4096");
  // The buffer holds the running top k plus the next candidates; a merge of
  // fixed-size rows smaller than kMergeCap - k sizes it to one pass, so small
  // merges keep several workgroups per CU.
  const int cap = seg_off ? kMergeCap
                          : (int)std::min<int64_t>(kMergeCap, std::max<int64_t>(ncand, 1) + k);
  const size_t sm = sizeof(uint64_t) * (cap + next_pow2(k)) + sizeof(SelectScratch64);
  static bool attr_set = false;
  if (!attr_set) {
    const size_t most = sizeof(uint64_t) * (kMergeCap + kMaxCandidates) + sizeof(SelectScratch64);
    MVG_HIP(hipFuncSetAttribute((const void *)merge_topk_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)most));
    attr_set = true;
  }
  hipLaunchKernelGGL(merge_topk_kernel, dim3((unsigned)nq), dim3(kNT), sm, s, cand, ncand, k,
                     negate, pad_dist, D, I64, I32, keys, seg_off, cap, kq, kmul);
  MVG_LAUNCH_CHECK("merge_topk_kernel");
}

static size_t chunk_smem(int d, int k) {
  return sizeof(float) * kTQ * d + sizeof(uint32_t) * kTQ * kCB + sizeof(uint32_t) * ((k + 3) & ~3) +
         sizeof(SelectScratch);
}

template <int D, bool IP>
static void launch_chunk(const float *base, int64_t n, int d, const float *q, int64_t nq, int k,
                         const uint8_t *excl, int64_t excl_len, uint64_t *part, int nchunks,
                         hipStream_t s) {
  dim3 grid(ceil_div(nq, kTQ), nchunks);
  size_t sm = chunk_smem(d, k);
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute((const void *)flat_chunk_topk_kernel<D, IP>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  hipLaunchKernelGGL((flat_chunk_topk_kernel<D, IP>), grid, dim3(kNT), sm, s, base, n, d, q, nq, k,
                     excl, excl_len, part, nchunks);
  MVG_LAUNCH_CHECK("flat_chunk_topk_kernel");
}

void flat_topk(const float *base, int64_t n, int d, int metric, const float *q, int64_t nq, int k,
               const uint8_t *excl, int64_t excl_len, DevBuf &part, float pad_dist, float *D,
               int64_t *I64, int32_t *I32, uint64_t *keys, hipStream_t s) {
  if (nq == 0) return;
  if (k > kMaxCandidates) throw Error(kInvalid, "k exceeds the device top-k limit of 4096");
  if (sizeof(float) * kTQ * d + sizeof(uint32_t) * kTQ * kCB + 16 * 1024 + sizeof(SelectScratch) >
      150 * 1024)
    throw Error(kInvalid, "dimension too large for the exact Flat kernel");
  const int nchunks = n > 0 ? ceil_div(n, kCB) : 0;
  if (nchunks > 0) {
    part.reserve(sizeof(uint64_t) * (size_t)nq * nchunks * k);
    uint64_t *pp = part.as<uint64_t>();
    const bool ip = metric == kMetricIP;
#define MVG_FLAT_CASE(DD)                                                                      \
  if (ip) launch_chunk<DD, true>(base, n, d, q, nq, k, excl, excl_len, pp, nchunks, s);       \
  else launch_chunk<DD, false>(base, n, d, q, nq, k, excl, excl_len, pp, nchunks, s);
    switch (d) {
      case 32: MVG_FLAT_CASE(32) break;
      case 64: MVG_FLAT_CASE(64) break;
      case 96: MVG_FLAT_CASE(96) break;
      case 128: MVG_FLAT_CASE(128) break;
      default: MVG_FLAT_CASE(0) break;
    }
#undef MVG_FLAT_CASE
  }
  if (nchunks == 1 && k <= 64) {
    // one chunk: the wavefront fold wrote each query's sorted top k already
    const int64_t total = nq * (int64_t)k;
    hipLaunchKernelGGL(sorted_topk_out_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s,
                       part.as<uint64_t>(), total, metric == kMetricIP ? 1 : 0, pad_dist, D, I64, I32, keys);
    MVG_LAUNCH_CHECK("sorted_topk_out_kernel");
    return;
  }
  launch_merge(nchunks > 0 ? part.as<uint64_t>() : nullptr, nq, (int64_t)nchunks * k, k,
               metric == kMetricIP ? 1 : 0, pad_dist, D, I64, I32, keys, s);
}

template <int D, bool IP, bool KM>
static void launch_argmin(const float *base, int64_t n, int d, const float *q, int64_t nq,
                          int32_t *out_id, float *out_dist, hipStream_t s) {
  size_t sm = sizeof(float) * kTQ * d + sizeof(uint64_t) * kTQ * kNW;
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute((const void *)flat_argmin_kernel<D, IP, KM>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  hipLaunchKernelGGL((flat_argmin_kernel<D, IP, KM>), dim3(ceil_div(nq, kTQ)), dim3(kNT), sm, s,
                     base, n, d, q, nq, out_id, out_dist);
  MVG_LAUNCH_CHECK("flat_argmin_kernel");
}

void flat_argmin(const float *base, int64_t n, int d, int metric, bool kmeans_mode, const float *q,
                 int64_t nq, int32_t *out_id, float *out_dist, hipStream_t s) {
  if (nq == 0) return;
  if (sizeof(float) * kTQ * d > 150 * 1024) throw Error(kInvalid, "dimension too large");
  const bool ip = metric == kMetricIP;
  // The d <= 4 k-means distance (short4) differs from l2_distance_simd[8]:
  // those sizes take the runtime-dimension path.
#define MVG_ARGMIN_CASE(DD)                                                              \
  if (ip) launch_argmin<DD, true, false>(base, n, d, q, nq, out_id, out_dist, s);        \
  else if (kmeans_mode) launch_argmin<DD, false, true>(base, n, d, q, nq, out_id, out_dist, s); \
  else launch_argmin<DD, false, false>(base, n, d, q, nq, out_id, out_dist, s);
  switch (d) {
    case 32: MVG_ARGMIN_CASE(32) break;
    case 64: MVG_ARGMIN_CASE(64) break;
    case 96: MVG_ARGMIN_CASE(96) break;
    case 128: MVG_ARGMIN_CASE(128) break;
    default: MVG_ARGMIN_CASE(0) break;
  }
#undef MVG_ARGMIN_CASE
}

}  // namespace mvg
