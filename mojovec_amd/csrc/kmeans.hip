// Lloyd k-means (coarse quantizer and PQ codebooks) and PQ encode / LUT.
//
// Restates KMeans._train (mojovec/clustering/kmeans.mojo:60-220): the data is
// cut into 32 fixed chunks of ceil(n/32) rows; each chunk accumulates its rows
// in row order, chunks are then reduced in chunk order, and a centroid is
// sum * (1/count) in FP32; an empty cluster is reseeded from row
// (sample_offset + (iter+1)*k + c) % n. The accumulation is FP64 here (the
// reference uses FP32), rounded to FP32 once before the scale.
//
// The sums keep exactly that order while spreading the work over every
// (centroid, column): the rows are stably sorted by their assigned centroid
// (LSD radix sort: within one centroid the rows stay in row order), then one
// thread per (subspace, centroid, column) walks its centroid's rows in row
// order, summing each chunk's rows from 0.0 and adding the chunk partials to
// the total in chunk order -- the same FP64 additions, in the same order, as
// one accumulator per (chunk, centroid, column) reduced over the chunks.
#include <rocprim/device/device_radix_sort.hpp>

#include "dist.h"
#include "kernels.h"

namespace mvg {

namespace {

constexpr int kChunks = 32;  // kmeans.mojo:84

// Sort keys of (row, subspace): subspace-major centroid index, and the row.
__global__ void kmeans_keys_kernel(const int32_t *__restrict__ assign, int64_t n, int S, int k,
                                   uint32_t *__restrict__ keys, int32_t *__restrict__ rows) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * S) return;
  const int64_t i = t / S;
  const int s = (int)(t - i * S);
  keys[t] = (uint32_t)s * (uint32_t)k + (uint32_t)assign[t];
  rows[t] = (int32_t)i;
}

// seg[key] = first sorted position with that key or a larger one (key in
// [0, nkeys]; seg[nkeys] = total).
__global__ void kmeans_segments_kernel(const uint32_t *__restrict__ sorted, int64_t total,
                                       int64_t nkeys, int64_t *__restrict__ seg) {
  const int64_t key = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (key > nkeys) return;
  int64_t lo = 0, hi = total;  // lower bound of key
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)sorted[mid] < key) lo = mid + 1; else hi = mid;
  }
  seg[key] = lo;
}

// One thread per (subspace, centroid, column): FP64 sum of the centroid's
// rows in row order, restarted at each 32-row-chunk boundary and added to the
// total in chunk order (kmeans.mojo:149-195), then sum * (1/count) in FP32
// (:197-210) or the reseed row of an empty cluster (:211-220).
__global__ void kmeans_update_kernel(const float *__restrict__ x, int64_t n, int S, int dsub, int k,
                                     const int32_t *__restrict__ rows,
                                     const int64_t *__restrict__ seg, int64_t chunk_size, int iter,
                                     int64_t sample_offset, float *__restrict__ cent) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)S * k * dsub;
  if (t >= total) return;
  const int j = (int)(t % dsub);
  const int64_t key = t / dsub;  // s * k + c
  const int s = (int)(key / k);
  const int64_t ld = (int64_t)S * dsub;
  const int64_t lo = seg[key], hi = seg[key + 1];
  float v;
  if (hi > lo) {
    double acc = 0.0, part = 0.0;
    int64_t cur = -1;
    for (int64_t p = lo; p < hi; p++) {
      const int64_t i = rows[p];
      const int64_t ch = i / chunk_size;
      if (ch != cur) {
        acc += part;
        part = 0.0;
        cur = ch;
      }
      part += (double)x[i * ld + (int64_t)s * dsub + j];
    }
    acc += part;
    const float inv = __fdiv_rn(1.0f, (float)(hi - lo));  // kmeans.mojo:201
    v = __fmul_rn((float)acc, inv);
  } else {
    const int c = (int)(key - (int64_t)s * k);
    const int64_t src = (sample_offset + (int64_t)(iter + 1) * k + c) % n;  // :214-216
    v = x[src * ld + (int64_t)s * dsub + j];
  }
  cent[key * dsub + j] = v;
}

// codes / assignment of (row, subspace m) = argmin over 256 centroids with
// the reference scan rule (start 1e38, strict <; pq.mojo:197-219).
template <int DSUB>
__global__ __launch_bounds__(256) void pq_assign_kernel(
    const float *__restrict__ x, int64_t n, int d, int M, int dsub_rt, const float *__restrict__ cb,
    const float *__restrict__ coarse, const int32_t *__restrict__ labels,
    uint8_t *__restrict__ codes, int32_t *__restrict__ assign) {
  extern __shared__ __align__(16) float cbs[];  // 256 * dsub
  const int dsub = DSUB > 0 ? DSUB : dsub_rt;
  const int m = blockIdx.y;
  for (int e = threadIdx.x; e < kKsub * dsub; e += blockDim.x)
    cbs[e] = cb[(int64_t)m * kKsub * dsub + e];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sub[DSUB > 0 ? DSUB : 64];
  const float *xs = x + i * d + (int64_t)m * dsub;
  if (coarse) {
    const float *cs = coarse + (int64_t)labels[i] * d + (int64_t)m * dsub;
    for (int j = 0; j < dsub; j++) sub[j] = xs[j] - cs[j];
  } else {
    for (int j = 0; j < dsub; j++) sub[j] = xs[j];
  }
  float best = 1e38f;
  int bk = -1;
  for (int kk = 0; kk < kKsub; kk++) {
    const float dist = l2_kmeans(sub, cbs + kk * dsub, dsub);
    if (dist < best) {
      best = dist;
      bk = kk;
    }
  }
  if (codes) codes[i * M + m] = (uint8_t)bk;       // UInt8(-1) == 255
  if (assign) assign[i * M + m] = bk > 0 ? bk : 0;  // kmeans.mojo:147
}

__global__ void pq_lut_kernel(const float *__restrict__ q, int d, int M, int dsub,
                              const float *__restrict__ cb, int metric, float *__restrict__ lut) {
  const int64_t qi = blockIdx.x;
  const float *qq = q + qi * d;
  for (int e = threadIdx.x; e < M * kKsub; e += blockDim.x) {
    const int m = e / kKsub;
    const float *c = cb + (int64_t)e * dsub;
    const float *sq = qq + (int64_t)m * dsub;
    lut[qi * M * kKsub + e] = metric == kMetricL2 ? l2_kmeans(sq, c, dsub) : -ip_pq(sq, c, dsub);
  }
}

// ProductQuantizer.decode (pq.mojo:221-250): row i, column m * dsub + j is
// codebook entry (m, codes[i][m]), component j. One thread per output float.
__global__ void pq_decode_kernel(const uint8_t *__restrict__ codes, int64_t n, int d, int M,
                                 const float *__restrict__ cb, float *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const int dsub = d / M;
  const int64_t i = t / d;
  const int c = (int)(t - i * d);
  const int m = c / dsub, j = c - m * dsub;
  out[t] = cb[((int64_t)m * kKsub + codes[i * M + m]) * dsub + j];
}

// Rows ids[i] of base [.][d] gathered into out [n][d].
__global__ void gather_rows_kernel(const float *__restrict__ base, int d,
                                   const int64_t *__restrict__ ids, int64_t n,
                                   float *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const int64_t i = t / d;
  out[t] = base[ids[i] * d + (t - i * d)];
}

__global__ void residual_kernel(const float *__restrict__ x, int64_t n, int d,
                                const float *__restrict__ coarse, const int32_t *__restrict__ labels,
                                float *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const int64_t i = t / d;
  const int j = (int)(t - i * d);
  out[t] = x[t] - coarse[(int64_t)labels[i] * d + j];
}

}  // namespace

void pq_assign(const float *x, int64_t n, int d, int M, const float *cb, const float *coarse,
               const int32_t *labels, uint8_t *codes, int32_t *assign, hipStream_t s) {
  if (n == 0) return;
  const int dsub = d / M;
  if (dsub > 64) throw Error(kInvalid, "PQ subvector dimension above 64 is not supported");
  dim3 grid(ceil_div(n, 256), M);
  const size_t sm = sizeof(float) * kKsub * dsub;
  if (sm > 64 * 1024) {
    const void *fn = dsub == 8 ? (const void *)pq_assign_kernel<8> : (const void *)pq_assign_kernel<0>;
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  }
  switch (dsub) {
    case 4: hipLaunchKernelGGL(pq_assign_kernel<4>, grid, dim3(256), sm, s, x, n, d, M, dsub, cb, coarse, labels, codes, assign); break;
    case 8: hipLaunchKernelGGL(pq_assign_kernel<8>, grid, dim3(256), sm, s, x, n, d, M, dsub, cb, coarse, labels, codes, assign); break;
    case 16: hipLaunchKernelGGL(pq_assign_kernel<16>, grid, dim3(256), sm, s, x, n, d, M, dsub, cb, coarse, labels, codes, assign); break;
    default: hipLaunchKernelGGL(pq_assign_kernel<0>, grid, dim3(256), sm, s, x, n, d, M, dsub, cb, coarse, labels, codes, assign); break;
  }
  MVG_LAUNCH_CHECK("pq_assign_kernel");
}

void pq_lut(const float *q, int64_t nq, int d, int M, const float *cb, int metric, float *lut,
            hipStream_t s) {
  if (nq == 0) return;
  hipLaunchKernelGGL(pq_lut_kernel, dim3((unsigned)nq), dim3(256), 0, s, q, d, M, d / M, cb, metric, lut);
  MVG_LAUNCH_CHECK("pq_lut_kernel");
}

void pq_decode(const uint8_t *codes, int64_t n, int d, int M, const float *cb, float *out,
               hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(pq_decode_kernel, dim3(ceil_div(n * d, 256)), dim3(256), 0, s, codes, n, d, M, cb,
                     out);
  MVG_LAUNCH_CHECK("pq_decode_kernel");
}

void gather_rows(const float *base, int d, const int64_t *ids, int64_t n, float *out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(ceil_div(n * d, 256)), dim3(256), 0, s, base, d, ids, n,
                     out);
  MVG_LAUNCH_CHECK("gather_rows_kernel");
}

void residuals(const float *x, int64_t n, int d, const float *coarse, const int32_t *labels,
               float *out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(residual_kernel, dim3(ceil_div(n * d, 256)), dim3(256), 0, s, x, n, d, coarse,
                     labels, out);
  MVG_LAUNCH_CHECK("residual_kernel");
}

void kmeans_fit(const float *x, int64_t n, int S, int dsub, int k, int niter,
                int64_t sample_offset, float *cent, KmeansWork &w, hipStream_t s) {
  if (n == 0 || niter <= 0) return;
  if (n * S > INT32_MAX) throw Error(kInvalid, "k-means holds at most 2^31-1 (row, subspace) pairs");
  const int64_t chunk_size = (n + kChunks - 1) / kChunks;
  const int64_t total = n * S, nkeys = (int64_t)S * k;
  unsigned end_bit = 1;
  while (end_bit < 32 && ((int64_t)1 << end_bit) < nkeys) end_bit++;
  w.assign.reserve(sizeof(int32_t) * (size_t)total);
  w.keys.reserve(sizeof(uint32_t) * (size_t)total);
  w.rows.reserve(sizeof(int32_t) * (size_t)total);
  w.keys_sorted.reserve(sizeof(uint32_t) * (size_t)total);
  w.rows_sorted.reserve(sizeof(int32_t) * (size_t)total);
  w.seg.reserve(sizeof(int64_t) * (size_t)(nkeys + 1));
  size_t tmp_bytes = 0;
  MVG_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, w.keys.as<uint32_t>(), w.keys_sorted.as<uint32_t>(),
                                    w.rows.as<int32_t>(), w.rows_sorted.as<int32_t>(), (size_t)total, 0,
                                    end_bit, s));
  w.sort_tmp.reserve(std::max<size_t>(tmp_bytes, 1));
  int32_t *assign = w.assign.as<int32_t>();
  for (int it = 0; it < niter; it++) {
    if (S == 1) {
      flat_argmin(cent, k, dsub, kMetricL2, true, x, n, assign, nullptr, s);
    } else {
      if (k != kKsub) throw Error(kInvalid, "PQ training uses 256 centroids per subspace");
      pq_assign(x, n, S * dsub, S, cent, nullptr, nullptr, nullptr, assign, s);
    }
    hipLaunchKernelGGL(kmeans_keys_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, s, assign, n, S, k,
                       w.keys.as<uint32_t>(), w.rows.as<int32_t>());
    MVG_LAUNCH_CHECK("kmeans_keys_kernel");
    size_t tb = w.sort_tmp.bytes;
    MVG_HIP(rocprim::radix_sort_pairs(w.sort_tmp.ptr, tb, w.keys.as<uint32_t>(), w.keys_sorted.as<uint32_t>(),
                                      w.rows.as<int32_t>(), w.rows_sorted.as<int32_t>(), (size_t)total, 0,
                                      end_bit, s));
    hipLaunchKernelGGL(kmeans_segments_kernel, dim3(ceil_div(nkeys + 1, 256)), dim3(256), 0, s,
                       w.keys_sorted.as<uint32_t>(), total, nkeys, w.seg.as<int64_t>());
    MVG_LAUNCH_CHECK("kmeans_segments_kernel");
    const int64_t nt = nkeys * dsub;
    hipLaunchKernelGGL(kmeans_update_kernel, dim3(ceil_div(nt, 256)), dim3(256), 0, s, x, n, S, dsub, k,
                       w.rows_sorted.as<int32_t>(), w.seg.as<int64_t>(), chunk_size, it, sample_offset,
                       cent);
    MVG_LAUNCH_CHECK("kmeans_update_kernel");
  }
}

}  // namespace mvg
