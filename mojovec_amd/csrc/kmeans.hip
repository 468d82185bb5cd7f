// Lloyd k-means (coarse quantizer and PQ codebooks) and PQ encode / LUT.
//
// Restates KMeans._train (mojovec/clustering/kmeans.mojo:60-220): the data is
// cut into 32 fixed chunks of ceil(n/32) rows; each chunk accumulates its rows
// in row order, chunks are then reduced in chunk order, and a centroid is
// sum * (1/count) in FP32; an empty cluster is reseeded from row
// (sample_offset + (iter+1)*k + c) % n. The accumulation is FP64 here (the
// reference uses FP32), rounded to FP32 once before the scale. One thread owns
// one (chunk, column) accumulator stream, so the sums are deterministic and
// equal to the CPU restatement's.
#include "dist.h"
#include "kernels.h"

namespace mvg {

namespace {

constexpr int kChunks = 32;  // kmeans.mojo:84

__global__ void kmeans_chunk_sums_kernel(const float *__restrict__ x, int64_t n, int S, int dsub,
                                         int k, const int32_t *__restrict__ assign,
                                         double *__restrict__ sums, int64_t *__restrict__ counts,
                                         int64_t chunk_size) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)S * kChunks * dsub;
  if (t >= total) return;
  const int j = (int)(t % dsub);
  const int ch = (int)((t / dsub) % kChunks);
  const int s = (int)(t / ((int64_t)dsub * kChunks));
  const int ld = S * dsub;
  const int64_t start = (int64_t)ch * chunk_size;
  const int64_t end = min(start + chunk_size, n);
  double *sm = sums + ((int64_t)(s * kChunks + ch) * k) * dsub;
  int64_t *cm = counts + (int64_t)(s * kChunks + ch) * k;
  for (int64_t i = start; i < end; i++) {
    const int c = assign[i * S + s];
    sm[(int64_t)c * dsub + j] += (double)x[i * ld + (int64_t)s * dsub + j];
    if (j == 0) cm[c] += 1;
  }
}

__global__ void kmeans_finalize_kernel(const float *__restrict__ x, int64_t n, int S, int dsub,
                                       int k, const double *__restrict__ sums,
                                       const int64_t *__restrict__ counts, int iter,
                                       int64_t sample_offset, float *__restrict__ cent) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)S * k * dsub;
  if (t >= total) return;
  const int j = (int)(t % dsub);
  const int c = (int)((t / dsub) % k);
  const int s = (int)(t / ((int64_t)dsub * k));
  int64_t count = 0;
  for (int ch = 0; ch < kChunks; ch++) count += counts[(int64_t)(s * kChunks + ch) * k + c];
  float v;
  if (count > 0) {
    double acc = 0.0;
    for (int ch = 0; ch < kChunks; ch++)
      acc += sums[((int64_t)(s * kChunks + ch) * k + c) * dsub + j];
    const float inv = __fdiv_rn(1.0f, (float)count);  // kmeans.mojo:201
    v = __fmul_rn((float)acc, inv);
  } else {
    const int64_t src = (sample_offset + (int64_t)(iter + 1) * k + c) % n;  // :214-216
    v = x[src * (int64_t)(S * dsub) + (int64_t)s * dsub + j];
  }
  cent[((int64_t)s * k + c) * dsub + j] = v;
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
  const int64_t chunk_size = (n + kChunks - 1) / kChunks;
  const size_t nsum = (size_t)S * kChunks * k * dsub;
  const size_t ncnt = (size_t)S * kChunks * k;
  w.sums.reserve(sizeof(double) * nsum);
  w.counts.reserve(sizeof(int64_t) * ncnt);
  w.assign.reserve(sizeof(int32_t) * (size_t)n * S);
  int32_t *assign = w.assign.as<int32_t>();
  for (int it = 0; it < niter; it++) {
    if (S == 1) {
      flat_argmin(cent, k, dsub, kMetricL2, true, x, n, assign, nullptr, s);
    } else {
      if (k != kKsub) throw Error(kInvalid, "PQ training uses 256 centroids per subspace");
      pq_assign(x, n, S * dsub, S, cent, nullptr, nullptr, nullptr, assign, s);
    }
    MVG_HIP(hipMemsetAsync(w.sums.ptr, 0, sizeof(double) * nsum, s));
    MVG_HIP(hipMemsetAsync(w.counts.ptr, 0, sizeof(int64_t) * ncnt, s));
    const int64_t nt1 = (int64_t)S * kChunks * dsub;
    hipLaunchKernelGGL(kmeans_chunk_sums_kernel, dim3(ceil_div(nt1, 64)), dim3(64), 0, s, x, n, S,
                       dsub, k, assign, w.sums.as<double>(), w.counts.as<int64_t>(), chunk_size);
    MVG_LAUNCH_CHECK("kmeans_chunk_sums_kernel");
    const int64_t nt2 = (int64_t)S * k * dsub;
    hipLaunchKernelGGL(kmeans_finalize_kernel, dim3(ceil_div(nt2, 256)), dim3(256), 0, s, x, n, S,
                       dsub, k, w.sums.as<double>(), w.counts.as<int64_t>(), it, sample_offset, cent);
    MVG_LAUNCH_CHECK("kmeans_finalize_kernel");
  }
}

}  // namespace mvg
