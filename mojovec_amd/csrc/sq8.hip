// SQ8 Flat: global-calibration 8-bit codes (mojovec/quantization/sq8.mojo,
// mojovec/index/index_flat_sq8.mojo), exact integer code dots, top-(k*k_factor)
// and the optional exact FP32 rerank.
//
//   L2 (affine UInt8): code = clamp(round((x - min) * inv), 0, 255);
//       dist = Float32(max(|q|^2 + |b|^2 - 2 q.b, 0)) * scale^2   (sq8.mojo:30-37, 77-79)
//   IP (symmetric Int8): code = clamp(round(x * inv), -127, 127);
//       dist = -Float32(q.b) * scale^2                           (sq8.mojo:22-29, 71-75)
// Code dots are exact in int32/uint32 (dot4 instructions accumulate exactly),
// so codes, norms and dots agree bit for bit with the CPU restatement.
#include "dist.h"
#include "kernels.h"
#include "select.h"

namespace mvg {

namespace {

constexpr int kTQ = 8;
constexpr int kCB = 1024;
constexpr int kSlice = 16;  // u32 words of a row held in registers per step

__global__ void minmax_kernel(const float *__restrict__ x, int64_t count, uint32_t *__restrict__ mm) {
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t k = f2key(x[i]);
    lo = min(lo, k);
    hi = max(hi, k);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, o, 64));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&mm[0], lo);
    atomicMax(&mm[1], hi);
  }
}

__global__ void init_minmax_kernel(uint32_t *mm) {
  mm[0] = 0xFFFFFFFFu;
  mm[1] = 0u;
}

// One wavefront per row: codes, the row's UInt32 code norm and (for the
// rebased int8 contraction) the UInt32 sum of the uint8 codes.
__global__ void sq8_encode_kernel(const float *__restrict__ x, int64_t n, int d, int ld,
                                  int symmetric, float minimum, float inv,
                                  uint8_t *__restrict__ codes, uint32_t *__restrict__ norms,
                                  uint32_t *__restrict__ sums) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  const float *xr = x + row * d;
  uint8_t *cr = codes + row * ld;
  uint32_t norm = 0, sum = 0;
  for (int j = lane; j < d; j += 64) {
    if (symmetric) {
      float s = roundf(__fmul_rn(xr[j], inv));
      s = fminf(fmaxf(s, -127.0f), 127.0f);
      const int8_t c = (int8_t)s;
      cr[j] = (uint8_t)c;
      norm += (uint32_t)((int32_t)c * (int32_t)c);
    } else {
      float s = roundf(__fmul_rn(__fsub_rn(xr[j], minimum), inv));
      s = fminf(fmaxf(s, 0.0f), 255.0f);
      const uint8_t c = (uint8_t)s;
      cr[j] = c;
      norm += (uint32_t)c * (uint32_t)c;
      sum += c;
    }
  }
  // zero padding up to the stride: contributes nothing to uint8 or int8 dots
  // (written here, not by a memset node, so a captured search replays it)
  for (int j = d + lane; j < ld; j += 64) cr[j] = 0;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    norm += (uint32_t)__shfl_xor((int)norm, o, 64);
    sum += (uint32_t)__shfl_xor((int)sum, o, 64);
  }
  if (lane == 0) {
    norms[row] = norm;
    if (sums) sums[row] = sum;
  }
}

// kTQ queries x one chunk of kCB rows: exact code dots, FP32 code distance,
// per-query top-kk of the chunk as u64 keys.
template <bool IP>
__global__ __launch_bounds__(kNT) void sq8_chunk_topk_kernel(
    const uint8_t *__restrict__ codes, const uint32_t *__restrict__ norms, int64_t n, int d,
    const uint8_t *__restrict__ qcodes, const uint32_t *__restrict__ qnorms, int64_t nq,
    float scale_sq, int kk, const uint8_t *__restrict__ excl, int64_t excl_len,
    uint64_t *__restrict__ part, int nchunks) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int dw = d / 4;  // d is the zero-padded code row stride (multiple of 64)
  uint32_t *qs = reinterpret_cast<uint32_t *>(smem);  // kTQ * dw
  uint32_t *keys = qs + kTQ * dw;                     // kTQ * kCB
  uint32_t *sel = keys + kTQ * kCB;                   // kk
  SelectScratch *scr = reinterpret_cast<SelectScratch *>(sel + ((kk + 3) & ~3));
  const int tid = threadIdx.x;
  const int64_t q0 = (int64_t)blockIdx.x * kTQ;
  const int chunk = blockIdx.y;
  const int64_t c0 = (int64_t)chunk * kCB;
  const int qcount = (int)min<int64_t>(kTQ, nq - q0);
  const uint32_t *qsrc = reinterpret_cast<const uint32_t *>(qcodes);
  for (int i = tid; i < kTQ * dw; i += kNT) {
    const int qi = i / dw;
    qs[i] = qi < qcount ? qsrc[(q0 + qi) * dw + (i - qi * dw)] : 0u;
  }
  __syncthreads();
  for (int r = 0; r < kCB / kNT; r++) {
    const int local = r * kNT + tid;
    const int64_t j = c0 + local;
    const bool live = j < n && !(excl && j < excl_len && excl[j] > 0);
    int32_t acc[kTQ];
#pragma unroll
    for (int qi = 0; qi < kTQ; qi++) acc[qi] = 0;
    if (live) {
      const uint32_t *row = reinterpret_cast<const uint32_t *>(codes + j * d);
      for (int w0 = 0; w0 < dw; w0 += kSlice) {
        uint32_t b[kSlice];
        const int cnt = min(kSlice, dw - w0);
#pragma unroll
        for (int t = 0; t < kSlice; t++) b[t] = t < cnt ? row[w0 + t] : 0u;
#pragma unroll
        for (int qi = 0; qi < kTQ; qi++) {
          int32_t a = acc[qi];
#pragma unroll
          for (int t = 0; t < kSlice; t++) {
            const uint32_t qv = t < cnt ? qs[qi * dw + w0 + t] : 0u;
            if (IP)
              a = __builtin_amdgcn_sdot4((int)qv, (int)b[t], a, false);
            else
              a = (int32_t)__builtin_amdgcn_udot4(qv, b[t], (uint32_t)a, false);
          }
          acc[qi] = a;
        }
      }
    }
    for (int qi = 0; qi < kTQ; qi++) {
      uint32_t key = kSentinel;
      if (live && qi < qcount) {
        float dist;
        if (IP) {
          dist = __fmul_rn(-(float)acc[qi], scale_sq);
        } else {
          const int64_t v = (int64_t)qnorms[q0 + qi] + (int64_t)norms[j] - 2 * (int64_t)(uint32_t)acc[qi];
          dist = __fmul_rn((float)(uint32_t)(v > 0 ? v : 0), scale_sq);  // sq8.mojo:77-79
        }
        key = f2key(dist);
      }
      keys[qi * kCB + local] = key;
    }
  }
  __syncthreads();
  const int clen = (int)min<int64_t>(kCB, n - c0);
  for (int qi = 0; qi < qcount; qi++) {
    const int cnt = select_topk_u32(keys + qi * kCB, clen, kk, *scr, sel);
    uint64_t *dst = part + ((q0 + qi) * nchunks + chunk) * (int64_t)kk;
    for (int t = tid; t < kk; t += kNT) {
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

}  // namespace

void sq8_minmax(const float *x, int64_t count, float *out2, DevBuf &tmp, hipStream_t s) {
  tmp.reserve(2 * sizeof(uint32_t));
  uint32_t *mm = tmp.as<uint32_t>();
  hipLaunchKernelGGL(init_minmax_kernel, dim3(1), dim3(1), 0, s, mm);
  const int blocks = (int)std::min<int64_t>(4096, std::max<int64_t>(1, ceil_div(count, 256)));
  hipLaunchKernelGGL(minmax_kernel, dim3(blocks), dim3(256), 0, s, x, count, mm);
  MVG_LAUNCH_CHECK("minmax_kernel");
  uint32_t h[2];
  MVG_HIP(hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, s));
  MVG_HIP(hipStreamSynchronize(s));
  out2[0] = key2f(h[0]);
  out2[1] = key2f(h[1]);
}

int sq8_row_stride(int d) { return (d + 63) & ~63; }

void sq8_encode(const float *x, int64_t n, int d, int symmetric, float minimum, float inv_scale,
                uint8_t *codes, uint32_t *norms, uint32_t *sums, hipStream_t s) {
  if (n == 0) return;
  const int ld = sq8_row_stride(d);
  hipLaunchKernelGGL(sq8_encode_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, s, x, n, d, ld,
                     symmetric, minimum, inv_scale, codes, norms, sums);
  MVG_LAUNCH_CHECK("sq8_encode_kernel");
}

void sq8_search(const Sq8SearchArgs &a, Sq8Work &w, hipStream_t s) {
  if (a.nq == 0) return;
  const int kf = a.k_factor > 1 ? a.k_factor : 1;
  const int64_t kk64 = (int64_t)a.k * kf;
  if (kk64 > kMaxCandidates) throw Error(kInvalid, "k * k_factor exceeds the device limit of 4096");
  const int kk = (int)kk64;
  const bool ip = a.metric == kMetricIP;
  const float inv = 1.0f / a.scale;
  const int ld = sq8_row_stride(a.d);
  w.qcodes.reserve((size_t)a.nq * ld);
  w.qnorms.reserve(sizeof(uint32_t) * a.nq);
  w.qsums.reserve(sizeof(uint32_t) * a.nq);
  sq8_encode(a.q, a.nq, a.d, ip ? 1 : 0, ip ? 0.0f : a.gmin, inv, w.qcodes.as<uint8_t>(),
             w.qnorms.as<uint32_t>(), w.qsums.as<uint32_t>(), s);
  const float ssq = a.scale * a.scale;
  if (a.use_mfma && a.n > 0) {
    int64_t ncand = 0;
    gemm_topk(ip ? kGemmI8IP : kGemmU8L2, w.qcodes.as<uint8_t>(), a.codes, a.nq, a.n, ld, nullptr,
              nullptr, w.qnorms.as<uint32_t>(), w.qsums.as<uint32_t>(), a.norms, a.sums, ssq, kk,
              a.excl, a.excl_len, w.part, ncand, s);
    if (kf == 1) {
      launch_merge(w.part.as<uint64_t>(), a.nq, ncand, a.k, ip ? 1 : 0, 0.0f, a.D, a.I,
                   nullptr, nullptr, s, nullptr, a.kq, 1);
    } else {
      w.keys.reserve(sizeof(uint64_t) * (size_t)a.nq * kk);
      launch_merge(w.part.as<uint64_t>(), a.nq, ncand, kk, 0, 0.0f, nullptr, nullptr,
                   nullptr, w.keys.as<uint64_t>(), s, nullptr, a.kq, kf);
      rerank(a.q, a.nq, a.d, a.metric, a.vectors, a.n, w.keys.as<uint64_t>(), kk, a.k, 0.0f, a.D, a.I, s,
             nullptr, RowMap{}, a.kq);
    }
    return;
  }
  const int nchunks = a.n > 0 ? ceil_div(a.n, kCB) : 0;
  if (nchunks > 0) {
    w.part.reserve(sizeof(uint64_t) * (size_t)a.nq * nchunks * kk);
    const size_t sm = sizeof(uint32_t) * kTQ * (ld / 4) + sizeof(uint32_t) * kTQ * kCB +
                      sizeof(uint32_t) * ((kk + 3) & ~3) + sizeof(SelectScratch);
    if (sm > 150 * 1024) throw Error(kInvalid, "SQ8 scan workspace exceeds LDS");
    const void *fn = ip ? (const void *)sq8_chunk_topk_kernel<true> : (const void *)sq8_chunk_topk_kernel<false>;
    if (sm > 64 * 1024)
      MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    dim3 grid(ceil_div(a.nq, kTQ), nchunks);
    if (ip)
      hipLaunchKernelGGL(sq8_chunk_topk_kernel<true>, grid, dim3(kNT), sm, s, a.codes, a.norms, a.n,
                         ld, w.qcodes.as<uint8_t>(), w.qnorms.as<uint32_t>(), a.nq, ssq, kk, a.excl,
                         a.excl_len, w.part.as<uint64_t>(), nchunks);
    else
      hipLaunchKernelGGL(sq8_chunk_topk_kernel<false>, grid, dim3(kNT), sm, s, a.codes, a.norms, a.n,
                         ld, w.qcodes.as<uint8_t>(), w.qnorms.as<uint32_t>(), a.nq, ssq, kk, a.excl,
                         a.excl_len, w.part.as<uint64_t>(), nchunks);
    MVG_LAUNCH_CHECK("sq8_chunk_topk_kernel");
  }
  const uint64_t *part = nchunks > 0 ? w.part.as<uint64_t>() : nullptr;
  if (kf == 1) {
    launch_merge(part, a.nq, (int64_t)nchunks * kk, a.k, ip ? 1 : 0, 0.0f, a.D, a.I, nullptr,
                 nullptr, s, nullptr, a.kq, 1);
  } else {
    w.keys.reserve(sizeof(uint64_t) * (size_t)a.nq * kk);
    launch_merge(part, a.nq, (int64_t)nchunks * kk, kk, 0, 0.0f, nullptr, nullptr, nullptr,
                 w.keys.as<uint64_t>(), s, nullptr, a.kq, kf);
    rerank(a.q, a.nq, a.d, a.metric, a.vectors, a.n, w.keys.as<uint64_t>(), kk, a.k, 0.0f, a.D, a.I, s,
           nullptr, RowMap{}, a.kq);
  }
}

}  // namespace mvg
