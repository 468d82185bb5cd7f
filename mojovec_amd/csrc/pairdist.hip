// Batched (query, row) distances for a CPU graph search that offloads its
// candidate evaluations: the distance computers of the Flat and SQ8 storages
// (FlatDistanceComputer, mojovec/index/index_flat.mojo:36-115;
// SQ8DistanceComputer, mojovec/index/index_flat_sq8.mojo:21-190), evaluated
// for many pairs at once.
//
// FP32 rows: a group of 8 lanes computes one pair, lane j holding SIMD lane j
// of the reference's 8-wide accumulators, so each step reads 32 contiguous
// bytes of the row and the result is bit-identical to the CPU routine:
//   order 0, distance():        l2_distance_simd[8] / inner_product_simd[8]
//     (distances.mojo:43-228): four accumulators over 32-element blocks,
//     ((a0 + a1) + a2) + a3, 8-element blocks, the halving tree, one 4-element
//     block, a scalar tail; d < 8 is scalar throughout;
//   order 1, distance_batch4(): l2_distance_simd_batch4[8] /
//     inner_product_simd_batch4[8] (distances.mojo:541-660): one accumulator
//     over 8-element blocks, the halving tree, a scalar tail.
// SQ8 codes: a group of 16 lanes forms the exact integer code dot with dot4
// instructions (any order is exact) and one lane applies the code distance
// (sq8.mojo:62-79).
#include "dist.h"
#include "kernels.h"

namespace mvg {

namespace {

// Halving tree over the 8 lanes of a group (v[j] + v[j + 4], then + 2, + 1);
// every lane ends with the result.
__device__ __forceinline__ float group_tree8(float v) {
  v = v + __shfl_xor(v, 4, 8);
  v = v + __shfl_xor(v, 2, 8);
  v = v + __shfl_xor(v, 1, 8);
  return v;
}

template <bool IP>
__device__ __forceinline__ float term(float a, float b) {
  if (IP) return a * b;
  const float df = a - b;
  return df * df;
}

template <bool IP>
__device__ __forceinline__ float fma_term(float a, float b, float acc) {
  if (IP) return fma_(a, b, acc);
  const float df = a - b;
  return fma_(df, df, acc);
}

// distance(): l2_distance_simd[8] / inner_product_simd[8] by lane j of a group.
template <bool IP>
__device__ float group_single8(const float *__restrict__ x, const float *__restrict__ y, int d, int j) {
  if (d < 8) {  // scalar path (distances.mojo:57-62)
    float r = 0.0f;
    for (int i = 0; i < d; i++) r = r + term<IP>(x[i], y[i]);
    return r;
  }
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  int i = 0;
  for (; i <= d - 32; i += 32) {
    a0 = fma_term<IP>(x[i + j], y[i + j], a0);
    a1 = fma_term<IP>(x[i + 8 + j], y[i + 8 + j], a1);
    a2 = fma_term<IP>(x[i + 16 + j], y[i + 16 + j], a2);
    a3 = fma_term<IP>(x[i + 24 + j], y[i + 24 + j], a3);
  }
  float v = ((a0 + a1) + a2) + a3;
  for (; i <= d - 8; i += 8) v = fma_term<IP>(x[i + j], y[i + j], v);
  float res = group_tree8(v);
  if (i <= d - 4) {  // one 4-element block: products, then the 4-lane tree
    float r[4];
#pragma unroll
    for (int t = 0; t < 4; t++) r[t] = IP ? fma_(x[i + t], y[i + t], 0.0f) : term<false>(x[i + t], y[i + t]);
    res += (r[0] + r[2]) + (r[1] + r[3]);
    i += 4;
  }
  for (; i < d; i++) res += term<IP>(x[i], y[i]);
  return res;
}

// distance_batch4(): l2_distance_simd_batch4[8] / inner_product_simd_batch4[8].
template <bool IP>
__device__ float group_batch8(const float *__restrict__ x, const float *__restrict__ y, int d, int j) {
  float acc = 0.0f;
  int i = 0;
  // one accumulator chain; four blocks' loads are issued ahead of their fmas
  for (; i + 32 <= d; i += 32) {
    float xv[4], yv[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      xv[b] = x[i + 8 * b + j];
      yv[b] = y[i + 8 * b + j];
    }
#pragma unroll
    for (int b = 0; b < 4; b++) acc = fma_term<IP>(xv[b], yv[b], acc);
  }
  for (; i + 8 <= d; i += 8) acc = fma_term<IP>(x[i + j], y[i + j], acc);
  float res = group_tree8(acc);
  for (; i < d; i++) res += term<IP>(x[i], y[i]);
  return res;
}

// Pair p: query row qid[p] (or p / per_query when qid is null) of qbase
// against row xid[p] of xbase. 8 lanes per pair; blockDim is a multiple of 8,
// so a group never straddles the end of the pairs.
template <bool IP, int ORDER>
__global__ __launch_bounds__(256) void flat_pair_dist_kernel(
    const float *__restrict__ qbase, const int64_t *__restrict__ qid, int64_t per_query,
    const float *__restrict__ xbase, const int64_t *__restrict__ xid, int64_t npairs, int d,
    float *__restrict__ out) {
  const int64_t p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
  const int j = threadIdx.x & 7;
  if (p >= npairs) return;
  const int64_t qi = qid ? qid[p] : p / per_query;
  const float *x = qbase + qi * d;
  const float *y = xbase + xid[p] * d;
  const float r = ORDER == 0 ? group_single8<IP>(x, y, d, j) : group_batch8<IP>(x, y, d, j);
  if (j == 0) out[p] = IP ? -r : r;
}

// SQ8: 16 lanes per pair over the zero-padded code rows (ld a multiple of 64).
template <bool IP>
__global__ __launch_bounds__(256) void sq8_pair_dist_kernel(
    const uint8_t *__restrict__ qcodes, const uint32_t *__restrict__ qnorms,
    const int64_t *__restrict__ qid, int64_t per_query, const uint8_t *__restrict__ codes,
    const uint32_t *__restrict__ norms, const int64_t *__restrict__ xid, int64_t npairs, int ld,
    float scale_sq, float *__restrict__ out) {
  const int64_t p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int j = threadIdx.x & 15;
  if (p >= npairs) return;
  const int64_t qi = qid ? qid[p] : p / per_query;
  const int64_t xi = xid[p];
  const uint4 *qa = reinterpret_cast<const uint4 *>(qcodes + qi * ld);
  const uint4 *xa = reinterpret_cast<const uint4 *>(codes + xi * ld);
  uint32_t acc = 0;
  for (int w = j; w < ld / 16; w += 16) {  // 16-byte chunks: 256 bytes per group step
    const uint4 a = qa[w], b = xa[w];
    if (IP) {
      int t = (int)acc;
      t = __builtin_amdgcn_sdot4((int)a.x, (int)b.x, t, false);
      t = __builtin_amdgcn_sdot4((int)a.y, (int)b.y, t, false);
      t = __builtin_amdgcn_sdot4((int)a.z, (int)b.z, t, false);
      t = __builtin_amdgcn_sdot4((int)a.w, (int)b.w, t, false);
      acc = (uint32_t)t;
    } else {
      acc = __builtin_amdgcn_udot4(a.x, b.x, acc, false);
      acc = __builtin_amdgcn_udot4(a.y, b.y, acc, false);
      acc = __builtin_amdgcn_udot4(a.z, b.z, acc, false);
      acc = __builtin_amdgcn_udot4(a.w, b.w, acc, false);
    }
  }
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) acc += (uint32_t)__shfl_xor((int)acc, o, 16);
  if (j != 0) return;
  float dist;
  if (IP) {
    dist = __fmul_rn(-(float)(int32_t)acc, scale_sq);
  } else {
    const int64_t v = (int64_t)qnorms[qi] + (int64_t)norms[xi] - 2 * (int64_t)acc;
    dist = __fmul_rn((float)(uint32_t)(v > 0 ? v : 0), scale_sq);
  }
  out[p] = dist;
}

}  // namespace

void flat_pair_distances(const float *qbase, const int64_t *qid, int64_t per_query,
                         const float *xbase, const int64_t *xid, int64_t npairs, int d, int metric,
                         int order, float *out, hipStream_t s) {
  if (npairs == 0) return;
  const dim3 grid((unsigned)ceil_div(npairs * 8, 256));
  const bool ip = metric == kMetricIP;
#define MVG_PD(IP, O)                                                                             \
  hipLaunchKernelGGL((flat_pair_dist_kernel<IP, O>), grid, dim3(256), 0, s, qbase, qid, per_query, \
                     xbase, xid, npairs, d, out)
  if (ip) {
    if (order == 0) MVG_PD(true, 0); else MVG_PD(true, 1);
  } else {
    if (order == 0) MVG_PD(false, 0); else MVG_PD(false, 1);
  }
#undef MVG_PD
  MVG_LAUNCH_CHECK("flat_pair_dist_kernel");
}

void sq8_pair_distances(const uint8_t *qcodes, const uint32_t *qnorms, const int64_t *qid,
                        int64_t per_query, const uint8_t *codes, const uint32_t *norms,
                        const int64_t *xid, int64_t npairs, int ld, int metric, float scale_sq,
                        float *out, hipStream_t s) {
  if (npairs == 0) return;
  const dim3 grid((unsigned)ceil_div(npairs * 16, 256));
  if (metric == kMetricIP)
    hipLaunchKernelGGL(sq8_pair_dist_kernel<true>, grid, dim3(256), 0, s, qcodes, qnorms, qid,
                       per_query, codes, norms, xid, npairs, ld, scale_sq, out);
  else
    hipLaunchKernelGGL(sq8_pair_dist_kernel<false>, grid, dim3(256), 0, s, qcodes, qnorms, qid,
                       per_query, codes, norms, xid, npairs, ld, scale_sq, out);
  MVG_LAUNCH_CHECK("sq8_pair_dist_kernel");
}

}  // namespace mvg
