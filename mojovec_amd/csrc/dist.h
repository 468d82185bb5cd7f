// Exact-order FP32 distance routines.
//
// The plain (FP32) form of every kernel reproduces the reference's operation
// order so that distances, argmins and top-k ids agree bit for bit with the
// CPU restatement:
//   l2_distance_simd[W] / inner_product_simd[W]  mojovec/utils/distances.mojo:43-228
//   l2_distance_short4 / inner_product_short4    mojovec/utils/distances.mojo:9-39
// Structure: four W-lane FMA accumulators over 4W blocks, lane-wise
// ((a0 + a1) + a2) + a3, W-blocks, then the lanes are folded by the halving
// tree v[j] += v[j + W/2] ... (SIMD.reduce_add), then narrower tails.
// Everything here is compiled with -ffp-contract=off: only __builtin_fmaf
// fuses.
#pragma once

#include "common.h"

namespace mvg {

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <int W>
__device__ __forceinline__ float tree_sum(float *v) {
#pragma unroll
  for (int h = W / 2; h >= 1; h >>= 1) {
#pragma unroll
    for (int j = 0; j < h; j++) v[j] = v[j] + v[j + h];
  }
  return v[0];
}

// l2_distance_simd[8] for a runtime dimension (flat / coarse / k-means).
template <typename PX, typename PY>
__device__ float l2_w8(PX x, PY y, int d) {
  if (d < 8) {
    float r = 0.0f;
    for (int i = 0; i < d; i++) {
      float df = x[i] - y[i];
      r = r + df * df;
    }
    return r;
  }
  float a0[8], a1[8], a2[8], a3[8];
#pragma unroll
  for (int j = 0; j < 8; j++) a0[j] = a1[j] = a2[j] = a3[j] = 0.0f;
  int i = 0;
  for (; i <= d - 32; i += 32) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float d0 = x[i + j] - y[i + j];
      float d1 = x[i + 8 + j] - y[i + 8 + j];
      float d2 = x[i + 16 + j] - y[i + 16 + j];
      float d3 = x[i + 24 + j] - y[i + 24 + j];
      a0[j] = fma_(d0, d0, a0[j]);
      a1[j] = fma_(d1, d1, a1[j]);
      a2[j] = fma_(d2, d2, a2[j]);
      a3[j] = fma_(d3, d3, a3[j]);
    }
  }
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = ((a0[j] + a1[j]) + a2[j]) + a3[j];
  for (; i <= d - 8; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float df = x[i + j] - y[i + j];
      v[j] = fma_(df, df, v[j]);
    }
  }
  float res = tree_sum<8>(v);
  for (; i <= d - 4; i += 4) {
    float r[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float df = x[i + j] - y[i + j];
      r[j] = df * df;
    }
    res += tree_sum<4>(r);
  }
  for (; i < d; i++) {
    float df = x[i] - y[i];
    res += df * df;
  }
  return res;
}

// inner_product_simd[W] for W in {4, 8} and a runtime dimension.
template <int W, typename PX, typename PY>
__device__ float ip_wn(PX x, PY y, int d) {
  if (d < W) {
    float r = 0.0f;
    for (int i = 0; i < d; i++) r = r + x[i] * y[i];
    return r;
  }
  float p0[W], p1[W], p2[W], p3[W];
#pragma unroll
  for (int j = 0; j < W; j++) p0[j] = p1[j] = p2[j] = p3[j] = 0.0f;
  int i = 0;
  for (; i <= d - 4 * W; i += 4 * W) {
#pragma unroll
    for (int j = 0; j < W; j++) {
      p0[j] = fma_(x[i + j], y[i + j], p0[j]);
      p1[j] = fma_(x[i + W + j], y[i + W + j], p1[j]);
      p2[j] = fma_(x[i + 2 * W + j], y[i + 2 * W + j], p2[j]);
      p3[j] = fma_(x[i + 3 * W + j], y[i + 3 * W + j], p3[j]);
    }
  }
  float v[W];
#pragma unroll
  for (int j = 0; j < W; j++) v[j] = ((p0[j] + p1[j]) + p2[j]) + p3[j];
  for (; i <= d - W; i += W) {
#pragma unroll
    for (int j = 0; j < W; j++) v[j] = fma_(x[i + j], y[i + j], v[j]);
  }
  float res = tree_sum<W>(v);
  if (W >= 8) {
    for (; i <= d - 4; i += 4) {
      float r[4];
#pragma unroll
      for (int j = 0; j < 4; j++) r[j] = fma_(x[i + j], y[i + j], 0.0f);
      res += tree_sum<4>(r);
    }
  }
  for (; i < d; i++) res += x[i] * y[i];
  return res;
}

// l2_distance_short4 / inner_product_short4 (d <= 4).
template <typename PX, typename PY>
__device__ __forceinline__ float l2_short4(PX x, PY y, int d) {
  if (d == 4) {
    float r[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float df = x[j] - y[j];
      r[j] = df * df;
    }
    return tree_sum<4>(r);
  }
  float res = 0.0f;
  for (int i = 0; i < d; i++) {
    float df = x[i] - y[i];
    res += df * df;
  }
  return res;
}

template <typename PX, typename PY>
__device__ __forceinline__ float ip_short4(PX x, PY y, int d) {
  if (d == 4) {
    float r[4];
#pragma unroll
    for (int j = 0; j < 4; j++) r[j] = x[j] * y[j];
    return tree_sum<4>(r);
  }
  float res = 0.0f;
  for (int i = 0; i < d; i++) res += x[i] * y[i];
  return res;
}

// K-means / PQ distance selector (kmeans.mojo:131-138, pq.mojo:206-213).
template <typename PX, typename PY>
__device__ __forceinline__ float l2_kmeans(PX x, PY y, int d) {
  return d <= 4 ? l2_short4(x, y, d) : l2_w8(x, y, d);
}

// PQ inner-product LUT entry (pq.mojo:285-292): -ip with W=4 or short4.
template <typename PX, typename PY>
__device__ __forceinline__ float ip_pq(PX x, PY y, int d) {
  return d <= 4 ? ip_short4(x, y, d) : ip_wn<4>(x, y, d);
}

// Compile-time-dimension forms of l2_distance_simd[8] / inner_product_simd[8]
// for D >= 8 or D <= 4 (the LUT entry of 4 < dsub < 8 takes the scalar path
// of the runtime routine instead). `xr` is usually a register array, `q` the
// LDS operand; the difference is q - xr as in the reference (x = query).
template <int D>
__device__ __forceinline__ float l2_w8_fixed(const float *xr, const float *q) {
  float a0[8], a1[8], a2[8], a3[8];
#pragma unroll
  for (int j = 0; j < 8; j++) a0[j] = a1[j] = a2[j] = a3[j] = 0.0f;
  constexpr int NB4 = D / 32;
#pragma unroll
  for (int b = 0; b < NB4; b++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int i = b * 32;
      float d0 = q[i + j] - xr[i + j];
      float d1 = q[i + 8 + j] - xr[i + 8 + j];
      float d2 = q[i + 16 + j] - xr[i + 16 + j];
      float d3 = q[i + 24 + j] - xr[i + 24 + j];
      a0[j] = fma_(d0, d0, a0[j]);
      a1[j] = fma_(d1, d1, a1[j]);
      a2[j] = fma_(d2, d2, a2[j]);
      a3[j] = fma_(d3, d3, a3[j]);
    }
  }
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = ((a0[j] + a1[j]) + a2[j]) + a3[j];
  constexpr int I1 = NB4 * 32;
  constexpr int NW8 = (D - I1) / 8;
#pragma unroll
  for (int w = 0; w < NW8; w++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float df = q[I1 + w * 8 + j] - xr[I1 + w * 8 + j];
      v[j] = fma_(df, df, v[j]);
    }
  }
  float res = tree_sum<8>(v);
  constexpr int I2 = I1 + NW8 * 8;
  constexpr int NW4 = (D - I2) / 4;
#pragma unroll
  for (int w = 0; w < NW4; w++) {
    float r[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float df = q[I2 + w * 4 + j] - xr[I2 + w * 4 + j];
      r[j] = df * df;
    }
    res += tree_sum<4>(r);
  }
#pragma unroll
  for (int i = I2 + NW4 * 4; i < D; i++) {
    float df = q[i] - xr[i];
    res += df * df;
  }
  return res;
}

template <int D>
__device__ __forceinline__ float ip_w8_fixed(const float *xr, const float *q) {
  float p0[8], p1[8], p2[8], p3[8];
#pragma unroll
  for (int j = 0; j < 8; j++) p0[j] = p1[j] = p2[j] = p3[j] = 0.0f;
  constexpr int NB4 = D / 32;
#pragma unroll
  for (int b = 0; b < NB4; b++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int i = b * 32;
      p0[j] = fma_(q[i + j], xr[i + j], p0[j]);
      p1[j] = fma_(q[i + 8 + j], xr[i + 8 + j], p1[j]);
      p2[j] = fma_(q[i + 16 + j], xr[i + 16 + j], p2[j]);
      p3[j] = fma_(q[i + 24 + j], xr[i + 24 + j], p3[j]);
    }
  }
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = ((p0[j] + p1[j]) + p2[j]) + p3[j];
  constexpr int I1 = NB4 * 32;
  constexpr int NW8 = (D - I1) / 8;
#pragma unroll
  for (int w = 0; w < NW8; w++) {
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = fma_(q[I1 + w * 8 + j], xr[I1 + w * 8 + j], v[j]);
  }
  float res = tree_sum<8>(v);
  constexpr int I2 = I1 + NW8 * 8;
  constexpr int NW4 = (D - I2) / 4;
#pragma unroll
  for (int w = 0; w < NW4; w++) {
    float r[4];
#pragma unroll
    for (int j = 0; j < 4; j++) r[j] = fma_(q[I2 + w * 4 + j], xr[I2 + w * 4 + j], 0.0f);
    res += tree_sum<4>(r);
  }
#pragma unroll
  for (int i = I2 + NW4 * 4; i < D; i++) res += q[i] * xr[i];
  return res;
}

// Compile-time-dimension form of inner_product_simd[4] (the IVF-Flat IP
// distance, index_ivf_flat.mojo:225) for D >= 4: four 4-lane accumulators
// over 16-element blocks, 4-element blocks, the 4-lane tree, a scalar tail.
template <int D>
__device__ __forceinline__ float ip_w4_fixed(const float *xr, const float *q) {
  float p0[4], p1[4], p2[4], p3[4];
#pragma unroll
  for (int j = 0; j < 4; j++) p0[j] = p1[j] = p2[j] = p3[j] = 0.0f;
  constexpr int NB4 = D / 16;
#pragma unroll
  for (int b = 0; b < NB4; b++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int i = b * 16;
      p0[j] = fma_(q[i + j], xr[i + j], p0[j]);
      p1[j] = fma_(q[i + 4 + j], xr[i + 4 + j], p1[j]);
      p2[j] = fma_(q[i + 8 + j], xr[i + 8 + j], p2[j]);
      p3[j] = fma_(q[i + 12 + j], xr[i + 12 + j], p3[j]);
    }
  }
  float v[4];
#pragma unroll
  for (int j = 0; j < 4; j++) v[j] = ((p0[j] + p1[j]) + p2[j]) + p3[j];
  constexpr int I1 = NB4 * 16;
  constexpr int NW4 = (D - I1) / 4;
#pragma unroll
  for (int w = 0; w < NW4; w++) {
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = fma_(q[I1 + w * 4 + j], xr[I1 + w * 4 + j], v[j]);
  }
  float res = tree_sum<4>(v);
#pragma unroll
  for (int i = I1 + NW4 * 4; i < D; i++) res += q[i] * xr[i];
  return res;
}

// ---------------------------------------------------------------------------
// Wavefront forms of the W = 64 routines used by the exact rerank
// (index_hnsw.mojo:456-465): lane j owns SIMD lane j; the halving tree is a
// shuffle-down ladder. Every lane returns the full result.

__device__ __forceinline__ float wave_tree64(float v) {
#pragma unroll
  for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
  return __shfl(v, 0, 64);
}

template <int N>
__device__ __forceinline__ float wave_tree_n(float v) {
#pragma unroll
  for (int h = N / 2; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
  return __shfl(v, 0, 64);
}

// The same W = 64 distance computed by a group of 16 lanes, each holding four
// consecutive lanes of the 64-wide accumulators as a float4, for d a positive
// multiple of 64 (no tails). The four-accumulator blocks, the W blocks and
// every pairing of the halving tree are those of wave_dist_w64: virtual
// strides 32, 16, 8 and 4 are lane strides 8, 4, 2 and 1, strides 2 and 1
// fold inside the float4. The result is bit-identical and valid in glane 0.
// q and x must be 16-byte aligned.
template <bool IP>
__device__ __forceinline__ void w64_step(float4 &acc, float4 qv, float4 xv) {
  if (IP) {
    acc.x = fma_(qv.x, xv.x, acc.x);
    acc.y = fma_(qv.y, xv.y, acc.y);
    acc.z = fma_(qv.z, xv.z, acc.z);
    acc.w = fma_(qv.w, xv.w, acc.w);
  } else {
    const float dx = qv.x - xv.x, dy = qv.y - xv.y, dz = qv.z - xv.z, dw = qv.w - xv.w;
    acc.x = fma_(dx, dx, acc.x);
    acc.y = fma_(dy, dy, acc.y);
    acc.z = fma_(dz, dz, acc.z);
    acc.w = fma_(dw, dw, acc.w);
  }
}

template <bool IP>
__device__ float group16_dist_w64(const float *q, const float *x, int d, int glane) {
  const float4 *q4 = reinterpret_cast<const float4 *>(q);
  const float4 *x4 = reinterpret_cast<const float4 *>(x);
  float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0, a3 = a0;
  const int n4 = d >> 2;
  int i = 0;  // float4 index; a 4W block is 64 float4, a W block 16
  for (; i + 64 <= n4; i += 64) {
    w64_step<IP>(a0, q4[i + glane], x4[i + glane]);
    w64_step<IP>(a1, q4[i + 16 + glane], x4[i + 16 + glane]);
    w64_step<IP>(a2, q4[i + 32 + glane], x4[i + 32 + glane]);
    w64_step<IP>(a3, q4[i + 48 + glane], x4[i + 48 + glane]);
  }
  float4 v;
  v.x = ((a0.x + a1.x) + a2.x) + a3.x;
  v.y = ((a0.y + a1.y) + a2.y) + a3.y;
  v.z = ((a0.z + a1.z) + a2.z) + a3.z;
  v.w = ((a0.w + a1.w) + a2.w) + a3.w;
  for (; i + 16 <= n4; i += 16) w64_step<IP>(v, q4[i + glane], x4[i + glane]);
#pragma unroll
  for (int h = 8; h >= 1; h >>= 1) {
    v.x = v.x + __shfl_down(v.x, h, 16);
    v.y = v.y + __shfl_down(v.y, h, 16);
    v.z = v.z + __shfl_down(v.z, h, 16);
    v.w = v.w + __shfl_down(v.w, h, 16);
  }
  return (v.x + v.z) + (v.y + v.w);
}

// group16_dist_w64 of R rows at once (each row's operations exactly as
// above; the rows' loads are issued together, so R rows are in flight per
// group). BIG: d may reach the four-accumulator blocks (d >= 256); without it
// those blocks are compiled out (d < 256 never enters them). Results valid in
// glane 0.
template <bool IP, int R, bool BIG>
__device__ __forceinline__ void group16_distR_w64(const float *q, const float *const (&x)[R], int d,
                                                  int glane, float (&r)[R]) {
  const float4 *q4 = reinterpret_cast<const float4 *>(q);
  const int n4 = d >> 2;
  int i = 0;
  float4 v[R];
  if constexpr (BIG) {
    float4 a[R][4];
#pragma unroll
    for (int t = 0; t < R; t++)
#pragma unroll
      for (int b = 0; b < 4; b++) a[t][b] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (; i + 64 <= n4; i += 64) {
      float4 xv[R][4];
#pragma unroll
      for (int t = 0; t < R; t++)
#pragma unroll
        for (int b = 0; b < 4; b++) xv[t][b] = reinterpret_cast<const float4 *>(x[t])[i + 16 * b + glane];
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const float4 qv = q4[i + 16 * b + glane];
#pragma unroll
        for (int t = 0; t < R; t++) w64_step<IP>(a[t][b], qv, xv[t][b]);
      }
    }
#pragma unroll
    for (int t = 0; t < R; t++) {
      v[t].x = ((a[t][0].x + a[t][1].x) + a[t][2].x) + a[t][3].x;
      v[t].y = ((a[t][0].y + a[t][1].y) + a[t][2].y) + a[t][3].y;
      v[t].z = ((a[t][0].z + a[t][1].z) + a[t][2].z) + a[t][3].z;
      v[t].w = ((a[t][0].w + a[t][1].w) + a[t][2].w) + a[t][3].w;
    }
  } else {
    // ((0 + 0) + 0) + 0 of the empty four-accumulator stage
#pragma unroll
    for (int t = 0; t < R; t++) v[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  for (; i + 16 <= n4; i += 16) {
    float4 xv[R];
#pragma unroll
    for (int t = 0; t < R; t++) xv[t] = reinterpret_cast<const float4 *>(x[t])[i + glane];
    const float4 qv = q4[i + glane];
#pragma unroll
    for (int t = 0; t < R; t++) w64_step<IP>(v[t], qv, xv[t]);
  }
#pragma unroll
  for (int h = 8; h >= 1; h >>= 1) {
#pragma unroll
    for (int t = 0; t < R; t++) {
      v[t].x = v[t].x + __shfl_down(v[t].x, h, 16);
      v[t].y = v[t].y + __shfl_down(v[t].y, h, 16);
      v[t].z = v[t].z + __shfl_down(v[t].z, h, 16);
      v[t].w = v[t].w + __shfl_down(v[t].w, h, 16);
    }
  }
#pragma unroll
  for (int t = 0; t < R; t++) r[t] = (v[t].x + v[t].z) + (v[t].y + v[t].w);
}

template <bool IP>
__device__ float wave_dist_w64(const float *q, const float *x, int d, int lane) {
  if (d < 64) {
    float r = 0.0f;
    for (int i = 0; i < d; i++) {
      if (IP) {
        r = r + q[i] * x[i];
      } else {
        float df = q[i] - x[i];
        r = r + df * df;
      }
    }
    return r;
  }
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  int i = 0;
  for (; i <= d - 256; i += 256) {
    if (IP) {
      a0 = fma_(q[i + lane], x[i + lane], a0);
      a1 = fma_(q[i + 64 + lane], x[i + 64 + lane], a1);
      a2 = fma_(q[i + 128 + lane], x[i + 128 + lane], a2);
      a3 = fma_(q[i + 192 + lane], x[i + 192 + lane], a3);
    } else {
      float d0 = q[i + lane] - x[i + lane];
      float d1 = q[i + 64 + lane] - x[i + 64 + lane];
      float d2 = q[i + 128 + lane] - x[i + 128 + lane];
      float d3 = q[i + 192 + lane] - x[i + 192 + lane];
      a0 = fma_(d0, d0, a0);
      a1 = fma_(d1, d1, a1);
      a2 = fma_(d2, d2, a2);
      a3 = fma_(d3, d3, a3);
    }
  }
  float v = ((a0 + a1) + a2) + a3;
  for (; i <= d - 64; i += 64) {
    if (IP) {
      v = fma_(q[i + lane], x[i + lane], v);
    } else {
      float df = q[i + lane] - x[i + lane];
      v = fma_(df, df, v);
    }
  }
  float res = wave_tree64(v);
  for (; i <= d - 16; i += 16) {
    float r = 0.0f;
    if (lane < 16) {
      if (IP) {
        r = fma_(q[i + lane], x[i + lane], 0.0f);
      } else {
        float df = q[i + lane] - x[i + lane];
        r = df * df;
      }
    }
    res += wave_tree_n<16>(r);
  }
  for (; i <= d - 8; i += 8) {
    float r = 0.0f;
    if (lane < 8) {
      if (IP) {
        r = fma_(q[i + lane], x[i + lane], 0.0f);
      } else {
        float df = q[i + lane] - x[i + lane];
        r = df * df;
      }
    }
    res += wave_tree_n<8>(r);
  }
  for (; i <= d - 4; i += 4) {
    float r = 0.0f;
    if (lane < 4) {
      if (IP) {
        r = fma_(q[i + lane], x[i + lane], 0.0f);
      } else {
        float df = q[i + lane] - x[i + lane];
        r = df * df;
      }
    }
    res += wave_tree_n<4>(r);
  }
  for (; i < d; i++) {
    if (IP) {
      res += q[i] * x[i];
    } else {
      float df = q[i] - x[i];
      res += df * df;
    }
  }
  return res;
}

}  // namespace mvg
