// GEMM-form distances on MFMA with a running per-query top-k.
//
// dist(q, x) comes from one dense contraction q . x on the matrix cores:
//   BF16_L2 : |q|^2 + |x|^2 - 2 q.x   (bf16 operands, FP32 accumulate)
//   BF16_IP : -q.x
//   I8_IP   : -Float32(q.x) * scale^2 with int8 codes, int32 accumulate
//             (sq8.mojo:71-75: exact)
//   U8_L2   : Float32(max(|q|^2 + |x|^2 - 2 q.x, 0)) * scale^2 with affine
//             uint8 codes rebased to int8 (s = u ^ 0x80 = u - 128):
//             q.x = s_q.s_x + 128 (sum s_q + sum s_x) + 16384 d, exact in int32
//             (sq8.mojo:76-79, distances.mojo:437-451)
// A workgroup owns 16 queries and one contiguous split of the base rows. Each
// of its four wavefronts computes 16 x 64 tiles with v_mfma 16x16 instructions
// (K = 32 bf16 or 64 int8 per instruction), and each query keeps its running
// top-kk in LDS behind a threshold, so only candidates that beat the current
// kk-th key are appended and merged.
#include "dist.h"
#include "kernels.h"
#include "select.h"

namespace mvg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx4 __attribute__((ext_vector_type(4)));

constexpr int kQT = 16;        // queries per workgroup
constexpr int kStepCols = 256;  // base rows per step (4 waves x 4 tiles x 16)
constexpr int kGCap = 384;      // per-query append capacity

enum GemmKind { BF16_L2 = 0, BF16_IP = 1, I8_IP = 2, U8_L2 = 3 };

struct GemmLayout {
  int qtile, top, tmp, cnt, thr, scr, total;
  __host__ __device__ GemmLayout(int row_bytes, int kk) {
    int np2 = 1;
    while (np2 < kk) np2 <<= 1;
    qtile = 0;
    top = (kQT * row_bytes + 15) & ~15;
    tmp = top + 8 * kQT * (kk + kGCap);
    cnt = tmp + 8 * np2;
    thr = cnt + 4 * kQT;
    thr = (thr + 7) & ~7;
    scr = thr + 8 * kQT;
    total = scr + (int)sizeof(SelectScratch64);
  }
};

// Merges query qi's append buffer into its sorted top set.
__device__ void gemm_flush(uint64_t *top, uint64_t *tmp, int kk, int cnt, SelectScratch64 &scr,
                           uint64_t &thr_out) {
  const int tid = threadIdx.x;
  const int n = kk + cnt;
  if (n <= 2 * kNT) {
    for (int i = tid; i < kk; i += kNT) tmp[i] = kSentinel64;
    __syncthreads();
    for (int i = tid; i < n; i += kNT) {
      const uint64_t key = top[i];
      if (key == kSentinel64) continue;
      int r = 0;
      for (int j = 0; j < n; j++) r += top[j] < key;
      if (r < kk) tmp[r] = key;
    }
    __syncthreads();
  } else {
    select_sort_u64(top, n, kk, scr, tmp);
  }
  for (int i = tid; i < kk; i += kNT) top[i] = tmp[i];
  __syncthreads();
  if (tid == 0) thr_out = top[kk - 1];
}

template <int KIND>
__global__ __launch_bounds__(kNT) void gemm_topk_kernel(
    const uint8_t *__restrict__ Q, const uint8_t *__restrict__ X, int64_t nq, int64_t n,
    int row_bytes, int d_true, const float *__restrict__ qnorm, const float *__restrict__ xnorm,
    const uint32_t *__restrict__ qn_u, const uint32_t *__restrict__ qs_u,
    const uint32_t *__restrict__ xn_u, const uint32_t *__restrict__ xs_u, float scale_sq, int kk,
    int64_t split_len, const uint8_t *__restrict__ excl, int64_t excl_len,
    uint64_t *__restrict__ part, int S) {
  extern __shared__ __align__(16) unsigned char smem[];
  const GemmLayout lay(row_bytes, kk);
  uint8_t *qtile = smem + lay.qtile;
  uint64_t *top = reinterpret_cast<uint64_t *>(smem + lay.top);  // [kQT][kk + kGCap]
  uint64_t *tmp = reinterpret_cast<uint64_t *>(smem + lay.tmp);
  int *cnt = reinterpret_cast<int *>(smem + lay.cnt);
  uint64_t *thr = reinterpret_cast<uint64_t *>(smem + lay.thr);
  SelectScratch64 *scr = reinterpret_cast<SelectScratch64 *>(smem + lay.scr);
  const int stride = kk + kGCap;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t q0 = (int64_t)blockIdx.x * kQT;
  const int split = blockIdx.y;
  const int64_t b_begin = (int64_t)split * split_len;
  const int64_t b_end = min(n, b_begin + split_len);
  const int qcount = (int)min<int64_t>(kQT, nq - q0);

  // Query tile -> LDS (zero rows past nq).
  const int row_words = row_bytes / 16;
  for (int i = tid; i < kQT * row_words; i += kNT) {
    const int qi = i / row_words;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (qi < qcount) v = reinterpret_cast<const uint4 *>(Q + (q0 + qi) * row_bytes)[i - qi * row_words];
    reinterpret_cast<uint4 *>(qtile)[i] = v;
  }
  for (int i = tid; i < kQT * stride; i += kNT) top[i] = kSentinel64;
  if (tid < kQT) {
    cnt[tid] = 0;
    thr[tid] = kSentinel64;
  }
  __syncthreads();

  const int nsteps = row_bytes / 64;
  const int lrow = lane & 15;   // A row / B column of this lane
  const int lgrp = lane >> 4;   // 16-byte group within the 64-byte k slice
  for (int64_t b0 = b_begin; b0 < b_end; b0 += kStepCols) {
    floatx4 accf[4];
    intx4 acci[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      accf[t] = floatx4{0.f, 0.f, 0.f, 0.f};
      acci[t] = intx4{0, 0, 0, 0};
    }
    int64_t rows[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int64_t r = b0 + wave * 64 + t * 16 + lrow;
      rows[t] = r < b_end ? r : b_end - 1;
    }
    for (int ks = 0; ks < nsteps; ks++) {
      const uint4 a = *reinterpret_cast<const uint4 *>(qtile + lrow * row_bytes + ks * 64 + lgrp * 16);
#pragma unroll
      for (int t = 0; t < 4; t++) {
        uint4 b = *reinterpret_cast<const uint4 *>(X + rows[t] * row_bytes + ks * 64 + lgrp * 16);
        if (KIND == BF16_L2 || KIND == BF16_IP) {
          accf[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                            __builtin_bit_cast(bf16x8, b), accf[t], 0, 0, 0);
        } else {
          uint4 aa = a;
          if (KIND == U8_L2) {  // u - 128 as int8
            aa.x ^= 0x80808080u; aa.y ^= 0x80808080u; aa.z ^= 0x80808080u; aa.w ^= 0x80808080u;
            b.x ^= 0x80808080u; b.y ^= 0x80808080u; b.z ^= 0x80808080u; b.w ^= 0x80808080u;
          }
          acci[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(intx4, aa),
                                                          __builtin_bit_cast(intx4, b), acci[t], 0, 0, 0);
        }
      }
    }
    // lane holds D[4 * lgrp + j][lrow] of each tile
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int64_t col = b0 + wave * 64 + t * 16 + lrow;
      const bool col_ok = col < b_end && !(excl && col < excl_len && excl[col] > 0);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int qi = 4 * lgrp + j;
        if (!col_ok || qi >= qcount) continue;
        float dist;
        if (KIND == BF16_L2) {
          dist = (qnorm[q0 + qi] + xnorm[col]) - 2.0f * accf[t][j];
        } else if (KIND == BF16_IP) {
          dist = -accf[t][j];
        } else if (KIND == I8_IP) {
          dist = __fmul_rn(-(float)acci[t][j], scale_sq);
        } else {
          const int64_t sq = (int64_t)qs_u[q0 + qi] - 128 * (int64_t)d_true;
          const int64_t sx = (int64_t)xs_u[col] - 128 * (int64_t)d_true;
          const int64_t dot = (int64_t)acci[t][j] + 128 * (sq + sx) + 16384 * (int64_t)d_true;
          const int64_t v = (int64_t)qn_u[q0 + qi] + (int64_t)xn_u[col] - 2 * dot;
          dist = __fmul_rn((float)(uint32_t)(v > 0 ? v : 0), scale_sq);
        }
        const uint64_t key = pack_key(dist, (uint32_t)col);
        if (key < thr[qi]) top[qi * stride + kk + atomicAdd(&cnt[qi], 1)] = key;
      }
    }
    __syncthreads();
    bool need = false;
    for (int qi = 0; qi < kQT; qi++) need |= cnt[qi] > kGCap - kStepCols;
    if (need || b0 + kStepCols >= b_end) {
      for (int qi = 0; qi < qcount; qi++) {
        const int c = cnt[qi];
        if (c == 0) continue;
        gemm_flush(top + qi * stride, tmp, kk, c, *scr, thr[qi]);
        __syncthreads();
        if (tid == 0) cnt[qi] = 0;
        __syncthreads();
      }
    }
  }
  for (int qi = 0; qi < qcount; qi++) {
    uint64_t *dst = part + ((q0 + qi) * S + split) * (int64_t)kk;
    for (int t = tid; t < kk; t += kNT) dst[t] = top[qi * stride + t];
  }
}

// FP32 rows -> bf16 rows padded to row_elems (round to nearest even), plus
// the FP32 squared norm of the rounded row.
__global__ void to_bf16_kernel(const float *__restrict__ x, int64_t n, int d, int row_elems,
                               uint16_t *__restrict__ out, float *__restrict__ norms) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  float acc = 0.0f;
  for (int j = lane; j < row_elems; j += 64) {
    uint16_t h = 0;
    if (j < d) {
      const __bf16 b = (__bf16)x[row * d + j];
      h = __builtin_bit_cast(uint16_t, b);
      const float r = (float)b;
      acc = fma_(r, r, acc);
    }
    out[row * row_elems + j] = h;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0 && norms) norms[row] = acc;
}

}  // namespace

int gemm_row_bytes(int kind, int d) {
  const int bytes = (kind == BF16_L2 || kind == BF16_IP) ? 2 * d : d;
  return (bytes + 63) & ~63;
}

void to_bf16(const float *x, int64_t n, int d, uint16_t *out, float *norms, hipStream_t s) {
  if (n == 0) return;
  const int row_elems = gemm_row_bytes(BF16_L2, d) / 2;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, s, x, n, d, row_elems, out,
                     norms);
  MVG_LAUNCH_CHECK("to_bf16_kernel");
}

void gemm_topk(int kind, const uint8_t *Q, const uint8_t *X, int64_t nq, int64_t n, int d,
               const float *qnorm, const float *xnorm, const uint32_t *qn_u, const uint32_t *qs_u,
               const uint32_t *xn_u, const uint32_t *xs_u, float scale_sq, int kk, const uint8_t *excl, int64_t excl_len, DevBuf &part,
               int &nsplit, hipStream_t s) {
  const int row_bytes = gemm_row_bytes(kind, d);
  const GemmLayout lay(row_bytes, kk);
  if (lay.total > 150 * 1024) throw Error(kInvalid, "GEMM top-k workspace exceeds LDS");
  const int qtiles = ceil_div(nq, kQT);
  // Split the base so that the grid holds ~2048 workgroups, at least one step each.
  int64_t S = std::max<int64_t>(1, std::min<int64_t>(ceil_div(2048, qtiles), ceil_div(n, kStepCols)));
  int64_t split_len = ((ceil_div(n, S) + kStepCols - 1) / kStepCols) * kStepCols;
  S = ceil_div(n, split_len);
  nsplit = (int)S;
  part.reserve(sizeof(uint64_t) * (size_t)nq * S * kk);
  const void *fn;
  switch (kind) {
    case BF16_L2: fn = (const void *)gemm_topk_kernel<BF16_L2>; break;
    case BF16_IP: fn = (const void *)gemm_topk_kernel<BF16_IP>; break;
    case I8_IP: fn = (const void *)gemm_topk_kernel<I8_IP>; break;
    default: fn = (const void *)gemm_topk_kernel<U8_L2>; break;
  }
  if (lay.total > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lay.total));
  dim3 grid(qtiles, (unsigned)S);
#define MVG_GEMM(K)                                                                               \
  hipLaunchKernelGGL(gemm_topk_kernel<K>, grid, dim3(kNT), lay.total, s, Q, X, nq, n, row_bytes, d, \
                     qnorm, xnorm, qn_u, qs_u, xn_u, xs_u, scale_sq, kk, split_len, excl,        \
                     excl_len,                                                                     \
                     part.as<uint64_t>(), (int)S);
  switch (kind) {
    case BF16_L2: MVG_GEMM(BF16_L2) break;
    case BF16_IP: MVG_GEMM(BF16_IP) break;
    case I8_IP: MVG_GEMM(I8_IP) break;
    default: MVG_GEMM(U8_L2) break;
  }
#undef MVG_GEMM
  MVG_LAUNCH_CHECK("gemm_topk_kernel");
}

}  // namespace mvg
