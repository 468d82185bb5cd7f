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
//   FP8_L2  : |q|^2 + |x|^2 - 2 q.x, FP8 E4M3 operands, FP32 accumulate; each
//             row is stored scaled by a power of two that puts its largest
//             magnitude in [224, 448], and q.x = acc * (1/s_q) * (1/s_x)
//   FP8_IP  : -q.x ~ c_q + c_x - q~.x~ with the same operands, where
//             c = (|v~|^2 - |v|^2) / 2 over the dequantized row v~: since
//             -q.x = (|q - x|^2 - |q|^2 - |x|^2) / 2, the estimate's error is
//             (x - q).(dq - dx) for quantization errors dq, dx -- small for
//             near neighbours, where -q~.x~ alone errs by q.dx + x.dq
// A workgroup owns 16 queries and one contiguous split of the base rows. Each
// of its four wavefronts computes 16 x 64 tiles with v_mfma 16x16 instructions
// (K = 32 bf16 or 64 int8 per instruction), and each query keeps its running
// top-kk in LDS behind a threshold, so only candidates that beat the current
// kk-th key are appended and merged.
#include <algorithm>
#include <utility>

#include "dist.h"
#include "kernels.h"
#include "select.h"

namespace mvg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef long longx1;

constexpr int kQT = 16;        // queries per workgroup
constexpr int kStepCols = 256;  // base rows per step (4 waves x 4 tiles x 16)
// Per-query append capacity: 384 keys; 768 above 64 candidates when that
// keeps as many workgroups per CU (LDS) and fits in 150 KB, since fewer
// flushes of a larger buffer cost less there (each is a radix select, O(n)).
constexpr int kGCap = 384, kGCapLarge = 768;

enum GemmKind { BF16_L2 = 0, BF16_IP = 1, I8_IP = 2, U8_L2 = 3, FP8_L2 = 4, FP8_IP = 5 };

struct GemmLayout {
  int qtile, top, tmp, cnt, thr, scr, total, cap;
  __host__ __device__ GemmLayout(int row_bytes, int kk) {
    place(row_bytes, kk, kGCap);
    if (kk > 64) {
      const int small = total;
      place(row_bytes, kk, kGCapLarge);
      if (total > 150 * 1024 || (160 * 1024) / total < (160 * 1024) / small) place(row_bytes, kk, kGCap);
    }
  }
  __host__ __device__ void place(int row_bytes, int kk, int c) {
    int np2 = 1;
    while (np2 < kk) np2 <<= 1;
    cap = c;
    qtile = 0;
    top = (kQT * row_bytes + 15) & ~15;
    tmp = top + 8 * kQT * (kk + cap);
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
  if (n <= kNT / 2) {  // rank merge for small sets, radix select above
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

// Workgroup -> (query tile, base split) for a 1-D grid of 8 * per workgroups.
// Workgroups go to the eight XCDs in turn (linear id mod 8) and each XCD has
// its own L2, so XCD j takes the contiguous range [j * per, (j + 1) * per) of
// the split-major order: the workgroups resident on one XCD walk the same base
// rows for different query tiles, and one memory fetch serves all of them.
// False for the padding workgroups past qtiles * S.
__device__ __forceinline__ bool xcd_tile(int64_t nq, int rows, int S, int64_t &qt, int &split) {
  const int64_t qtiles = (nq + rows - 1) / rows;
  const int64_t per = gridDim.x >> 3;
  const int64_t L = blockIdx.x;
  const int64_t G = (L & 7) * per + (L >> 3);
  if (G >= qtiles * S) return false;
  split = (int)(G / qtiles);
  qt = G - (int64_t)split * qtiles;
  return true;
}

// Grid size for xcd_tile: qtiles * S rounded up to a multiple of 8.
inline unsigned xcd_grid(int64_t qtiles, int64_t S) { return (unsigned)(8 * ((qtiles * S + 7) / 8)); }

// Rows of one base split in the wide forms. P == 1: split s is the contiguous
// range [s * split_len, (s + 1) * split_len). P > 1 (strided splits): split s
// is residue r = s % R of chunk c = s / R (R = P: every residue; R = 1: only
// the rows 0, P, 2 P, ..., a 1/P sample), i.e. the rows r, r + P, r + 2 P, ...
// whose local index j = (row - r) / P lies in [c * split_len, (c + 1) *
// split_len). Local row j is base row goff + j * P.
struct SplitRows {
  int64_t goff, b_begin, b_end;
  __device__ __forceinline__ SplitRows(int64_t n, int split, int P, int R, int64_t split_len) {
    const int r = split % R;
    goff = r;
    b_begin = (int64_t)(split / R) * split_len;
    b_end = min((n - r + P - 1) / P, b_begin + split_len);
  }
};

// A split without rows (a residue's last chunk can be one): its queries' top
// lists are empty.
__device__ __forceinline__ void empty_split(uint64_t *part, int64_t q0, int qcount, int S, int split, int kk) {
  for (int i = threadIdx.x; i < qcount * kk; i += kNT)
    part[((q0 + i / kk) * S + split) * (int64_t)kk + i % kk] = kSentinel64;
}

template <int KIND>
__global__ __launch_bounds__(kNT) void gemm_topk_kernel(
    const uint8_t *__restrict__ Q, const uint8_t *__restrict__ X, int64_t nq, int64_t n,
    int row_bytes, int d_true, const float *__restrict__ qnorm, const float *__restrict__ xnorm,
    const uint32_t *__restrict__ qn_u, const uint32_t *__restrict__ qs_u,
    const uint32_t *__restrict__ xn_u, const uint32_t *__restrict__ xs_u, float scale_sq, int kk,
    int64_t split_len, const uint8_t *__restrict__ excl, int64_t excl_len,
    uint64_t *__restrict__ part, int S, const float *__restrict__ qinv,
    const float *__restrict__ xinv, const uint8_t *__restrict__ only) {
  extern __shared__ __align__(16) unsigned char smem[];
  const GemmLayout lay(row_bytes, kk);
  uint8_t *qtile = smem + lay.qtile;
  uint64_t *top = reinterpret_cast<uint64_t *>(smem + lay.top);  // [kQT][kk + lay.cap]
  uint64_t *tmp = reinterpret_cast<uint64_t *>(smem + lay.tmp);
  int *cnt = reinterpret_cast<int *>(smem + lay.cnt);
  uint64_t *thr = reinterpret_cast<uint64_t *>(smem + lay.thr);
  SelectScratch64 *scr = reinterpret_cast<SelectScratch64 *>(smem + lay.scr);
  const int stride = kk + lay.cap;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int64_t qt;
  int split;
  if (!xcd_tile(nq, kQT, S, qt, split)) return;
  const int64_t q0 = qt * kQT;
  const int64_t b_begin = (int64_t)split * split_len;
  const int64_t b_end = min(n, b_begin + split_len);
  const int qcount = (int)min<int64_t>(kQT, nq - q0);
  if (only) {  // only the tiles that hold a flagged query
    bool any = false;
    for (int qi = 0; qi < qcount; qi++) any |= only[q0 + qi] != 0;
    if (!any) return;
  }

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
    // four k steps per iteration: their 16 row loads are issued together
#pragma unroll 4
    for (int ks = 0; ks < nsteps; ks++) {
      const uint4 a = *reinterpret_cast<const uint4 *>(qtile + lrow * row_bytes + ks * 64 + lgrp * 16);
#pragma unroll
      for (int t = 0; t < 4; t++) {
        uint4 b = *reinterpret_cast<const uint4 *>(X + rows[t] * row_bytes + ks * 64 + lgrp * 16);
        if (KIND == BF16_L2 || KIND == BF16_IP) {
          accf[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                            __builtin_bit_cast(bf16x8, b), accf[t], 0, 0, 0);
        } else if (KIND == FP8_L2 || KIND == FP8_IP) {
          // two K = 32 steps over the lane's 16 bytes; A and B lanes hold the
          // same byte positions, so each k is paired exactly once
          const long alo = (long)(((uint64_t)a.y << 32) | a.x), ahi = (long)(((uint64_t)a.w << 32) | a.z);
          const long blo = (long)(((uint64_t)b.y << 32) | b.x), bhi = (long)(((uint64_t)b.w << 32) | b.z);
          accf[t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(alo, blo, accf[t], 0, 0, 0);
          accf[t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(ahi, bhi, accf[t], 0, 0, 0);
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
        } else if (KIND == FP8_L2) {
          dist = (qnorm[q0 + qi] + xnorm[col]) - 2.0f * (accf[t][j] * (qinv[q0 + qi] * xinv[col]));
        } else if (KIND == FP8_IP) {
          dist = (qnorm[q0 + qi] + xnorm[col]) - accf[t][j] * (qinv[q0 + qi] * xinv[col]);
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
    for (int qi = 0; qi < kQT; qi++) need |= cnt[qi] > lay.cap - kStepCols;
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

// ---------------------------------------------------------------------------
// Wide form for kk <= 64: a workgroup owns 64 queries (16 per wavefront) and
// one split of the base rows, so each base byte fetched serves 64 queries
// (four times the reuse of the 16-query form above). The 256 rows of a step
// are staged through LDS 128 bytes of k at a time (32 KB per slice, double
// buffered, the next slice's global loads in flight while the current one
// feeds the MFMAs; one barrier per slice), and every wavefront meets the 256
// columns with its 16 query rows in 16 MFMA tiles. Selection stays inside the
// wavefront: a distance is compared with its query's running kk-th distance
// in registers, the few that pass are appended to a 32-slot LDS buffer per
// query, and a buffer near full is sorted and merged into the query's sorted
// top kk with the lane-wide bitonic networks (no workgroup barrier).
constexpr int kWQ = 64;                   // queries per workgroup
constexpr int kWCap = 32;                 // append slots per query
constexpr int kWMaxK = 64;                // largest kk of the wide form (LDS permitting)
constexpr int kBPlane = kStepCols * 16;   // one 16-byte group of a staged slice
constexpr int kBSlice = 8 * kBPlane;      // 128 bytes of k for 256 rows

struct WideLayout {
  int apitch, a, b, colv, top, buf, cnt, total;
  __host__ __device__ WideLayout(int row_bytes, int kk) {
    apitch = row_bytes + 16;  // padded rows: the 16 lanes of a fragment read hit distinct banks
    a = 0;
    b = (kWQ * apitch + 15) & ~15;
    colv = b + 2 * kBSlice;    // per-column terms of a step, three steps in turn
    top = colv + 3 * kStepCols * 8;
    buf = top + 8 * kWQ * kk;
    cnt = buf + 8 * kWQ * kWCap;
    total = cnt + 4 * kWQ;
  }
};

// Column swizzle of group plane g (0..7): the eight groups of one column, and
// of two neighbouring columns, land in distinct banks when written, while a
// fragment read (16 consecutive columns of one plane) stays one contiguous
// 256-byte run.
__device__ __forceinline__ int wide_swz(int g) { return (int)((0xFCA96530u >> (4 * g)) & 0xFu); }

// Merges query row qq's append buffer into its sorted top kk; the whole
// wavefront calls it with the same qq. RANK: the rank merge below when the
// two fit in one wavefront (the uint8 kind keeps the sorting network: with
// the rank merge its kernels spilled).
template <bool RANK>
__device__ __forceinline__ void wide_flush(uint64_t *top, uint64_t *buf, int *cnt, int qq, int kk) {
  const int lane = threadIdx.x & 63;
  const int c = cnt[qq];
  if (RANK && kk + c <= 64) {
    // Rank merge: lanes 0..kk-1 hold the top set, the next c lanes the
    // appended keys. Each lane counts the keys below its own, read as
    // same-address broadcasts from LDS (independent reads; the sorting
    // network's shuffles are LDS round trips in a dependent chain, about
    // twenty times slower here), and stores its key at that rank. Real keys
    // are distinct (one per column); slots past the real keys were the
    // sentinel already and stay so.
    const uint64_t *t = top + qq * kk, *b = buf + qq * kWCap;
    const uint64_t v = lane < kk ? t[lane] : (lane < kk + c ? b[lane - kk] : kSentinel64);
    int rank = 0;
#pragma unroll 4
    for (int i = 0; i < kk; i++) rank += t[i] < v;
#pragma unroll 4
    for (int i = 0; i < c; i++) rank += b[i] < v;
    if (v != kSentinel64 && rank < kk) top[qq * kk + rank] = v;
    if (lane == 0) cnt[qq] = 0;
    return;
  }
  uint64_t cand = lane < c ? buf[qq * kWCap + lane] : kSentinel64;
  cand = wave_sort64(cand);
  const uint64_t cur = lane < kk ? top[qq * kk + lane] : kSentinel64;
  const uint64_t m = wave_merge64(cur, cand);
  if (lane < kk) top[qq * kk + lane] = m;
  if (lane == 0) cnt[qq] = 0;
}

// Per-lane selection state of one wavefront's 16 query rows in the wide
// forms: the lane's four queries are rows 4 * lgrp + j of the wavefront. The
// integer kinds test candidates with an exact integer prefilter derived from
// the running kk-th distance (a superset: the rounding of the FP32 distance
// is covered by a margin), then the exact key on the rare pass.
template <int KIND, bool STR, bool REG>
struct WideSel {
  float thr_f[4], qa[4], qb[4];
  uint64_t thr_k[4];  // each query's kk-th key (the sentinel while fewer): changes at merges only
  int64_t qc[4], qsx[4];
  int ithr[4];
  bool qok[4];
  int cnt_r[4];
  // STR (strided splits): local column j of the split is row goff + j *
  // stride of the base; otherwise the local column is the row
  int64_t goff = 0, stride = 1;
  // a fixed bound on each query's keys (strict; the sentinel when none):
  // the running threshold never rises above it
  uint64_t bnd_k[4];

  __device__ __forceinline__ void init(int64_t q0, int wq, int qcount, int lgrp, const float *qnorm,
                                       const float *qinv, const uint32_t *qn_u, const uint32_t *qs_u,
                                       int d_true) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int qi = wq + 4 * lgrp + j;
      qok[j] = qi < qcount;
      const int64_t qg = q0 + (qok[j] ? qi : 0);
      thr_f[j] = INFINITY;
      thr_k[j] = kSentinel64;
      qa[j] = (KIND == BF16_L2 || KIND == FP8_L2 || KIND == FP8_IP) ? qnorm[qg] : 0.0f;
      qb[j] = (KIND == FP8_L2 || KIND == FP8_IP) ? qinv[qg] : 0.0f;
      qc[j] = KIND == U8_L2 ? (int64_t)qn_u[qg] : 0;
      qsx[j] = KIND == U8_L2 ? (int64_t)qs_u[qg] - 128 * (int64_t)d_true : 0;
      ithr[j] = KIND == I8_IP ? -(1 << 30) : -600000000;  // no bound yet: everything passes
      cnt_r[j] = 0;
      bnd_k[j] = kSentinel64;
    }
  }

  // Keys up to bound[q * bstride] (inclusive) pass from the start; the
  // split's top list then holds all of its keys up to that bound while it is
  // not full.
  __device__ __forceinline__ void set_bound(const uint64_t *bound, int bstride, int64_t q0, int wq, int lgrp,
                                            float scale_sq, int d_true) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (!qok[j]) continue;
      const uint64_t b = bound[(q0 + wq + 4 * lgrp + j) * bstride];
      if (b == kSentinel64) continue;
      bnd_k[j] = thr_k[j] = b + 1;
      thr_f[j] = key2f((uint32_t)(b >> 32));
      int_thr(j, scale_sq, d_true);
    }
  }

  // I8_IP: dist = fmul_rn(-acc, s) <= thr  =>  acc >= floor(-thr / s) - 2.
  // U8_L2: dist = fmul_rn(max(v, 0), s) with v = qn + xn - 2 acc - 256 (sq + sx)
  // - 32768 d, and dist <= thr  =>  v <= vthr = floor(thr / s (1 + 2^-20)) + 4,
  // i.e. 2 acc >= ithr + (xn - 256 sx) with ithr = qn - 256 sq - 32768 d - vthr
  // (clamped: |acc| <= 128^2 d and d <= 1264 keep every sum inside int32).
  __device__ __forceinline__ void int_thr(int j, float scale_sq, int d_true) {
    if (KIND == I8_IP) {
      const float qt = -thr_f[j] / scale_sq;
      ithr[j] = thr_f[j] == INFINITY ? -(1 << 30)
                                     : (int)fminf(fmaxf(floorf(qt), -1073741824.0f), 1073741824.0f) - 2;
    } else if (KIND == U8_L2) {
      int64_t vthr;
      if (thr_f[j] == INFINITY) vthr = (int64_t)1 << 40;
      else if (thr_f[j] < 0.0f) vthr = -((int64_t)1 << 40);
      else {  // r is NaN for scale 0: no bound
        const float r = thr_f[j] / scale_sq;
        vthr = r <= 1.0e12f ? (int64_t)floorf(r + r * 0x1p-20f) + 4 : ((int64_t)1 << 40);
      }
      const int64_t c = qc[j] - 256 * qsx[j] - 32768 * (int64_t)d_true - vthr;
      ithr[j] = (int)max<int64_t>(-600000000, min<int64_t>(c, 300000000));
    }
  }

  // Candidates of tiles [T0, T1) of one step: tile t holds columns pb0 +
  // t * 16 + lrow (valid up to pb0 + lim), cv the step's per-column terms at
  // this lane's column. The tests of all the chunk's tiles come first and
  // take no branch while nothing passes (straight-line code the scheduler
  // can interleave with MFMAs); passing keys go to the queries' append
  // buffers, and a buffer near full is merged into its query's top kk and
  // the bounds are refreshed. EX: an exclusion mask may be given; without
  // one the column test is a single compare, with no branch per tile (int8,
  // uint8 and FP8 kinds at d = 768, BF16 at d = 128: 5-11% less time per
  // search). REG: the register-tile kernel's selection.
  template <int T0, int T1>
  __device__ __forceinline__ void tiles(const floatx4 (&accf)[16], const intx4 (&acci)[16], int64_t pb0,
                                        const uint2 *cv, int lim, int64_t b_begin, const uint8_t *excl,
                                        int64_t excl_len, const uint32_t *xn_u, const uint32_t *xs_u,
                                        float scale_sq, int d_true, uint64_t *top, uint64_t *buf, int *cnt,
                                        int kk) {
    // the BF16 register-tile kernel keeps the general test: it ran slower
    // with the mask-free one (register allocation)
    if (excl || (REG && (KIND == BF16_L2 || KIND == BF16_IP)))
      tiles_x<T0, T1, true>(accf, acci, pb0, cv, lim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true,
                            top, buf, cnt, kk);
    else
      tiles_x<T0, T1, false>(accf, acci, pb0, cv, lim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true,
                             top, buf, cnt, kk);
  }
  template <int T0, int T1, bool EX>
  __device__ __forceinline__ void tiles_x(const floatx4 (&accf)[16], const intx4 (&acci)[16], int64_t pb0,
                                          const uint2 *cv, int lim, int64_t b_begin, const uint8_t *excl,
                                          int64_t excl_len, const uint32_t *xn_u, const uint32_t *xs_u,
                                          float scale_sq, int d_true, uint64_t *top, uint64_t *buf, int *cnt,
                                          int kk) {
    constexpr bool kColTerms = KIND == BF16_L2 || KIND == FP8_L2 || KIND == FP8_IP || KIND == U8_L2;
    constexpr int NT = T1 - T0;
    const int lane = threadIdx.x & 63, lrow = lane & 15, lgrp = lane >> 4;
    float dist[NT][4];
    bool pass[NT][4];
    int xc[NT];  // U8_L2: the column's xn - 256 sx, exact in int32 (d <= 1264)
    bool any_chunk = false;
#pragma unroll
    for (int u = 0; u < NT; u++) {
      const int t = T0 + u;
      const int64_t col = STR ? goff + (pb0 + t * 16 + lrow) * stride : pb0 + t * 16 + lrow;
      const bool col_ok = t * 16 + lrow <= lim && !(EX && col < excl_len && excl[col] > 0);
      float xa = 0.0f, xb = 0.0f;
      int xcorr = 0;
      if (kColTerms) {
        const uint2 v = cv[t * 16];
        xa = __uint_as_float(v.x);
        xb = __uint_as_float(v.y);
        xcorr = (int)v.x;
      }
      xc[u] = xcorr;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (KIND == I8_IP) {
          pass[u][j] = acci[t][j] >= ithr[j];
        } else if (KIND == U8_L2) {
          pass[u][j] = 2 * acci[t][j] >= ithr[j] + xcorr;
        } else {
          if (KIND == BF16_L2) {
            dist[u][j] = (qa[j] + xa) - 2.0f * accf[t][j];
          } else if (KIND == BF16_IP) {
            dist[u][j] = -accf[t][j];
          } else if (KIND == FP8_L2) {
            dist[u][j] = (qa[j] + xa) - 2.0f * (accf[t][j] * (qb[j] * xb));
          } else {
            dist[u][j] = (qa[j] + xa) - accf[t][j] * (qb[j] * xb);
          }
          pass[u][j] = dist[u][j] <= thr_f[j];
        }
        pass[u][j] = pass[u][j] && col_ok && qok[j];
        any_chunk |= pass[u][j];
      }
    }
    if (__ballot(any_chunk) == 0) return;
#pragma unroll
    for (int u = 0; u < NT; u++) {
      const int t = T0 + u;
      bool any = false;
#pragma unroll
      for (int j = 0; j < 4; j++) any |= pass[u][j];
      if (__ballot(any) == 0) continue;
      const int64_t col = STR ? goff + (pb0 + t * 16 + lrow) * stride : pb0 + t * 16 + lrow;
      const uint64_t gmask = 0xFFFFull << (16 * lgrp);
      const uint64_t below = (1ull << lane) - 1;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (KIND == I8_IP) {
          dist[u][j] = __fmul_rn(-(float)acci[t][j], scale_sq);
        } else if (KIND == U8_L2) {
          // qn + xn - 2 (acc + 128 (sq + sx) + 16384 d), from the staged
          // column term xn - 256 sx (no reload of the column's norm and sum)
          const int64_t v = qc[j] - 256 * qsx[j] - 32768 * (int64_t)d_true + (int64_t)xc[u] -
                            2 * (int64_t)acci[t][j];
          dist[u][j] = __fmul_rn((float)(uint32_t)(v > 0 ? v : 0), scale_sq);
        }
        const uint64_t key = pack_key(dist[u][j], (uint32_t)col);
        // exact key order against the kk-th key (ties: lower id), held in
        // registers (an LDS read here put its latency on every append)
        const bool take = pass[u][j] && dist[u][j] <= thr_f[j] && key < thr_k[j];
        const uint64_t m = __ballot(take);
        if (m == 0) continue;
        if (take) buf[(4 * lgrp + j) * kWCap + cnt_r[j] + __popcll(m & gmask & below)] = key;
        cnt_r[j] += __popcll(m & gmask);
      }
      // room for the next tile's 16 columns per query, else flush
      bool full = false;
#pragma unroll
      for (int j = 0; j < 4; j++) full |= cnt_r[j] > kWCap - 16;
      if (__ballot(full) != 0) {
        if (lrow == 0) {
#pragma unroll
          for (int j = 0; j < 4; j++) cnt[4 * lgrp + j] = cnt_r[j];
        }
        // one copy of the merge (an unrolled loop inlined 16 and spilled)
#pragma unroll 1
        for (int qq = 0; qq < 16; qq++)
          if (cnt[qq] > kWCap - 16) wide_flush<KIND != U8_L2>(top, buf, cnt, qq, kk);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          cnt_r[j] = cnt[4 * lgrp + j];
          const uint64_t th = min(top[(4 * lgrp + j) * kk + kk - 1], bnd_k[j]);
          thr_k[j] = th;
          thr_f[j] = th == kSentinel64 ? INFINITY : key2f((uint32_t)(th >> 32));
          int_thr(j, scale_sq, d_true);
        }
      }
    }
  }

  // All 16 tiles of a step, four at a time.
  __device__ __forceinline__ void step(const floatx4 (&accf)[16], const intx4 (&acci)[16], int64_t pb0,
                                       const uint2 *cv, int lim, int64_t b_begin, const uint8_t *excl,
                                       int64_t excl_len, const uint32_t *xn_u, const uint32_t *xs_u,
                                       float scale_sq, int d_true, uint64_t *top, uint64_t *buf, int *cnt,
                                       int kk) {
    tiles<0, 4>(accf, acci, pb0, cv, lim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true, top, buf, cnt, kk);
    tiles<4, 8>(accf, acci, pb0, cv, lim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true, top, buf, cnt, kk);
    tiles<8, 12>(accf, acci, pb0, cv, lim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true, top, buf, cnt, kk);
    tiles<12, 16>(accf, acci, pb0, cv, lim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true, top, buf, cnt, kk);
  }

  // Last merges, then the wavefront's 16 queries' top kk of this split.
  __device__ __forceinline__ void finish(uint64_t *top, uint64_t *buf, int *cnt, int kk, int64_t q0, int wq,
                                         int qcount, uint64_t *part, int S, int split) {
    const int lane = threadIdx.x & 63, lrow = lane & 15, lgrp = lane >> 4;
    if (lrow == 0) {
#pragma unroll
      for (int j = 0; j < 4; j++) cnt[4 * lgrp + j] = cnt_r[j];
    }
#pragma unroll 1
    for (int qq = 0; qq < 16; qq++)
      if (cnt[qq] > 0) wide_flush<KIND != U8_L2>(top, buf, cnt, qq, kk);
    for (int qq = 0; qq < 16; qq++) {
      if (wq + qq >= qcount) break;
      if (lane < kk) part[((q0 + wq + qq) * S + split) * (int64_t)kk + lane] = top[qq * kk + lane];
    }
  }
};

template <int KIND, bool STR>
__global__ __launch_bounds__(kNT, 1) void gemm_wide_kernel(
    const uint8_t *__restrict__ Q, const uint8_t *__restrict__ X, int64_t nq, int64_t n,
    int row_bytes, int d_true, const float *__restrict__ qnorm, const float *__restrict__ xnorm,
    const uint32_t *__restrict__ qn_u, const uint32_t *__restrict__ qs_u,
    const uint32_t *__restrict__ xn_u, const uint32_t *__restrict__ xs_u, float scale_sq, int kk,
    int64_t split_len, const uint8_t *__restrict__ excl, int64_t excl_len,
    uint64_t *__restrict__ part, int S, const float *__restrict__ qinv,
    const float *__restrict__ xinv, int P, int R, const uint64_t *__restrict__ bound, int bstride) {
  if (!STR) P = R = 1;  // contiguous splits: the row arithmetic folds away
  extern __shared__ __align__(16) unsigned char smem[];
  const WideLayout lay(row_bytes, kk);
  uint8_t *atile = smem + lay.a;
  uint8_t *bst = smem + lay.b;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lgrp = lane >> 4;
  // this wavefront's 16 query rows
  uint64_t *top = reinterpret_cast<uint64_t *>(smem + lay.top) + wave * 16 * kk;
  uint64_t *buf = reinterpret_cast<uint64_t *>(smem + lay.buf) + wave * 16 * kWCap;
  int *cnt = reinterpret_cast<int *>(smem + lay.cnt) + wave * 16;

  int64_t qt;
  int split;
  if (!xcd_tile(nq, kWQ, S, qt, split)) return;
  const int64_t q0 = qt * kWQ;
  const SplitRows sr(n, split, P, R, split_len);
  const int64_t b_begin = sr.b_begin, b_end = sr.b_end;
  const int qcount = (int)min<int64_t>(kWQ, nq - q0);
  if (b_end <= b_begin) {
    empty_split(part, q0, qcount, S, split, kk);
    return;
  }
  const int wq = wave * 16;  // first query row of this wavefront in the tile
  const uint8_t *Xs = X + sr.goff * row_bytes;    // local row 0 of the split
  const int64_t pitch = (int64_t)P * row_bytes;   // bytes between local rows

  const int row_words = row_bytes / 16;
  for (int i = tid; i < kWQ * row_words; i += kNT) {
    const int qi = i / row_words, wi = i - qi * row_words;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (qi < qcount) v = reinterpret_cast<const uint4 *>(Q + (q0 + qi) * row_bytes)[wi];
    if (KIND == U8_L2) {  // u - 128 as int8
      v.x ^= 0x80808080u; v.y ^= 0x80808080u; v.z ^= 0x80808080u; v.w ^= 0x80808080u;
    }
    *reinterpret_cast<uint4 *>(atile + qi * lay.apitch + wi * 16) = v;
  }
  for (int i = lane; i < 16 * kk; i += 64) top[i] = kSentinel64;
  if (lane < 16) cnt[lane] = 0;

  WideSel<KIND, STR, false> sel;
  sel.init(q0, wq, qcount, lgrp, qnorm, qinv, qn_u, qs_u, d_true);
  sel.goff = sr.goff;
  sel.stride = P;
  if (STR && bound) sel.set_bound(bound, bstride, q0, wq, lgrp, scale_sq, d_true);

  const int nks = row_bytes / 64;    // 64-byte k steps of a row
  const int nsl = (nks + 1) >> 1;    // 128-byte slices per step
  const int srow = tid >> 3, sgrp = tid & 7;  // staging: 8 columns x 16 bytes per thread
  const int sswz = wide_swz(sgrp);
  uint4 p0, p1, p2, p3, p4, p5, p6, p7;
  // loads slice SL of the step at B0 (the group past an odd row end re-reads
  // the row's last group; its MFMA is skipped)
#define MVG_LOAD_SLICE(B0, SL)                                                          \
  {                                                                                     \
    const int ko = min((SL) * 128 + sgrp * 16, row_bytes - 16);                          \
    const uint8_t *src = Xs + ko;                                                        \
    const int64_t lastc = b_end - 1;                                                     \
    const int64_t c0 = (B0) + srow;                                                      \
    p0 = *reinterpret_cast<const uint4 *>(src + min(c0, lastc) * pitch);                 \
    p1 = *reinterpret_cast<const uint4 *>(src + min(c0 + 32, lastc) * pitch);            \
    p2 = *reinterpret_cast<const uint4 *>(src + min(c0 + 64, lastc) * pitch);            \
    p3 = *reinterpret_cast<const uint4 *>(src + min(c0 + 96, lastc) * pitch);            \
    p4 = *reinterpret_cast<const uint4 *>(src + min(c0 + 128, lastc) * pitch);           \
    p5 = *reinterpret_cast<const uint4 *>(src + min(c0 + 160, lastc) * pitch);           \
    p6 = *reinterpret_cast<const uint4 *>(src + min(c0 + 192, lastc) * pitch);           \
    p7 = *reinterpret_cast<const uint4 *>(src + min(c0 + 224, lastc) * pitch);           \
  }
#define MVG_STORE_SLICE(SB)                                                              \
  {                                                                                      \
    if (KIND == U8_L2) {                                                                 \
      p0.x ^= 0x80808080u; p0.y ^= 0x80808080u; p0.z ^= 0x80808080u; p0.w ^= 0x80808080u; \
      p1.x ^= 0x80808080u; p1.y ^= 0x80808080u; p1.z ^= 0x80808080u; p1.w ^= 0x80808080u; \
      p2.x ^= 0x80808080u; p2.y ^= 0x80808080u; p2.z ^= 0x80808080u; p2.w ^= 0x80808080u; \
      p3.x ^= 0x80808080u; p3.y ^= 0x80808080u; p3.z ^= 0x80808080u; p3.w ^= 0x80808080u; \
      p4.x ^= 0x80808080u; p4.y ^= 0x80808080u; p4.z ^= 0x80808080u; p4.w ^= 0x80808080u; \
      p5.x ^= 0x80808080u; p5.y ^= 0x80808080u; p5.z ^= 0x80808080u; p5.w ^= 0x80808080u; \
      p6.x ^= 0x80808080u; p6.y ^= 0x80808080u; p6.z ^= 0x80808080u; p6.w ^= 0x80808080u; \
      p7.x ^= 0x80808080u; p7.y ^= 0x80808080u; p7.z ^= 0x80808080u; p7.w ^= 0x80808080u; \
    }                                                                                    \
    uint8_t *plane = bst + (SB) * kBSlice + sgrp * kBPlane;                              \
    *reinterpret_cast<uint4 *>(plane + ((srow ^ sswz) * 16)) = p0;                       \
    *reinterpret_cast<uint4 *>(plane + (((srow + 32) ^ sswz) * 16)) = p1;                \
    *reinterpret_cast<uint4 *>(plane + (((srow + 64) ^ sswz) * 16)) = p2;                \
    *reinterpret_cast<uint4 *>(plane + (((srow + 96) ^ sswz) * 16)) = p3;                \
    *reinterpret_cast<uint4 *>(plane + (((srow + 128) ^ sswz) * 16)) = p4;               \
    *reinterpret_cast<uint4 *>(plane + (((srow + 160) ^ sswz) * 16)) = p5;               \
    *reinterpret_cast<uint4 *>(plane + (((srow + 192) ^ sswz) * 16)) = p6;               \
    *reinterpret_cast<uint4 *>(plane + (((srow + 224) ^ sswz) * 16)) = p7;               \
  }

  // Per-column terms of a step (norm, scale or code-sum correction), loaded
  // with the step's first slice (one column per thread) and kept in LDS until
  // the step's epilogue. Three buffers in turn: with one slice per step the
  // terms of step s + 2 are written while a slower wavefront may still read
  // those of step s.
  constexpr bool kColTerms = KIND == BF16_L2 || KIND == FP8_L2 || KIND == FP8_IP || KIND == U8_L2;
  uint2 *colv = reinterpret_cast<uint2 *>(smem + lay.colv);
  uint2 cterm = make_uint2(0, 0);
#define MVG_LOAD_TERMS(B0)                                                                   \
  if (kColTerms) {                                                                           \
    const int64_t c = (B0) + tid;                                                            \
    cterm = make_uint2(0, 0);                                                                \
    if (c < b_end) {                                                                         \
      const int64_t g = sr.goff + c * P;                                                     \
      if (KIND == U8_L2)                                                                     \
        cterm.x = (uint32_t)((int)xn_u[g] - 256 * ((int)xs_u[g] - 128 * d_true));           \
      else                                                                                   \
        cterm.x = __float_as_uint(xnorm[g]);                                                 \
      if (KIND == FP8_L2 || KIND == FP8_IP) cterm.y = __float_as_uint(xinv[g]);              \
    }                                                                                        \
  }
  int sp = 0;  // buffer of the current step's terms (0..2)
  MVG_LOAD_SLICE(b_begin, 0)
  MVG_LOAD_TERMS(b_begin)
  MVG_STORE_SLICE(0)
  if (kColTerms) colv[tid] = cterm;
  __syncthreads();
  int sb = 0;
  const uint8_t *arow = atile + (wq + lrow) * lay.apitch + lgrp * 16;
  floatx4 accf[16];
  intx4 acci[16];
#pragma unroll
  for (int t = 0; t < 16; t++) {
    accf[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    acci[t] = intx4{0, 0, 0, 0};
  }
  for (int64_t b0 = b_begin; b0 < b_end; b0 += kStepCols) {
    for (int sl = 0; sl < nsl; sl++) {
      int nsl1 = sl + 1;
      int64_t nb0 = b0;
      if (nsl1 == nsl) {
        nsl1 = 0;
        nb0 = b0 + kStepCols;
      }
      const bool more = nb0 < b_end;
      if (more) MVG_LOAD_SLICE(nb0, nsl1)
      if (more && nsl1 == 0) MVG_LOAD_TERMS(nb0)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int ks = 2 * sl + h;
      if (ks < nks) {
        const uint4 a = *reinterpret_cast<const uint4 *>(arow + ks * 64);
        const int g = 4 * h + lgrp;
        const uint8_t *plane = bst + sb * kBSlice + g * kBPlane + ((lrow ^ wide_swz(g)) * 16);
        // the 16 fragment reads are issued together, ahead of their MFMAs
        uint4 bfr[16];
#pragma unroll
        for (int t = 0; t < 16; t++) bfr[t] = *reinterpret_cast<const uint4 *>(plane + t * 256);
#pragma unroll
        for (int t = 0; t < 16; t++) {
          const uint4 b = bfr[t];
          if (KIND == BF16_L2 || KIND == BF16_IP) {
            accf[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                              __builtin_bit_cast(bf16x8, b), accf[t], 0, 0, 0);
          } else if (KIND == FP8_L2 || KIND == FP8_IP) {
            const long alo = (long)(((uint64_t)a.y << 32) | a.x), ahi = (long)(((uint64_t)a.w << 32) | a.z);
            const long blo = (long)(((uint64_t)b.y << 32) | b.x), bhi = (long)(((uint64_t)b.w << 32) | b.z);
            accf[t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(alo, blo, accf[t], 0, 0, 0);
            accf[t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(ahi, bhi, accf[t], 0, 0, 0);
          } else {  // int8 codes (U8_L2 rebased when staged)
            acci[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(intx4, a),
                                                            __builtin_bit_cast(intx4, b), acci[t], 0, 0, 0);
          }
        }
        // schedule: the A and 16 B fragment reads first, then the MFMAs
        __builtin_amdgcn_sched_group_barrier(0x100, 17, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, KIND == FP8_L2 || KIND == FP8_IP ? 32 : 16, 0);
      }
    }
      if (more) MVG_STORE_SLICE(sb ^ 1)
      if (kColTerms && more && nsl1 == 0) colv[(sp == 2 ? 0 : sp + 1) * kStepCols + tid] = cterm;
      __syncthreads();
      sb ^= 1;
    }
    {
      const uint2 *cv = colv + sp * kStepCols + lrow;
      sp = sp == 2 ? 0 : sp + 1;
      sel.step(accf, acci, b0, cv, (int)min<int64_t>(kStepCols, b_end - b0) - 1, b_begin, excl,
               excl_len, xn_u, xs_u, scale_sq, d_true, top, buf, cnt, kk);
#pragma unroll
      for (int t = 0; t < 16; t++) {
        accf[t] = floatx4{0.f, 0.f, 0.f, 0.f};
        acci[t] = intx4{0, 0, 0, 0};
      }
    }
  }
#undef MVG_LOAD_TERMS
#undef MVG_LOAD_SLICE
#undef MVG_STORE_SLICE
  sel.finish(top, buf, cnt, kk, q0, wq, qcount, part, S, split);
}

// Wide form for row sizes fixed at compile time: NKS 64-byte k steps per row
// (the 1M x 768 layouts: int8 and FP8 rows NKS 12, BF16 rows NKS 24). Each
// lane holds its query row's fragments for the whole row in registers, so
// LDS keeps only two staged base slices and the selection state, and a
// 64-query tile of 1,536-byte BF16 rows fits. A step of 256 base rows is
// unrolled; slice loads run DEPTH slices ahead through a ring of registers,
// and the barrier after each slice waits for the LDS writes only
// (__syncthreads' fence would also wait for the loads in flight). FP8 runs
// the 16x16x128 f8f6f4 MFMA with unit scales: one instruction per 128 bytes
// of k. Its lane group g holds bytes [16 g, 16 g + 16) and [64 + 16 g,
// 80 + 16 g) of the slice for both operands (any k order common to A and B
// gives the same dot product; these are the planes the int8 and BF16 reads
// take, without the bank conflicts of planes 2 g and 2 g + 1). Base slices
// load through a buffer resource over the split's rows: 32-bit offsets, and
// rows past the split's end read as zeros (the epilogue drops them).
struct WideRegLayout {
  int b, colv, top, buf, cnt, total;
  __host__ __device__ explicit WideRegLayout(int kk) {
    b = 0;
    colv = b + 2 * kBSlice;
    top = colv + 3 * kStepCols * 8;
    buf = top + 8 * kWQ * kk;
    cnt = buf + 8 * kWQ * kWCap;
    total = cnt + 4 * kWQ;
  }
};

typedef int intx8 __attribute__((ext_vector_type(8)));

// f(I) for I = 0 .. N-1 with I a compile-time constant (std::integral_constant).
template <typename F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

template <int KIND, int NKS, bool STR>
__global__ __launch_bounds__(kNT, 1) void gemm_wide_reg_kernel(
    const uint8_t *__restrict__ Q, const uint8_t *__restrict__ X, int64_t nq, int64_t n,
    int d_true, const float *__restrict__ qnorm, const float *__restrict__ xnorm,
    const uint32_t *__restrict__ qn_u, const uint32_t *__restrict__ qs_u,
    const uint32_t *__restrict__ xn_u, const uint32_t *__restrict__ xs_u, float scale_sq, int kk,
    int64_t split_len, const uint8_t *__restrict__ excl, int64_t excl_len,
    uint64_t *__restrict__ part, int S, const float *__restrict__ qinv,
    const float *__restrict__ xinv, int P, int R, const uint64_t *__restrict__ bound, int bstride) {
  if (!STR) P = R = 1;  // contiguous splits: the row arithmetic folds away
  static_assert(NKS % 2 == 0, "whole 128-byte slices");
  constexpr int NSL = NKS / 2;                                    // slices per row
  constexpr bool F8 = KIND == FP8_L2 || KIND == FP8_IP;
  // ring of slice loads: three deep where the registers allow (the FP8
  // operands of the 128-k MFMA take eight registers each and spill at three)
  constexpr int DEPTH = (NKS <= 12 && NSL % 3 == 0 && !F8) ? 3 : 2;
  static_assert(NSL % DEPTH == 0, "a step's slices in whole turns of the ring");
  constexpr int ROW = NKS * 64;                                   // row bytes
  constexpr bool kColTerms = KIND == BF16_L2 || KIND == FP8_L2 || KIND == FP8_IP || KIND == U8_L2;
  // PIPE: a step's selection runs in chunks during the next step's slices,
  // beside their MFMAs (BF16 rows: 12 slices of MFMA work per step to hide
  // it behind; measured slower for the 6-slice int8 / FP8 steps)
  constexpr bool PIPE = NSL >= 12;
  extern __shared__ __align__(16) unsigned char smem[];
  const WideRegLayout lay(kk);
  uint8_t *bst = smem + lay.b;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lgrp = lane >> 4;
  uint64_t *top = reinterpret_cast<uint64_t *>(smem + lay.top) + wave * 16 * kk;
  uint64_t *buf = reinterpret_cast<uint64_t *>(smem + lay.buf) + wave * 16 * kWCap;
  int *cnt = reinterpret_cast<int *>(smem + lay.cnt) + wave * 16;

  int64_t qt;
  int split;
  if (!xcd_tile(nq, kWQ, S, qt, split)) return;
  const int64_t q0 = qt * kWQ;
  const SplitRows sr(n, split, P, R, split_len);
  const int64_t b_begin = sr.b_begin, b_end = sr.b_end;
  const int qcount = (int)min<int64_t>(kWQ, nq - q0);
  if (b_end <= b_begin) {
    empty_split(part, q0, qcount, S, split, kk);
    return;
  }
  const int wq = wave * 16;
  const int pitch = P * ROW;  // bytes between local rows (the host keeps a split's span below 2 GiB)

  // the lane's query row (wq + lrow): fragment h of slice sl at areg[2 sl + h]
  // (FP8: the 32 bytes of slice sl at a8[sl])
  uint4 areg[F8 ? 1 : NKS];
  intx8 a8[F8 ? NSL : 1];
  {
    const int qi = wq + lrow;
    const bool ok = qi < qcount;
    const uint8_t *qr = Q + (q0 + (ok ? qi : 0)) * ROW;
    if constexpr (F8) {
#pragma unroll
      for (int sl = 0; sl < NSL; sl++) {
        const intx4 lo = ok ? *reinterpret_cast<const intx4 *>(qr + sl * 128 + lgrp * 16) : intx4{0, 0, 0, 0};
        const intx4 hi = ok ? *reinterpret_cast<const intx4 *>(qr + sl * 128 + 64 + lgrp * 16) : intx4{0, 0, 0, 0};
        a8[sl] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
      }
    } else {
#pragma unroll
      for (int i = 0; i < NKS; i++) {
        uint4 v = ok ? *reinterpret_cast<const uint4 *>(qr + i * 64 + lgrp * 16) : make_uint4(0, 0, 0, 0);
        if (KIND == U8_L2) {  // u - 128 as int8
          v.x ^= 0x80808080u; v.y ^= 0x80808080u; v.z ^= 0x80808080u; v.w ^= 0x80808080u;
        }
        areg[i] = v;
      }
    }
  }
  for (int i = lane; i < 16 * kk; i += 64) top[i] = kSentinel64;
  if (lane < 16) cnt[lane] = 0;
  WideSel<KIND, STR, true> sel;
  sel.init(q0, wq, qcount, lgrp, qnorm, qinv, qn_u, qs_u, d_true);
  sel.goff = sr.goff;
  sel.stride = P;
  if (STR && bound) sel.set_bound(bound, bstride, q0, wq, lgrp, scale_sq, d_true);

  const int srow = tid >> 3, sgrp = tid & 7;  // staging: 8 columns x 16 bytes per thread
  const int sswz = wide_swz(sgrp);
  uint4 pr[DEPTH][8];
  // the split's local rows [b_begin, b_end) as one buffer that ends with the
  // last one's bytes; flags word: raw byte addressing, as for gfx9 buffers.
  // Rows past the split's end read as zeros: strided splits put the whole
  // offset in the vector operand, which the range check covers.
  const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(X + (sr.goff + b_begin * P) * ROW), (short)0,
      (int)(((b_end - b_begin - 1) * P + 1) * ROW), 0x00020000);
  auto load_slice = [&](uint4(&p)[8], int64_t B0, int SL) __attribute__((always_inline)) {
    const int vo = (int)(B0 - b_begin + srow) * pitch + SL * 128 + sgrp * 16;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if constexpr (STR)
        p[i] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xrs, vo + i * 32 * pitch, 0, 0));
      else
        p[i] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xrs, vo, i * 32 * ROW, 0));
    }
  };
  auto store_slice = [&](uint4(&p)[8], int SB) __attribute__((always_inline)) {
    uint8_t *plane = bst + SB * kBSlice + sgrp * kBPlane;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint4 v = p[i];
      if (KIND == U8_L2) {
        v.x ^= 0x80808080u; v.y ^= 0x80808080u; v.z ^= 0x80808080u; v.w ^= 0x80808080u;
      }
      *reinterpret_cast<uint4 *>(plane + (((srow + 32 * i) ^ sswz) * 16)) = v;
    }
  };
  uint2 *colv = reinterpret_cast<uint2 *>(smem + lay.colv);
  uint2 cterm = make_uint2(0, 0);
  // column terms of the step at B0 (columns past b_end read the last row's:
  // the epilogue drops those columns)
  auto load_terms = [&](int64_t B0) __attribute__((always_inline)) {
    if (kColTerms) {
      const int64_t c = sr.goff + min(B0 + tid, b_end - 1) * P;
      if (KIND == U8_L2)
        cterm.x = (uint32_t)((int)xn_u[c] - 256 * ((int)xs_u[c] - 128 * d_true));
      else
        cterm.x = __float_as_uint(xnorm[c]);
      if (F8) cterm.y = __float_as_uint(xinv[c]);
    }
  };

  // slices 0 .. DEPTH - 1 of the first step in flight, slice 0 staged
  load_terms(b_begin);
#pragma unroll
  for (int r = 0; r < DEPTH; r++) load_slice(pr[r], b_begin, r);
  store_slice(pr[0], 0);
  if (kColTerms) colv[tid] = cterm;
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");

  floatx4 accf[16];
  intx4 acci[16];
  // the previous step's accumulators, selected in chunks during this step's
  // slices (its MFMAs run meanwhile); plim -1: no previous step yet
  floatx4 paccf[16];
  intx4 pacci[16];
#pragma unroll
  for (int t = 0; t < 16; t++) {
    accf[t] = paccf[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    acci[t] = pacci[t] = intx4{0, 0, 0, 0};
  }
  int64_t pb0 = b_begin;
  int plim = -1;
  const uint2 *pcv = colv + lrow;
  int sb = 0, sp = 0;
  for (int64_t b0 = b_begin; b0 < b_end; b0 += kStepCols) {
    static_for<NSL>([&](auto slc) __attribute__((always_inline)) {
      constexpr int sl = decltype(slc)::value;
      // Slice sl + DEPTH (of this step or the next) into the ring slot that
      // slice sl leaves. Past the split's end the loads re-read its last row
      // and nothing uses them: no branch, so that a step stays one block (at
      // a branch per slice the compiler sank the step's MFMAs below its last
      // barrier and held every slice's fragments at once).
      {
        const int tsl = sl + DEPTH < NSL ? sl + DEPTH : sl + DEPTH - NSL;
        const int64_t tb0 = min(sl + DEPTH < NSL ? b0 : b0 + kStepCols, b_end);
        load_slice(pr[sl % DEPTH], tb0, tsl);
        if (tsl == 0) load_terms(tb0);
      }
      const uint8_t *slice = bst + sb * kBSlice;
      if constexpr (F8) {
        const int g0 = lgrp, g1 = lgrp + 4;
        const uint8_t *p0 = slice + g0 * kBPlane + ((lrow ^ wide_swz(g0)) * 16);
        const uint8_t *p1 = slice + g1 * kBPlane + ((lrow ^ wide_swz(g1)) * 16);
#pragma unroll
        for (int quarter = 0; quarter < 4; quarter++) {
          intx8 bv[4];
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const intx4 lo = *reinterpret_cast<const intx4 *>(p0 + (4 * quarter + t) * 256);
            const intx4 hi = *reinterpret_cast<const intx4 *>(p1 + (4 * quarter + t) * 256);
            bv[t] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
          }
#pragma unroll
          for (int t = 0; t < 4; t++)
            accf[4 * quarter + t] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                a8[sl], bv[t], accf[4 * quarter + t], 0, 0, 0, 127, 0, 127);
          __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
        }
      } else {
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const uint4 a = areg[2 * sl + h];
          const int g = 4 * h + lgrp;
          const uint8_t *plane = slice + g * kBPlane + ((lrow ^ wide_swz(g)) * 16);
#pragma unroll
          for (int half = 0; half < 2; half++) {
            uint4 bfr[8];
#pragma unroll
            for (int t = 0; t < 8; t++) bfr[t] = *reinterpret_cast<const uint4 *>(plane + (8 * half + t) * 256);
#pragma unroll
            for (int t = 0; t < 8; t++) {
              const int tt = 8 * half + t;
              if (KIND == BF16_L2 || KIND == BF16_IP)
                accf[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                                   __builtin_bit_cast(bf16x8, bfr[t]),
                                                                   accf[tt], 0, 0, 0);
              else
                acci[tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(intx4, a),
                                                                 __builtin_bit_cast(intx4, bfr[t]),
                                                                 acci[tt], 0, 0, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
          }
        }
      }
      // tiles [16 sl / NSL, 16 (sl + 1) / NSL) of the previous step
      if constexpr (PIPE)
        sel.template tiles<(16 * sl) / NSL, (16 * (sl + 1)) / NSL>(
            paccf, pacci, pb0, pcv, plim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true, top, buf,
            cnt, kk);
      // stage slice sl + 1 (the next step's slice 0 after the last; past the
      // split's end a buffer nobody reads)
      store_slice(pr[(sl + 1) % DEPTH], sb ^ 1);
      if (kColTerms && sl + 1 == NSL) colv[(sp == 2 ? 0 : sp + 1) * kStepCols + tid] = cterm;
      // the slice's MFMAs issue before its barrier (the scheduler had moved
      // the FP8 ones below the step's last barrier, every fragment live; for
      // those the accumulators pass through an empty volatile asm here)
      if constexpr (F8) {
#pragma unroll
        for (int t = 0; t < 16; t += 4)
          asm volatile("" : "+a"(accf[t]), "+a"(accf[t + 1]), "+a"(accf[t + 2]), "+a"(accf[t + 3]));
      }
      __builtin_amdgcn_sched_barrier(0);
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      sb ^= 1;
    });
    pcv = colv + sp * kStepCols + lrow;
    sp = sp == 2 ? 0 : sp + 1;
    pb0 = b0;
    plim = (int)min<int64_t>(kStepCols, b_end - b0) - 1;
    if constexpr (!PIPE)
      sel.step(accf, acci, pb0, pcv, plim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true, top, buf,
               cnt, kk);
#pragma unroll
    for (int t = 0; t < 16; t++) {
      if constexpr (PIPE) {
        paccf[t] = accf[t];
        pacci[t] = acci[t];
      }
      accf[t] = floatx4{0.f, 0.f, 0.f, 0.f};
      acci[t] = intx4{0, 0, 0, 0};
    }
  }
  if constexpr (PIPE)
    sel.step(paccf, pacci, pb0, pcv, plim, b_begin, excl, excl_len, xn_u, xs_u, scale_sq, d_true, top, buf,
             cnt, kk);
  sel.finish(top, buf, cnt, kk, q0, wq, qcount, part, S, split);
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

// FP32 rows -> FP8 E4M3 rows padded to row_bytes (zero bytes), each scaled
// by the power of two 2^e that puts its largest magnitude in [224, 448];
// inv[row] = 2^-e, and norms (nullable) = FP32 squared norm of the
// dequantized row v~ (half_diff false) or (|v~|^2 - |v|^2) / 2 (half_diff
// true, summed as (v~ - v)(v~ + v) per element). One wavefront per row.
__global__ void to_fp8_kernel(const float *__restrict__ x, int64_t n, int d, int row_bytes,
                              uint8_t *__restrict__ out, float *__restrict__ inv,
                              float *__restrict__ norms, bool half_diff) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  const float *xr = x + row * d;
  float mx = 0.0f;
  for (int j = lane; j < d; j += 64) mx = fmaxf(mx, fabsf(xr[j]));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  int e = 0;
  if (mx > 0.0f) {
    int ex;
    (void)frexpf(mx, &ex);  // mx in [2^(ex-1), 2^ex)
    e = 9 - ex;             // mx * 2^e in [256, 512) ... one step down if >= 448
    if (ldexpf(mx, e) > 448.0f) e -= 1;
  }
  const float s = ldexpf(1.0f, e), is = ldexpf(1.0f, -e);
  float acc = 0.0f;
  uint8_t *orow = out + row * row_bytes;
  for (int j0 = 4 * lane; j0 < row_bytes; j0 += 256) {
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = j0 + c < d ? xr[j0 + c] * s : 0.0f;
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
    const float r[4] = {__builtin_amdgcn_cvt_f32_fp8(w, 0) * is, __builtin_amdgcn_cvt_f32_fp8(w, 1) * is,
                        __builtin_amdgcn_cvt_f32_fp8(w, 2) * is, __builtin_amdgcn_cvt_f32_fp8(w, 3) * is};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (half_diff) {
        const float xv = v[c] * is;  // the original value (exact: power-of-two scaling)
        acc = fma_(r[c] - xv, r[c] + xv, acc);
      } else {
        acc = fma_(r[c], r[c], acc);
      }
    }
    *reinterpret_cast<int *>(orow + j0) = w;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) {
    inv[row] = is;
    if (norms) norms[row] = half_diff ? 0.5f * acc : acc;
  }
}

}  // namespace

void to_fp8(const float *x, int64_t n, int d, uint8_t *out, float *inv, float *norms,
            hipStream_t s, bool half_diff) {
  if (n == 0) return;
  hipLaunchKernelGGL(to_fp8_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, s, x, n, d,
                     gemm_row_bytes(FP8_L2, d), out, inv, norms, half_diff);
  MVG_LAUNCH_CHECK("to_fp8_kernel");
}

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

namespace {

// Strided splits keep each split's exact top kper (sorted, the sentinel past
// a split's last row) among its keys up to the query's bound B (the sentinel
// without one). Let M be min(B, the smallest kper-th key over the splits
// whose list is full). Every split's list holds all of that split's keys up
// to M -- a full list ends at or above M, and a list that is not full holds
// all its split's keys up to B -- so the keys <= M in the union are exactly
// the base's keys <= M. If there are at least kk of them, the union's kk
// smallest are the base's; otherwise (a split held more than kper of the
// query's kk nearest rows, or fewer than kk keys lie below B) flag[q] = 1 and
// the query is searched again. One wavefront per query.
__global__ void gemm_certify_kernel(const uint64_t *__restrict__ part, int64_t nq, int S, int kper, int kk,
                                    const uint64_t *__restrict__ bound, int bstride,
                                    uint8_t *__restrict__ flag) {
  const int lane = threadIdx.x & 63;
  const int64_t q = (int64_t)blockIdx.x * (kNT / 64) + (threadIdx.x >> 6);
  if (q >= nq) return;
  const uint64_t *row = part + q * (int64_t)S * kper;
  uint64_t m = kSentinel64;
  for (int s = lane; s < S; s += 64) m = min(m, row[(int64_t)s * kper + kper - 1]);
#pragma unroll
  for (int j = 1; j < 64; j <<= 1) m = min(m, xor_lane64(m, j));
  if (bound) m = min(m, bound[q * bstride]);  // no key above it was kept
  int c = 0;
  if (m != kSentinel64) {
    for (int s = lane; s < S; s += 64) {
      const uint64_t *l = row + (int64_t)s * kper;
      int lo = 0, hi = kper;  // first index whose key exceeds m
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (l[mid] <= m) lo = mid + 1;
        else hi = mid;
      }
      c += lo;
    }
#pragma unroll
    for (int j = 1; j < 64; j <<= 1) c += __shfl_xor(c, j);
  }
  if (lane == 0) flag[q] = m != kSentinel64 && c < kk;
}

// A flagged query's candidate row (ncand keys) becomes its exact top kk from
// the 16-query search (redo, kk keys per query) followed by sentinels.
__global__ void gemm_fixup_kernel(uint64_t *__restrict__ part, int64_t ncand,
                                  const uint64_t *__restrict__ redo, int kk,
                                  const uint8_t *__restrict__ flag) {
  const int64_t q = blockIdx.x;
  if (!flag[q]) return;
  uint64_t *dst = part + q * ncand;
  for (int64_t i = threadIdx.x; i < ncand; i += kNT) dst[i] = i < kk ? redo[q * kk + i] : kSentinel64;
}

}  // namespace

GemmPlan gemm_plan(int kind, int64_t nq, int64_t n, int d, int kk) {
  GemmPlan pl{};
  const int row_bytes = gemm_row_bytes(kind, d);
  const GemmLayout glay(row_bytes, kk);
  // Above 64 candidates the wide forms run on P strided splits (rows r, r +
  // P, ...): each keeps its exact top kw = 64, the union is certified per
  // query, and the 16-query form searches the rejected queries again (its
  // layout must fit). A sample pass first (rows 0, Ps, 2 Ps, ...) takes each
  // query's kks-th key over the sample, about its (kks Ps)-th = (1.25 kk +
  // 16 Ps)-th over the base; at least kk base keys lie below it unless more
  // than kks of the kk nearest rows fall in the sample (for kk = 200: 32
  // against 12.5 expected, 5.7 standard deviations). That key bounds the
  // strided lists: with P >= kks Ps / 40 each split receives about 40 keys
  // instead of filling and merging a top 64, and with kk / P <= 24 of the kk
  // nearest rows per split, more than 64 in one is far in the tail unless
  // the rows are laid out with period P.
  const bool strided = kk > kWMaxK;
  const int kw = strided ? kWMaxK : kk;
  const int Ps = std::max(16, ceil_div(5 * kk, 4 * 48));
  const int kks = std::min(kWMaxK, ceil_div(5 * kk, 4 * Ps) + 16);
  const int P = strided ? std::max({2, ceil_div(kk, 24), ceil_div(kks * Ps, 40)}) : 1;
  const WideLayout wlay(row_bytes, kw);
  // the register-resident query tile for the 768-byte int8 / FP8 rows and the
  // 1,536-byte BF16 rows (d = 768)
  const bool reg12 = row_bytes == 768 && (kind == I8_IP || kind == U8_L2 || kind == FP8_L2 || kind == FP8_IP);
  const bool reg24 = row_bytes == 1536 && (kind == BF16_L2 || kind == BF16_IP);
  const WideRegLayout rlay(kw);
  const bool reg = (reg12 || reg24) && rlay.total <= 160 * 1024;
  if (nq > kQT && (reg || wlay.total <= 160 * 1024) &&
      (!strided || (glay.total <= 150 * 1024 && n >= (int64_t)P * kStepCols))) {
    const int qtiles = ceil_div(nq, kWQ);
    const int64_t nloc = ceil_div(n, P);  // local rows of the longest residue
    // Chunks per residue (the splits, when P = 1): one workgroup per CU at a
    // time, so a CU streams ceil(qtiles * P * C / cus) / (P * C) of the base;
    // every split costs a selection warm-up and kw more candidates to merge,
    // so take the fewest chunks within 10% of the best streaming balance.
    const int cus = std::max(1, device_cus());
    const int64_t cmax = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(nloc, kStepCols)));
    double best = 1e30;
    for (int64_t c = 1; c <= cmax; c++) best = std::min(best, (double)ceil_div(qtiles * P * c, cus) / c);
    int64_t C = 1;
    while (C < cmax && (double)ceil_div(qtiles * P * C, cus) / C > 1.1 * best) C++;
    // the register form addresses a split's rows with 32-bit buffer offsets:
    // its span, a step past its end included, stays below 2 GiB
    const int64_t prow = (int64_t)P * row_bytes;
    if (reg) C = std::max<int64_t>(C, ceil_div(nloc * prow, (int64_t(1) << 31) - 512 * prow));
    pl.split_len = ((ceil_div(nloc, C) + kStepCols - 1) / kStepCols) * kStepCols;
    C = ceil_div(nloc, pl.split_len);
    pl.wide = true;
    pl.reg = reg;
    pl.strided = strided;
    pl.P = P;
    pl.kw = kw;
    pl.S = C * P;
    pl.ncand = pl.S * kw;
    // [nq][S][kw] split lists, then (strided) the 16-query form's [nq][kk],
    // one flag byte per query (rounded up to 8 bytes) and the sample pass's
    // lists and keys
    pl.bytes = sizeof(uint64_t) * (size_t)nq * pl.S * kw +
               (strided ? sizeof(uint64_t) * (size_t)nq * kk + (size_t)((nq + 7) & ~int64_t(7)) : 0);
    // the sample pass (without enough rows for one, the lists start unbounded)
    if (strided && n >= (int64_t)Ps * kStepCols) {
      pl.Ps = Ps;
      pl.kks = kks;
      const int64_t ns = ceil_div(n, Ps);
      const int64_t csmax = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(ns, kStepCols)));
      double bs = 1e30;
      for (int64_t c = 1; c <= csmax; c++) bs = std::min(bs, (double)ceil_div(qtiles * c, cus) / c);
      int64_t Cs = 1;
      while (Cs < csmax && (double)ceil_div(qtiles * Cs, cus) / Cs > 1.1 * bs) Cs++;
      const int64_t psrow = (int64_t)Ps * row_bytes;
      if (reg) Cs = std::max<int64_t>(Cs, ceil_div(ns * psrow, (int64_t(1) << 31) - 512 * psrow));
      pl.split_len_s = ((ceil_div(ns, Cs) + kStepCols - 1) / kStepCols) * kStepCols;
      pl.Ss = ceil_div(ns, pl.split_len_s);
      pl.bytes += sizeof(uint64_t) * (size_t)nq * (pl.Ss + 1) * pl.kks;
    }
    return pl;
  }
  if (glay.total > 150 * 1024) throw Error(kInvalid, "GEMM top-k workspace exceeds LDS");
  const int qtiles = ceil_div(nq, kQT);
  // Splits of the base per query tile. Up to 64 candidates per query, a grid
  // of ~2048 workgroups hides the row loads best. Above that every split
  // fills its own top kk and pays its own flushes (about kk (1 + ln(split /
  // kk)) appends per query), so take the fewest splits within 10% of the best
  // balance over the CUs' resident workgroups (LDS-bound).
  int64_t S = std::max<int64_t>(1, std::min<int64_t>(ceil_div(2048, qtiles), ceil_div(n, kStepCols)));
  if (kk > kWMaxK) {
    const int cus = std::max(1, device_cus());
    const int per_cu = std::max(1, std::min(8, (160 * 1024) / std::max(1, glay.total)));
    const int64_t slots = (int64_t)cus * per_cu;
    const int64_t smax = std::max<int64_t>(1, std::min<int64_t>(256, ceil_div(n, kStepCols)));
    double best = 1e30;
    for (int64_t c = 1; c <= smax; c++) best = std::min(best, (double)ceil_div(qtiles * c, slots) / c);
    S = 1;
    while (S < smax && (double)ceil_div(qtiles * S, slots) / S > 1.1 * best) S++;
  }
  pl.split_len = ((ceil_div(n, S) + kStepCols - 1) / kStepCols) * kStepCols;
  pl.S = ceil_div(n, pl.split_len);
  pl.P = 1;
  pl.kw = kk;
  pl.ncand = pl.S * kk;
  pl.bytes = sizeof(uint64_t) * (size_t)nq * pl.S * kk;
  return pl;
}

void gemm_topk(int kind, const uint8_t *Q, const uint8_t *X, int64_t nq, int64_t n, int d,
               const float *qnorm, const float *xnorm, const uint32_t *qn_u, const uint32_t *qs_u,
               const uint32_t *xn_u, const uint32_t *xs_u, float scale_sq, int kk, const uint8_t *excl,
               int64_t excl_len, DevBuf &part, int64_t &ncand, hipStream_t s, const float *qinv,
               const float *xinv) {
  const GemmPlan pl = gemm_plan(kind, nq, n, d, kk);
  const int row_bytes = gemm_row_bytes(kind, d);
  const GemmLayout glay(row_bytes, kk);
  ncand = pl.ncand;
  part.reserve(pl.bytes);
  uint64_t *pk = part.as<uint64_t>();
  const void *fn16;
  switch (kind) {
    case BF16_L2: fn16 = (const void *)gemm_topk_kernel<BF16_L2>; break;
    case BF16_IP: fn16 = (const void *)gemm_topk_kernel<BF16_IP>; break;
    case I8_IP: fn16 = (const void *)gemm_topk_kernel<I8_IP>; break;
    case FP8_L2: fn16 = (const void *)gemm_topk_kernel<FP8_L2>; break;
    case FP8_IP: fn16 = (const void *)gemm_topk_kernel<FP8_IP>; break;
    default: fn16 = (const void *)gemm_topk_kernel<U8_L2>; break;
  }
  // the 16-query form over S16 contiguous splits into out[nq][S16][kk]
  // (only: the tiles holding a flagged query)
  auto run16 = [&](int64_t S16, int64_t split_len16, uint64_t *out, const uint8_t *only) {
    if (glay.total > 64 * 1024)
      MVG_HIP(hipFuncSetAttribute(fn16, hipFuncAttributeMaxDynamicSharedMemorySize, glay.total));
    const dim3 grid(xcd_grid(ceil_div(nq, kQT), S16));
#define MVG_GEMM(K)                                                                                \
  hipLaunchKernelGGL(gemm_topk_kernel<K>, grid, dim3(kNT), glay.total, s, Q, X, nq, n, row_bytes, d, \
                     qnorm, xnorm, qn_u, qs_u, xn_u, xs_u, scale_sq, kk, split_len16, excl,       \
                     excl_len, out, (int)S16, qinv, xinv, only);
    switch (kind) {
      case BF16_L2: MVG_GEMM(BF16_L2) break;
      case BF16_IP: MVG_GEMM(BF16_IP) break;
      case I8_IP: MVG_GEMM(I8_IP) break;
      case FP8_L2: MVG_GEMM(FP8_L2) break;
      case FP8_IP: MVG_GEMM(FP8_IP) break;
      default: MVG_GEMM(U8_L2) break;
    }
#undef MVG_GEMM
    MVG_LAUNCH_CHECK("gemm_topk_kernel");
  };
  if (!pl.wide) {
    run16(pl.S, pl.split_len, pk, nullptr);
    return;
  }
  const int qtiles = ceil_div(nq, kWQ);
  // one launch of the 64-query form: S_ splits of split_len_ local rows, top
  // kw_ per split into out; P_ / R_ as in SplitRows; bnd: per-query bound
  auto wide = [&](bool str, int64_t S_, int64_t split_len_, int kw_, int P_, int R_, uint64_t *out,
                  const uint64_t *bnd, int bstride) {
    const dim3 grid(xcd_grid(qtiles, S_));
    if (pl.reg) {
      const WideRegLayout rlay(kw_);
#define MVG_WREG(K, N, T)                                                                          \
  {                                                                                                \
    const void *rf = (const void *)gemm_wide_reg_kernel<K, N, T>;                                  \
    if (rlay.total > 64 * 1024)                                                                    \
      MVG_HIP(hipFuncSetAttribute(rf, hipFuncAttributeMaxDynamicSharedMemorySize, rlay.total));    \
    hipLaunchKernelGGL((gemm_wide_reg_kernel<K, N, T>), grid, dim3(kNT), rlay.total, s, Q, X, nq, n, \
                       d, qnorm, xnorm, qn_u, qs_u, xn_u, xs_u, scale_sq, kw_, split_len_, excl,   \
                       excl_len, out, (int)S_, qinv, xinv, P_, R_, bnd, bstride);                  \
  }
#define MVG_WREG_KIND(T)                          \
  switch (kind) {                                 \
    case BF16_L2: MVG_WREG(BF16_L2, 24, T) break; \
    case BF16_IP: MVG_WREG(BF16_IP, 24, T) break; \
    case I8_IP: MVG_WREG(I8_IP, 12, T) break;     \
    case FP8_L2: MVG_WREG(FP8_L2, 12, T) break;   \
    case FP8_IP: MVG_WREG(FP8_IP, 12, T) break;   \
    default: MVG_WREG(U8_L2, 12, T) break;        \
  }
      if (str) MVG_WREG_KIND(true) else MVG_WREG_KIND(false)
#undef MVG_WREG_KIND
#undef MVG_WREG
      MVG_LAUNCH_CHECK("gemm_wide_reg_kernel");
    } else {
      const WideLayout wlay(row_bytes, kw_);
#define MVG_WIDE(K, T)                                                                             \
  {                                                                                                \
    const void *fn = (const void *)gemm_wide_kernel<K, T>;                                         \
    if (wlay.total > 64 * 1024)                                                                    \
      MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, wlay.total));    \
    hipLaunchKernelGGL((gemm_wide_kernel<K, T>), grid, dim3(kNT), wlay.total, s, Q, X, nq, n,      \
                       row_bytes, d, qnorm, xnorm, qn_u, qs_u, xn_u, xs_u, scale_sq, kw_,           \
                       split_len_, excl, excl_len, out, (int)S_, qinv, xinv, P_, R_, bnd, bstride); \
  }
#define MVG_WIDE_KIND(T)                      \
  switch (kind) {                             \
    case BF16_L2: MVG_WIDE(BF16_L2, T) break; \
    case BF16_IP: MVG_WIDE(BF16_IP, T) break; \
    case I8_IP: MVG_WIDE(I8_IP, T) break;     \
    case FP8_L2: MVG_WIDE(FP8_L2, T) break;   \
    case FP8_IP: MVG_WIDE(FP8_IP, T) break;   \
    default: MVG_WIDE(U8_L2, T) break;        \
  }
      if (str) MVG_WIDE_KIND(true) else MVG_WIDE_KIND(false)
#undef MVG_WIDE_KIND
#undef MVG_WIDE
      MVG_LAUNCH_CHECK("gemm_wide_kernel");
    }
  };
  if (!pl.strided) {
    wide(false, pl.S, pl.split_len, pl.kw, 1, 1, pk, nullptr, 0);
    return;
  }
  const int64_t S = pl.S;
  const int kw = pl.kw;
  uint64_t *redo = pk + (size_t)nq * S * kw;
  uint8_t *flag = reinterpret_cast<uint8_t *>(redo + (size_t)nq * kk);
  const uint64_t *bound = nullptr;
  if (pl.Ps > 0) {
    // sample pass: rows 0, Ps, 2 Ps, ... -> each query's sorted top kks,
    // whose last key bounds the strided pass
    uint64_t *spart = reinterpret_cast<uint64_t *>(flag + ((nq + 7) & ~int64_t(7)));
    uint64_t *skeys = spart + (size_t)nq * pl.Ss * pl.kks;
    wide(true, pl.Ss, pl.split_len_s, pl.kks, pl.Ps, 1, spart, nullptr, 0);
    launch_merge(spart, nq, pl.Ss * pl.kks, pl.kks, 0, 0.0f, nullptr, nullptr, nullptr, skeys, s);
    bound = skeys + pl.kks - 1;
  }
  wide(true, S, pl.split_len, kw, pl.P, pl.P, pk, bound, pl.kks);
  hipLaunchKernelGGL(gemm_certify_kernel, dim3((unsigned)ceil_div(nq, kNT / 64)), dim3(kNT), 0, s, pk, nq, (int)S,
                     kw, kk, bound, pl.kks, flag);
  MVG_LAUNCH_CHECK("gemm_certify_kernel");
  // one split over every row: the rejected queries' top kk, sorted
  run16(1, ((n + kStepCols - 1) / kStepCols) * kStepCols, redo, flag);
  hipLaunchKernelGGL(gemm_fixup_kernel, dim3((unsigned)nq), dim3(kNT), 0, s, pk, ncand, redo, kk, flag);
  MVG_LAUNCH_CHECK("gemm_fixup_kernel");
}

}  // namespace mvg
