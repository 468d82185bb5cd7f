// IVF-PQ search: coarse top-nprobe, per-(query, list) LUT build in LDS, ADC
// scan of the list's codes, per-list top-(k*k_factor), per-query merge and the
// optional exact FP32 rerank.
//
// Plain form of IndexIVFPQ.search (mojovec/index/index_ivf_pq.mojo:225-360):
//   L2: LUT of the residual q - c_l (pq.mojo:252-292), dist = sum_m LUT[m][code_m]
//       summed sequentially in m starting from 0.0 (:336-341).
//   IP: LUT of q, dist = -<q, c_l> + sum_m LUT[m][code_m] (:316-324).
// Ids inside one list are ascending with position (append order of internal
// ids), so the list-local (key, position) order is the canonical (dist, id)
// order.
#include "dist.h"
#include "kernels.h"
#include "select.h"

namespace mvg {

namespace {

constexpr int kScanChunk = 4096;  // list positions scored per select round

template <int MF>
__device__ __forceinline__ float adc_sum(const uint8_t *__restrict__ code, const float *lut, int M,
                                         float init) {
  float dist = init;
  if constexpr (MF == 16) {
    const uint4 c = *reinterpret_cast<const uint4 *>(code);
    const uint32_t w[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
    for (int m = 0; m < 16; m++) dist += lut[m * kKsub + ((w[m >> 2] >> (8 * (m & 3))) & 0xFF)];
  } else if constexpr (MF == 12 || MF == 8) {
    uint32_t w[MF / 4];
#pragma unroll
    for (int t = 0; t < MF / 4; t++) w[t] = reinterpret_cast<const uint32_t *>(code)[t];
#pragma unroll
    for (int m = 0; m < MF; m++) dist += lut[m * kKsub + ((w[m >> 2] >> (8 * (m & 3))) & 0xFF)];
  } else {
    for (int m = 0; m < M; m++) dist += lut[m * kKsub + code[m]];
  }
  return dist;
}

// Compile-time bool passed to generic lambdas as an argument.
template <bool B>
struct BoolC {
  static constexpr bool value = B;
};

// Code words of one entry for the fixed-M scans (MF = 8, 12 or 16), loaded
// ahead of the LUT lookups so that a thread keeps several entries' loads in
// flight; adc_words sums them in the same order as adc_sum.
template <int MF>
struct Code {
  uint32_t w[MF > 0 ? MF / 4 : 1];
};

template <int MF>
__device__ __forceinline__ Code<MF> load_code(const uint8_t *__restrict__ codes, int64_t entry) {
  Code<MF> c;
  if constexpr (MF == 16) {
    const uint4 v = *reinterpret_cast<const uint4 *>(codes + entry * 16);
    c.w[0] = v.x; c.w[1] = v.y; c.w[2] = v.z; c.w[3] = v.w;
  } else if constexpr (MF == 8) {
    const uint2 v = *reinterpret_cast<const uint2 *>(codes + entry * 8);
    c.w[0] = v.x; c.w[1] = v.y;
  } else if constexpr (MF == 12) {
    const uint32_t *p = reinterpret_cast<const uint32_t *>(codes + entry * 12);
    c.w[0] = p[0]; c.w[1] = p[1]; c.w[2] = p[2];
  } else {
    c.w[0] = 0u;
  }
  return c;
}

// load_code with a 32-bit byte offset (the index's code bytes below 4 GiB):
// base + 32-bit offset addressing instead of a 64-bit address per load.
template <int MF>
__device__ __forceinline__ Code<MF> load_code32(const uint8_t *__restrict__ codes, uint32_t entry) {
  Code<MF> c;
  const uint8_t *p = codes + (uint32_t)(entry * (uint32_t)MF);
  if constexpr (MF == 16) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    c.w[0] = v.x; c.w[1] = v.y; c.w[2] = v.z; c.w[3] = v.w;
  } else if constexpr (MF == 8) {
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    c.w[0] = v.x; c.w[1] = v.y;
  } else if constexpr (MF == 12) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(p);
    c.w[0] = w[0]; c.w[1] = w[1]; c.w[2] = w[2];
  } else {
    c.w[0] = 0u;
  }
  return c;
}

template <int MF>
__device__ __forceinline__ float adc_words(const Code<MF> &c, const float *lut, float init) {
  float dist = init;
#pragma unroll
  for (int m = 0; m < MF; m++) dist += lut[m * kKsub + ((c.w[m >> 2] >> (8 * (m & 3))) & 0xFF)];
  return dist;
}

// LDS word at byte address a (address space 3). The scan kernels keep their
// LUT at the start of dynamic LDS and declare no static LDS, so the table's
// LDS address is 0: with the address a known constant plus the code byte
// times 4, the compiler selects and shifts the byte in one SDWA op and puts
// the subquantizer offset in the ds_read immediate. (Through the generic
// pointer it kept a separate add of the late-resolved dynamic-LDS base.)
typedef __attribute__((address_space(3))) const float lds_float;
__device__ __forceinline__ float lds_word(uint32_t a) {
  return *reinterpret_cast<lds_float *>(static_cast<uintptr_t>(a));
}

// Byte 0 of w times 4 in one SDWA op. The compiler selects SDWA for bytes
// 1-3 but writes byte 0 as a shift and a mask (two ops).
__device__ __forceinline__ uint32_t byte0_x4(uint32_t w) {
  uint32_t a;
  asm("v_lshlrev_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
      : "=v"(a)
      : "v"(2u), "v"(w));
  return a;
}

template <int MF, int U, int M_>
__device__ __forceinline__ void adc_step(const Code<MF> (&c)[U], float (&out)[U]) {
  if constexpr (M_ < MF) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t a = (M_ & 3) == 0 ? byte0_x4(c[u].w[M_ >> 2])
                                       : (((c[u].w[M_ >> 2] >> (8 * (M_ & 3))) & 0xFFu) << 2);
      out[u] += lds_word(M_ * kKsub * 4 + a);
    }
    adc_step<MF, U, M_ + 1>(c, out);
  }
}

// adc_words of U entries at once: each entry's sum keeps the m order, and the
// U independent chains interleave so that U LUT lookups are in flight. lut
// must be the LUT at LDS address 0 (the callers' layouts put it first in
// dynamic LDS, and they declare no static LDS).
template <int MF, int U>
__device__ __forceinline__ void adc_words_multi(const Code<MF> (&c)[U], const float *lut,
                                                const float (&init)[U], float (&out)[U]) {
  (void)lut;
#pragma unroll
  for (int u = 0; u < U; u++) out[u] = init[u];
  adc_step<MF, U, 0>(c, out);
}

// 8-bit tables (mixed L2 deep-pool form): byte at LDS address a, with the
// table at LDS address 0 as for lds_word. A 256-entry byte table spans 64
// dwords, so a wavefront's 64 lookups into one subquantizer's table meet at
// most two dwords per bank, against four for the FP32 table.
typedef __attribute__((address_space(3))) const uint8_t lds_uint8;
__device__ __forceinline__ uint32_t lds_byte(uint32_t a) {
  return *reinterpret_cast<lds_uint8 *>(static_cast<uintptr_t>(a));
}

// Sums of the U entries' 8-bit table values, exact in integers (at most
// 16 * 255), so their order is free: the lookups are issued first, then
// added pairwise.
template <int MF, int U>
__device__ __forceinline__ void adc_q8_multi(const Code<MF> (&c)[U], uint32_t (&out)[U]) {
  uint32_t t[U][MF];
#pragma unroll
  for (int m = 0; m < MF; m++)
#pragma unroll
    for (int u = 0; u < U; u++) {
      t[u][m] = lds_byte(m * kKsub + ((c[u].w[m >> 2] >> (8 * (m & 3))) & 0xFFu));
      // ds_read_u8 zero-extends; without this (empty) operand barrier the
      // compiler narrowed the sums to 16 bits and masked every byte again
      asm("" : "+v"(t[u][m]));
    }
#pragma unroll
  for (int u = 0; u < U; u++) {
    uint32_t acc = 0u;
#pragma unroll
    for (int m = 0; m < MF; m += 2) acc += t[u][m] + t[u][m + 1];
    out[u] = acc;
  }
}

// One pool word (key << 32 | entry) at segment slot i.
template <bool OFF32>
__device__ __forceinline__ void seg_put(uint64_t *seg, int i, uint32_t key, uint32_t entry) {
  const uint64_t v = ((uint64_t)key << 32) | entry;
  if constexpr (OFF32) *reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(seg) + ((uint32_t)i << 3)) = v;
  else seg[i] = v;
}

__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// One workgroup per (query, probe rank).
template <int MF, bool IP>
__global__ __launch_bounds__(kNT) void ivfpq_scan_kernel(
    const float *__restrict__ q, int d, int M, const float *__restrict__ coarse,
    const float *__restrict__ pq, const int64_t *__restrict__ list_off,
    const uint8_t *__restrict__ codes, const int32_t *__restrict__ ids,
    const int32_t *__restrict__ probes, int maxp, const int32_t *__restrict__ nprobe_q,
    int nprobe_all, int kk, const uint8_t *__restrict__ excl, int64_t excl_len,
    uint64_t *__restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *lut = reinterpret_cast<float *>(smem);                    // M * 256
  float *rv = lut + M * kKsub;                                     // d (+ 1 for coarse term)
  uint32_t *keys = reinterpret_cast<uint32_t *>(rv + ((d + 4) & ~3));  // kk + kScanChunk
  uint32_t *sel = keys + kk + kScanChunk;                          // kk
  uint32_t *prev_pos = sel + kk;                                   // kk
  SelectScratch *scr = reinterpret_cast<SelectScratch *>(prev_pos + ((kk + 3) & ~3));

  const int tid = threadIdx.x;
  const int64_t w = blockIdx.x;
  const int64_t qi = w / maxp;
  const int p = (int)(w - qi * maxp);
  uint64_t *dst = part + w * kk;
  int np = nprobe_q ? nprobe_q[qi] : nprobe_all;
  np = min(max(np, 1), maxp);
  const int l = p < np ? probes[qi * maxp + p] : -1;
  if (l < 0) {
    for (int t = tid; t < kk; t += kNT) dst[t] = kSentinel64;
    return;
  }
  const int64_t off = list_off[l];
  const int64_t L = list_off[l + 1] - off;
  if (L == 0) {
    for (int t = tid; t < kk; t += kNT) dst[t] = kSentinel64;
    return;
  }
  const int dsub = d / M;
  const float *qq = q + qi * d;
  const float *cl = coarse + (int64_t)l * d;
  for (int j = tid; j < d; j += kNT) rv[j] = IP ? qq[j] : qq[j] - cl[j];  // :306-310
  __syncthreads();
  for (int e = tid; e < M * kKsub; e += kNT) {
    const int m = e >> 8;
    const float *c = pq + (int64_t)e * dsub;
    lut[e] = IP ? -ip_pq(rv + m * dsub, c, dsub) : l2_kmeans(rv + m * dsub, c, dsub);
  }
  if (IP && tid == 0) rv[d] = -ip_wn<8>(qq, cl, d);  // :322-324
  __syncthreads();
  const float coarse_d = IP ? rv[d] : 0.0f;

  int nprev = 0;
  for (int64_t cb = 0; cb < L; cb += kScanChunk) {
    const int len = (int)min<int64_t>(kScanChunk, L - cb);
    for (int t = tid; t < len; t += kNT) {
      const int64_t pos = off + cb + t;
      uint32_t key = kSentinel;
      bool skip = false;
      if (excl) {
        const int32_t id = ids[pos];
        skip = id >= 0 && id < excl_len && excl[id] != 0;  // :328-334
      }
      if (!skip) key = f2key(adc_sum<MF>(codes + pos * M, lut, M, coarse_d));
      keys[nprev + t] = key;
    }
    __syncthreads();
    const int cnt = select_topk_u32(keys, nprev + len, kk, *scr, sel);
    // Compact the winners to the front, mapping indices to list positions.
    uint32_t kv[(kMaxCandidates + kNT - 1) / kNT];
    uint32_t pv[(kMaxCandidates + kNT - 1) / kNT];
    int nm = 0;
    for (int t = tid; t < cnt; t += kNT, nm++) {
      const uint32_t idx = sel[t];
      kv[nm] = keys[idx];
      pv[nm] = idx < (uint32_t)nprev ? prev_pos[idx] : (uint32_t)(cb + idx - nprev);
    }
    __syncthreads();
    nm = 0;
    for (int t = tid; t < cnt; t += kNT, nm++) {
      keys[t] = kv[nm];
      prev_pos[t] = pv[nm];
    }
    nprev = cnt;
    __syncthreads();
  }
  for (int t = tid; t < kk; t += kNT) {
    uint64_t v = kSentinel64;
    if (t < nprev) v = ((uint64_t)keys[t] << 32) | (uint32_t)ids[off + prev_pos[t]];
    dst[t] = v;
  }
}

// Query of a rerank workgroup. Without an order, workgroup b reranks query b.
// With one (the clustered order of the deep-pool scans), the grid is
// kRerankRanges * ceil(nq / kRerankRanges) and the workgroups of one XCD
// (consecutive workgroups go to consecutive XCDs) take a contiguous stretch of
// the order, so queries that share lists share their rows in that XCD's L2.
// Returns -1 for the grid's padding.
constexpr int kRerankRanges = 8;
__device__ __forceinline__ int64_t rerank_query(int64_t nq, const int32_t *__restrict__ order) {
  if (!order) return blockIdx.x;
  const int64_t per = (nq + kRerankRanges - 1) / kRerankRanges;
  const int64_t t = (int64_t)(blockIdx.x % kRerankRanges) * per + blockIdx.x / kRerankRanges;
  return t < nq ? (int64_t)order[t] : -1;
}

// Row of candidate id, or -1 when it has none: the id itself, or with MAP
// (explicit ids) rm.row_of[id]. A template flag so that the unmapped
// kernels keep one register per candidate.
template <bool MAP>
__device__ __forceinline__ int64_t cand_row(uint32_t id, RowMap rm, int64_t nrows) {
  int64_t r = (int64_t)id;
  if (MAP) r = r < rm.nids ? (int64_t)rm.row_of[id] : -1;
  return (r >= 0 && r < nrows) ? r : -1;
}

// Writes query qi's sorted top keys (cnt valid) as k result slots, padding
// slots at or past cnt and past the query's own count kq[qi].
template <bool IP>
__device__ __forceinline__ void write_results(const uint64_t *top, int cnt, int64_t qi, int k,
                                              const int32_t *__restrict__ kq, float pad_dist,
                                              float *__restrict__ D, int64_t *__restrict__ I) {
  const int kv = kq ? min(kq[qi], k) : k;
  for (int t = threadIdx.x; t < k; t += kNT) {
    float dv = pad_dist;
    int64_t iv = -1;
    if (t < cnt && t < kv) {
      dv = key2f((uint32_t)(top[t] >> 32));
      if (IP) dv = -dv;
      iv = (int64_t)(uint32_t)(top[t] & 0xFFFFFFFFu);
    }
    D[qi * k + t] = dv;
    I[qi * k + t] = iv;
  }
}

template <bool IP, bool MAP>
__global__ __launch_bounds__(kNT) void rerank_kernel(const float *__restrict__ q, int d,
                                                     const float *__restrict__ vectors,
                                                     int64_t nrows,
                                                     const uint64_t *__restrict__ cand, int kk,
                                                     int k, int buf_n, float pad_dist,
                                                     float *__restrict__ D,
                                                     int64_t *__restrict__ I, int64_t nq,
                                                     const int32_t *__restrict__ order, RowMap rm,
                                                     const int32_t *__restrict__ kq) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *qs = reinterpret_cast<float *>(smem);
  uint64_t *buf = reinterpret_cast<uint64_t *>(qs + ((d + 1) & ~1));
  uint64_t *top = buf + buf_n;
  SelectScratch *scr = reinterpret_cast<SelectScratch *>(top + buf_n);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t qi = rerank_query(nq, order);
  if (qi < 0) return;
  for (int j = tid; j < d; j += kNT) qs[j] = q[qi * d + j];
  for (int t = tid; t < buf_n; t += kNT) buf[t] = kSentinel64;
  __syncthreads();
  const uint64_t *cq = cand + qi * kk;
  for (int c = wave; c < kk; c += kNW) {
    const uint64_t key = cq[c];
    const uint32_t id = (uint32_t)(key & 0xFFFFFFFFu);
    const int64_t row = key == kSentinel64 ? -1 : cand_row<MAP>(id, rm, nrows);
    if (row < 0) continue;
    const float ex = wave_dist_w64<IP>(qs, vectors + row * d, d, lane);
    if (lane == 0) buf[c] = pack_key(IP ? -ex : ex, id);
  }
  __syncthreads();
  const int cnt = select_sort_u64(buf, kk, k, *scr, top);
  write_results<IP>(top, cnt, qi, k, kq, pad_dist, D, I);
}

// Exact distances of a query's kk candidates into buf (sentinel for empty or
// invalid slots): sixteen 16-lane groups, each scoring R candidates per step
// so that R rows per group are in flight.
template <bool IP, int R, bool BIG, bool MAP>
__device__ __forceinline__ void rerank_rows(const float *qs, const float *__restrict__ vectors,
                                            int64_t nrows, const uint64_t *__restrict__ cq, int kk,
                                            int d, uint64_t *buf, RowMap rm) {
  constexpr int G = kNT / 16;
  const int glane = threadIdx.x & 15, group = threadIdx.x >> 4;
  for (int c = group; c < kk; c += R * G) {
    uint32_t id[R];
    bool ok[R];
    const float *x[R];
    bool any = false;
#pragma unroll
    for (int t = 0; t < R; t++) {
      const int ct = c + t * G;
      const uint64_t kv = ct < kk ? cq[ct] : kSentinel64;
      id[t] = (uint32_t)(kv & 0xFFFFFFFFu);
      const int64_t row = kv == kSentinel64 ? -1 : cand_row<MAP>(id[t], rm, nrows);
      ok[t] = row >= 0;
      x[t] = ok[t] ? vectors + row * d : qs;  // an invalid slot reads the query row
      any |= ok[t];
    }
    float e[R];
#pragma unroll
    for (int t = 0; t < R; t++) e[t] = 0.0f;
    if (any) group16_distR_w64<IP, R, BIG>(qs, x, d, glane, e);
    if (glane == 0) {
#pragma unroll
      for (int t = 0; t < R; t++) {
        const int ct = c + t * G;
        if (ct < kk) buf[ct] = ok[t] ? pack_key(IP ? -e[t] : e[t], id[t]) : kSentinel64;
      }
    }
  }
}

// Rerank for d a multiple of 64: sixteen 16-lane groups per workgroup each
// score one candidate row at a time (group16_dist_w64, bit-identical to the
// wavefront form), so 16 rows per workgroup are in flight instead of 4.
template <bool IP, bool SMALLD, bool MAP>
__global__ __launch_bounds__(kNT) void rerank_group_kernel(const float *__restrict__ q, int d,
                                                           const float *__restrict__ vectors,
                                                           int64_t nrows,
                                                           const uint64_t *__restrict__ cand,
                                                           int kk, int k, float pad_dist,
                                                           float *__restrict__ D,
                                                           int64_t *__restrict__ I, int64_t nq,
                                                           const int32_t *__restrict__ order,
                                                           RowMap rm,
                                                           const int32_t *__restrict__ kq) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *qs = reinterpret_cast<float *>(smem);                      // d
  uint64_t *buf = reinterpret_cast<uint64_t *>(qs + d);             // kk
  uint64_t *top = buf + ((kk + 1) & ~1);                            // next_pow2(k)
  SelectScratch64 *scr = reinterpret_cast<SelectScratch64 *>(top + next_pow2(k));
  const int tid = threadIdx.x;
  const int64_t qi = rerank_query(nq, order);
  if (qi < 0) return;
  for (int j = tid; j < d; j += kNT) qs[j] = q[qi * d + j];
  __syncthreads();
  const uint64_t *cq = cand + qi * kk;
  if constexpr (SMALLD)  // d < 256: four rows per group, no four-accumulator blocks
    rerank_rows<IP, 4, false, MAP>(qs, vectors, nrows, cq, kk, d, buf, rm);
  else
    rerank_rows<IP, 2, true, MAP>(qs, vectors, nrows, cq, kk, d, buf, rm);
  __syncthreads();
  const int cnt = select_sort_u64(buf, kk, k, *scr, top);
  write_results<IP>(top, cnt, qi, k, kq, pad_dist, D, I);
}

// Mixed-form pre-rerank: the SQ8 code distance of each of a query's kk
// candidates (the integer SQ8 distance of index_flat_sq8.mojo:80-177 over
// codes of the stored rows), then the kk2 best of them, unordered, as the
// candidates of the exact FP32 rerank. Eight lanes per candidate row, 16-byte
// code chunks, eight rows per group in flight; keys are (integer distance << 32
// | id), so equal code distances keep the lower id.
template <bool IP, bool MAP>
__global__ __launch_bounds__(kNT) void sq8_prefilter_kernel(
    const uint8_t *__restrict__ qcodes, const uint32_t *__restrict__ qnorms, int ld,
    const uint8_t *__restrict__ vcodes, const uint32_t *__restrict__ vnorms, int64_t nrows,
    const uint64_t *__restrict__ cand, int kk, int kk2, uint64_t *__restrict__ out, int64_t nq,
    const int32_t *__restrict__ order, RowMap rm, const int32_t *__restrict__ kq, int kf,
    int per_k) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t *buf = reinterpret_cast<uint64_t *>(smem);  // kk
  uint64_t *top = buf + ((kk + 1) & ~1);               // kk2
  BucketScratch *scr = reinterpret_cast<BucketScratch *>(top + ((kk2 + 1) & ~1));
  const int64_t qi = rerank_query(nq, order);
  if (qi < 0) return;
  constexpr int G = kNT / 8, R = 8;
  const int tid = threadIdx.x, gl = tid & 7, group = tid >> 3;
  const uint4 *qa = reinterpret_cast<const uint4 *>(qcodes + qi * ld);
  const int words = ld / 16;
  const uint64_t *cq = cand + qi * kk;
  for (int c = group; c < kk; c += R * G) {
    uint32_t id[R], acc[R];
    int32_t row[R];  // rows < 2^31 (the device id width)
    bool ok[R];
#pragma unroll
    for (int t = 0; t < R; t++) {
      const int ct = c + t * G;
      const uint64_t kv = ct < kk ? cq[ct] : kSentinel64;
      id[t] = (uint32_t)(kv & 0xFFFFFFFFu);
      if constexpr (MAP) {
        row[t] = kv == kSentinel64 ? -1 : (int32_t)cand_row<true>(id[t], rm, nrows);
        ok[t] = row[t] >= 0;
      } else {  // the row is the id: one register
        row[t] = (int32_t)id[t];
        ok[t] = kv != kSentinel64 && (int64_t)id[t] < nrows;
      }
      acc[t] = 0;
    }
    // every row's chunk loaded without a branch (an invalid candidate reads
    // row 0 and is dropped below), so the R loads issue back to back
#pragma unroll
    for (int t = 0; t < R; t++) row[t] = ok[t] ? row[t] : 0;
    for (int w = gl; w < words; w += 8) {
      const uint4 a = qa[w];
      uint4 bv[R];
#pragma unroll
      for (int t = 0; t < R; t++) bv[t] = reinterpret_cast<const uint4 *>(vcodes + (int64_t)row[t] * ld)[w];
#pragma unroll
      for (int t = 0; t < R; t++) {
        const uint4 b = bv[t];
        if (IP) {
          int v = (int)acc[t];
          v = __builtin_amdgcn_sdot4((int)a.x, (int)b.x, v, false);
          v = __builtin_amdgcn_sdot4((int)a.y, (int)b.y, v, false);
          v = __builtin_amdgcn_sdot4((int)a.z, (int)b.z, v, false);
          v = __builtin_amdgcn_sdot4((int)a.w, (int)b.w, v, false);
          acc[t] = (uint32_t)v;
        } else {
          acc[t] = __builtin_amdgcn_udot4(a.x, b.x, acc[t], false);
          acc[t] = __builtin_amdgcn_udot4(a.y, b.y, acc[t], false);
          acc[t] = __builtin_amdgcn_udot4(a.z, b.z, acc[t], false);
          acc[t] = __builtin_amdgcn_udot4(a.w, b.w, acc[t], false);
        }
      }
    }
    // sum over the row's 8 lanes (one half-row of the wavefront; a group's
    // lanes are active together): DPP lane xor 1 and xor 2 within each quad,
    // then the half-row mirror (lane i with lane 7 - i: the other quad)
#pragma unroll
    for (int t = 0; t < R; t++) {
      acc[t] += (uint32_t)__builtin_amdgcn_mov_dpp((int)acc[t], 0xB1, 0xF, 0xF, false);
      acc[t] += (uint32_t)__builtin_amdgcn_mov_dpp((int)acc[t], 0x4E, 0xF, 0xF, false);
      acc[t] += (uint32_t)__builtin_amdgcn_mov_dpp((int)acc[t], 0x141, 0xF, 0xF, false);
    }
    if (gl == 0) {
#pragma unroll
      for (int t = 0; t < R; t++) {
        const int ct = c + t * G;
        if (ct >= kk) continue;
        uint32_t key = 0xFFFFFFFFu;
        if (ok[t]) {
          if (IP) {
            key = (uint32_t)((int64_t)0x80000000LL - (int64_t)(int32_t)acc[t]);
          } else {
            const int64_t v = (int64_t)qnorms[qi] + (int64_t)vnorms[row[t]] - 2 * (int64_t)acc[t];
            key = (uint32_t)(v > 0 ? v : 0);
          }
        }
        buf[ct] = ok[t] ? ((uint64_t)key << 32 | id[t]) : kSentinel64;
      }
    }
  }
  __syncthreads();
  // per-query k: the width a search with k = kq[qi] alone would keep
  int w2 = kk2;
  if (kq) {
    const int64_t kqi = kq[qi], kkq = kqi * kf;
    w2 = (int)(per_k > 0 ? min(kkq, max((int64_t)64, (int64_t)per_k * kqi)) : kkq);
    w2 = min(w2, kk2);
  }
  // the w2 smallest (code distance, id) keys, unordered: the rerank selects
  // over its candidates
  const int cnt = bucket_select_u64(buf, kk, w2, top, *scr);
  for (int t = tid; t < kk2; t += kNT) out[qi * kk2 + t] = t < cnt ? top[t] : kSentinel64;
}

// ---------------------------------------------------------------------------
// Per-query scan: one workgroup walks the query's probed lists in coarse
// order and keeps a running canonical top-kk of u64 (key << 32 | id) in LDS.
// A candidate enters the append buffer only if it beats the current kk-th
// key, which can only shrink, so the filter never drops a final winner. The
// buffer is merged into the top set by a rank merge (small sets) or the radix
// select (large sets).
//
// LUT per (query, list):
//   plain L2  : the exact residual table of pq.mojo:252-292 (q - c_l), as the
//               reference computes it;
//   mixed L2  : T[l][m][k] + Q[q][m][k] with T = |c_mk|^2 + 2<c_l,m, c_mk>
//               (precomputed per list) and Q = -2<q_m, c_mk> (per query), on
//               top of the exact FP32 coarse term |q - c_l|^2;
//   IP (both) : -<q_m, c_mk> once per query plus -<q, c_l> per list
//               (index_ivf_pq.mojo:288-293, 316-324).

constexpr int kCandCap = 2048;  // append-buffer capacity (u64 keys)
constexpr int kRound = 4 * kNT;  // list entries scored between buffer checks

struct QueryScanShared {
  int cnt;
  uint64_t thr;
  float base;
  int64_t item;
};

// Publishes this workgroup's kk-th key to the query's shared threshold and
// takes the tighter of the two (called right after a flush).
__device__ void share_threshold(uint64_t *g, QueryScanShared &st) {
  if (threadIdx.x == 0) {
    const uint64_t loc = st.thr;
    const uint64_t old = atomicMin(reinterpret_cast<unsigned long long *>(g), (unsigned long long)loc);
    if (old < loc) st.thr = old;
  }
  __syncthreads();
}

// Work-queue setup: per query, the entry offsets of its probes in coarse
// order, the item count ceil(total / R) clamped to [1, kMaxItemsPerQuery],
// the per-item entry count and a fresh threshold.
constexpr int kMaxItemsPerQuery = 8;
__global__ void wq_setup_kernel(const int32_t *__restrict__ probes, int64_t nq, int maxp,
                                const int32_t *__restrict__ nprobe_q, int nprobe_all,
                                const int64_t *__restrict__ list_off, int64_t R,
                                int64_t *__restrict__ poff, int *__restrict__ nitems,
                                int64_t *__restrict__ rq, uint64_t *__restrict__ gthr) {
  const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= nq) return;
  int np = nprobe_q ? nprobe_q[qi] : nprobe_all;
  np = min(max(np, 1), maxp);
  int64_t acc = 0;
  int64_t *po = poff + qi * (int64_t)(maxp + 1);
  for (int p = 0; p < np; p++) {
    po[p] = acc;
    const int l = probes[qi * maxp + p];
    if (l >= 0) acc += list_off[l + 1] - list_off[l];
  }
  po[np] = acc;
  const int64_t n_items = min<int64_t>(kMaxItemsPerQuery, max<int64_t>(1, (acc + R - 1) / R));
  nitems[qi] = (int)n_items;
  rq[qi] = max<int64_t>(1, (acc + n_items - 1) / n_items);
  gthr[qi] = kSentinel64;
}

// Exclusive scan of the item counts (one workgroup) and queue reset.
__global__ __launch_bounds__(kNT) void wq_offsets_kernel(const int *__restrict__ nitems, int64_t nq,
                                                         int64_t *__restrict__ item_off,
                                                         unsigned *__restrict__ counter) {
  __shared__ int sh[kNW + 1];
  int64_t carry = 0;
  for (int64_t b = 0; b < nq; b += kNT) {
    const int64_t i = b + threadIdx.x;
    const int v = i < nq ? nitems[i] : 0;
    int total;
    const int ex = block_excl_scan(v, sh, total);
    if (i < nq) item_off[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) {
    item_off[nq] = carry;
    *counter = 0u;
  }
}

// Merges the append buffer cand[kk .. kk+cnt) into the sorted top set
// cand[0 .. kk) (sentinel padded). Leaves cand[0..kk) sorted, cnt = 0.
__device__ void flush_candidates(uint64_t *cand, uint64_t *tmp, int kk, QueryScanShared &st,
                                 SelectScratch64 &scr) {
  const int tid = threadIdx.x;
  __syncthreads();
  const int n = kk + st.cnt;
  if (n <= kNT / 2) {
    // rank merge (O(n^2), small sets): keys are unique, sentinels sort last
    // and are dropped
    for (int i = tid; i < kk; i += kNT) tmp[i] = kSentinel64;
    __syncthreads();
    for (int i = tid; i < n; i += kNT) {
      const uint64_t key = cand[i];
      if (key == kSentinel64) continue;
      int r = 0;
      for (int j = 0; j < n; j++) r += cand[j] < key;
      if (r < kk) tmp[r] = key;
    }
    __syncthreads();
  } else {
    select_sort_u64(cand, n, kk, scr, tmp);  // tmp[0..np2) sorted, sentinel tail
  }
  for (int i = tid; i < kk; i += kNT) cand[i] = tmp[i];
  __syncthreads();
  if (tid == 0) {
    st.cnt = 0;
    st.thr = cand[kk - 1];  // sentinel until kk candidates exist
  }
  __syncthreads();
}

// LDS layout of the per-query scan (bytes), shared by host and device.
struct QueryScanLayout {
  int lut, qtab, rv, cand, tmp, scr, st, total;
  __host__ __device__ QueryScanLayout(int M, int d, int kk, bool has_qtab) {
    const int nlut = M * kKsub * 4;
    lut = 0;
    qtab = lut + nlut;
    rv = qtab + (has_qtab ? nlut : 0);
    cand = (rv + 4 * d + 15) & ~15;
    tmp = cand + 8 * (kk + kCandCap);
    scr = tmp + 8 * next_pow2(kk);
    st = (scr + (int)sizeof(SelectScratch64) + 15) & ~15;
    total = st + (int)sizeof(QueryScanShared);
  }
};

// ---- LUT construction shared by the per-query scan kernels --------------
//
// kRegCb shapes (plain L2, MF and DF > 0) keep PQ centroid `tid` of every
// subspace in registers (DF floats per thread), so each per-list LUT rebuild
// reads only the LDS residual.
template <int MF, bool IP, bool MIXED, int DF>
struct LutShape {
  static constexpr bool kRegCb = !IP && !MIXED && DF > 0 && MF > 0;
  static constexpr int kDsub = (DF > 0 && MF > 0) ? DF / MF : 1;
  static constexpr int kRegs = kRegCb ? DF : 1;
};

template <int MF, bool IP, bool MIXED, int DF>
__device__ __forceinline__ void load_reg_codebook(float *cbr, const float *__restrict__ pq) {
  using S = LutShape<MF, IP, MIXED, DF>;
  if constexpr (S::kRegCb) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int m = 0; m < MF; m++) {
#pragma unroll
      for (int j = 0; j < S::kDsub; j++)
        cbr[m * S::kDsub + j] = pq[((int64_t)m * kKsub + tid) * S::kDsub + j];
    }
  }
}

// Per-query tables: the IP LUT (pq.mojo:285-292) or the mixed -2<q_m, c_mk>.
// Ends with a barrier.
template <bool IP, bool MIXED>
__device__ __forceinline__ void build_query_tables(const float *__restrict__ qq, int d, int M,
                                                   const float *__restrict__ pq, float *rv,
                                                   float *lut, float *qtab) {
  const int tid = threadIdx.x, nlut = M * kKsub, dsub = d / M;
  if (IP || MIXED) {
    for (int j = tid; j < d; j += kNT) rv[j] = qq[j];
    __syncthreads();
    for (int e = tid; e < nlut; e += kNT) {
      const int m = e >> 8;
      const float *c = pq + (int64_t)e * dsub;
      if (IP) {
        lut[e] = -ip_pq(rv + m * dsub, c, dsub);
      } else {
        float dot = 0.0f;
        for (int j = 0; j < dsub; j++) dot = fma_(rv[m * dsub + j], c[j], dot);
        qtab[e] = -2.0f * dot;
      }
    }
  }
  __syncthreads();
}

// LUT of probed list l in LDS; returns the list's additive base (IP: -<q,
// c_l>; mixed L2: |q - c_l|^2; plain L2: 0). The caller must have finished
// every read of the previous list's LUT and base. Ends with a barrier.
template <int MF, bool IP, bool MIXED, int DF>
__device__ __forceinline__ float build_list_lut(const float *__restrict__ qq,
                                                const float *__restrict__ coarse, int l, int d,
                                                int M, const float *__restrict__ pq,
                                                const float *__restrict__ ptab, float *rv,
                                                float *lut, const float *qtab, const float *cbr,
                                                float *base_slot) {
  using S = LutShape<MF, IP, MIXED, DF>;
  const int tid = threadIdx.x, nlut = M * kKsub, dsub = d / M;
  const float *cl = coarse + (int64_t)l * d;
  if (!IP) {
    if (!MIXED) {
      for (int j = tid; j < d; j += kNT) rv[j] = qq[j] - cl[j];  // index_ivf_pq.mojo:306-310
      __syncthreads();
      if constexpr (S::kRegCb) {
#pragma unroll
        for (int m = 0; m < MF; m++)
          lut[m * kKsub + tid] = l2_w8_fixed<S::kDsub>(cbr + m * S::kDsub, rv + m * S::kDsub);
      } else {
        for (int e = tid; e < nlut; e += kNT) {
          const int m = e >> 8;
          lut[e] = l2_kmeans(rv + m * dsub, pq + (int64_t)e * dsub, dsub);
        }
      }
    } else {
      const float *tl = ptab + (int64_t)l * nlut;
      for (int e = tid; e < nlut; e += kNT) lut[e] = tl[e] + qtab[e];
      if (tid == 0) *base_slot = l2_w8(qq, cl, d);
    }
  } else if (tid == 0) {
    *base_slot = -ip_wn<8>(qq, cl, d);  // index_ivf_pq.mojo:322-324
  }
  __syncthreads();
  return (IP || MIXED) ? *base_slot : 0.0f;
}

// DF > 0 (with MF > 0): register-resident codebook, see LutShape.
//
// WQ (work queue): the workgroups are persistent and pull items from a
// device counter. An item is a range of at most R consecutive entries of one
// query's probed lists (in coarse order), so a skewed list is shared by
// several workgroups. Each item writes its top-kk to part[item]; a segmented
// merge reduces the items of a query. Items of one query share a threshold
// (gthr[q], atomicMin of each item's kk-th key): a final winner is never
// above it, since every id lives in exactly one item.
struct WorkQueue {
  int64_t nq;
  const int64_t *item_off;  // [nq + 1] first item of each query
  const int64_t *poff;      // [nq][maxp + 1] entry offsets of the probes
  const int64_t *rq;        // [nq] entries per item of each query
  unsigned *counter;        // next item
  uint64_t *gthr;           // [nq] shared threshold
  uint64_t *part;           // [items][kk]
};

template <int MF, bool IP, bool MIXED, int DF, bool WQ>
__global__ __launch_bounds__(kNT) void ivfpq_query_scan_kernel(
    const float *__restrict__ q, int d, int M, const float *__restrict__ coarse,
    const float *__restrict__ pq, const float *__restrict__ ptab,
    const int64_t *__restrict__ list_off, const uint8_t *__restrict__ codes,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ probes, int maxp,
    const int32_t *__restrict__ nprobe_q, int nprobe_all, int kk, int k, const uint8_t *__restrict__ excl,
    int64_t excl_len, uint64_t *__restrict__ out_keys, float *__restrict__ D,
    int64_t *__restrict__ I, float pad_dist, WorkQueue wq) {
  extern __shared__ __align__(16) unsigned char smem[];
  const QueryScanLayout lay(M, d, kk, MIXED && !IP);
  float *lut = reinterpret_cast<float *>(smem + lay.lut);
  float *qtab = reinterpret_cast<float *>(smem + lay.qtab);
  float *rv = reinterpret_cast<float *>(smem + lay.rv);
  uint64_t *cand = reinterpret_cast<uint64_t *>(smem + lay.cand);
  uint64_t *tmp = reinterpret_cast<uint64_t *>(smem + lay.tmp);
  SelectScratch64 *scr = reinterpret_cast<SelectScratch64 *>(smem + lay.scr);
  QueryScanShared *st = reinterpret_cast<QueryScanShared *>(smem + lay.st);

  const int tid = threadIdx.x;
  float cbr[LutShape<MF, IP, MIXED, DF>::kRegs];
  load_reg_codebook<MF, IP, MIXED, DF>(cbr, pq);
  for (;;) {
  // ---- pick the work: a whole query, or one queue item
  int64_t qi, item = 0, begin = 0, end = INT64_MAX;
  if constexpr (WQ) {
    __syncthreads();
    if (tid == 0) st->item = (int64_t)atomicAdd(wq.counter, 1u);
    __syncthreads();
    item = st->item;
    const int64_t nq_total = wq.nq;
    if (item >= wq.item_off[nq_total]) break;
    int64_t lo = 0, hi = nq_total;  // last query whose first item <= item
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (wq.item_off[mid] <= item) lo = mid; else hi = mid;
    }
    qi = lo;
    begin = (item - wq.item_off[qi]) * wq.rq[qi];
    end = begin + wq.rq[qi];
  } else {
    qi = blockIdx.x;
  }
  const float *qq = q + qi * d;
  int np = nprobe_q ? nprobe_q[qi] : nprobe_all;
  np = min(max(np, 1), maxp);
  const int64_t *poff = WQ ? wq.poff + qi * (int64_t)(maxp + 1) : nullptr;

  for (int i = tid; i < kk; i += kNT) cand[i] = kSentinel64;
  if (tid == 0) {
    st->cnt = 0;
    st->thr = WQ ? wq.gthr[qi] : kSentinel64;
  }
  build_query_tables<IP, MIXED>(qq, d, M, pq, rv, lut, qtab);

  int p = 0;
  if constexpr (WQ) {  // first probe whose entries reach `begin`
    while (p + 1 < np && poff[p + 1] <= begin) p++;
  }
  for (; p < np; p++) {
    const int l = probes[qi * maxp + p];
    if (l < 0) continue;
    const int64_t off = list_off[l];
    int64_t L = list_off[l + 1] - off;
    int64_t seg_lo = 0;
    if constexpr (WQ) {
      if (poff[p] >= end) break;
      seg_lo = max<int64_t>(begin - poff[p], 0);
      L = min<int64_t>(L, end - poff[p]);
    }
    if (L <= seg_lo) continue;
    const float base = build_list_lut<MF, IP, MIXED, DF>(qq, coarse, l, d, M, pq, ptab, rv, lut,
                                                         qtab, cbr, &st->base);
    for (int64_t r0 = seg_lo; r0 < L; r0 += kRound) {
      const uint64_t thr = st->thr;
      constexpr int U = kRound / kNT;
      if constexpr (MF > 0) {
        Code<MF> cw[U];
        int32_t idv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {  // every load of the round in flight first
          const int64_t pos = r0 + u * kNT + tid;
          cw[u] = Code<MF>{};
          idv[u] = -1;
          if (pos < L) {
            cw[u] = load_code<MF>(codes, off + pos);
            idv[u] = ids[off + pos];
          }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int64_t pos = r0 + u * kNT + tid;
          if (pos < L) {
            const int32_t id = idv[u];
            const bool skip = excl && id >= 0 && id < excl_len && excl[id] != 0;  // :328-334
            if (!skip) {
              const uint64_t key = pack_key(adc_words<MF>(cw[u], lut, base), (uint32_t)id);
              if (key < thr) cand[kk + atomicAdd(&st->cnt, 1)] = key;
            }
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int64_t pos = r0 + u * kNT + tid;
          if (pos < L) {
            const int32_t id = ids[off + pos];
            const bool skip = excl && id >= 0 && id < excl_len && excl[id] != 0;  // :328-334
            if (!skip) {
              const float dist = adc_sum<MF>(codes + (off + pos) * M, lut, M, base);
              const uint64_t key = pack_key(dist, (uint32_t)id);
              if (key < thr) cand[kk + atomicAdd(&st->cnt, 1)] = key;
            }
          }
        }
      }
      __syncthreads();
      if (st->cnt > kCandCap - kRound) {
        flush_candidates(cand, tmp, kk, *st, *scr);
        if (WQ) share_threshold(wq.gthr + qi, *st);
      }
    }
    if (st->cnt > 0) {
      flush_candidates(cand, tmp, kk, *st, *scr);
      if (WQ) share_threshold(wq.gthr + qi, *st);
    }
  }
  __syncthreads();
  if constexpr (WQ) {
    uint64_t *dst = wq.part + item * kk;
    for (int t = tid; t < kk; t += kNT) dst[t] = cand[t];
    continue;
  }
  // Output: sorted top-kk keys for the rerank, or final distances / ids.
  if (out_keys) {
    for (int t = tid; t < kk; t += kNT) out_keys[qi * kk + t] = cand[t];
  } else {
    for (int t = tid; t < k; t += kNT) {
      const uint64_t kv = cand[t];
      float dv = pad_dist;
      int64_t iv = -1;
      if (kv != kSentinel64) {
        dv = key2f((uint32_t)(kv >> 32));
        if (IP) dv = -dv;  // :357-360
        iv = (int64_t)(uint32_t)(kv & 0xFFFFFFFFu);
      }
      D[qi * k + t] = dv;
      I[qi * k + t] = iv;
    }
  }
  break;
  }  // for (;;)
}

// ---------------------------------------------------------------------------
// Deep candidate pools (k * k_factor in the thousands): a running top-kk
// would re-select and re-sort thousands of keys on every buffer flush, so
// the selection is done once per query instead.
//   1. scan: an entry's u32 distance key is appended, with its CSR entry
//      index, to the workgroup's private segment in HBM -- and its top
//      12-bit digit counted in an LDS histogram -- when that digit is at most
//      a running bound on the digit of the kk-th smallest key; the bound is
//      recomputed from the histogram as the scan goes and only tightens, so
//      every key that can be among the kk smallest is appended (a fraction
//      of the entries once it settles);
//   2. two more digit passes over the segment find the kk-th smallest key T
//      and how many keys equal to T it takes (MSD radix select);
//   3. if T is shared by more keys than are taken, three passes over the
//      ids of those keys find the largest id taken (lower ids win ties);
//   4. one pass over the segment emits (key, id) of the winners, which are
//      sorted in LDS and written out (unsorted when the exact rerank, which
//      selects over all of them, follows).
// The workgroups are persistent: each owns one segment of seg_cap words,
// where seg_cap is the sum of the max_nprobe largest lists (the most any
// query can visit), and takes queries from per-XCD ranges of a clustered
// query order (next_query).
constexpr int kDeepMinCandidates = 513;  // default: the deep form for k * k_factor >= this
constexpr int kDeepBatch = 4;            // list entries per thread with loads in flight
constexpr int kMfmaCoarseMinLists = 4096;  // mixed form: MFMA coarse above this nlist
// Top digit of the deep-pool keys (eighth-octave bins of the distance); the
// other 20 bits are two 10-bit digits.
constexpr int kTopBits = 12;
constexpr int kTopBins = 1 << kTopBits;
constexpr int kTopShift = 32 - kTopBits;
constexpr uint32_t kNoBound = kTopBins - 1;

struct DeepShared {
  int sh[16];  // block_excl_scan / digit search scratch
  int cnt, valid, need_eq, ties;
  int nseg;       // compacted segment: keys written so far
  uint32_t bound;  // top-digit bound of the keys worth writing (stream scan)
  int64_t q;
  float base;
  uint32_t T, Tid;
};

// LDS of the deep-pool scans. The winners (kk keys, or next_pow2(kk) when
// they are sorted in place) overlay the tables and the digit histogram,
// which are dead once the selection has found its cut.
struct DeepLayout {
  int lut, qtab, rv, out, hist, st, total;
  __host__ __device__ DeepLayout(int M, int d, int kk, bool has_qtab, bool sorted) {
    const int nlut = M * kKsub * 4;
    lut = 0;
    qtab = lut + nlut;
    rv = qtab + (has_qtab ? nlut : 0);
    hist = (rv + 4 * (d + 4) + 15) & ~15;
    const int tables_end = hist + 4 * kTopBins;
    out = 0;
    const int out_end = 8 * (sorted ? next_pow2(kk) : kk);
    st = ((tables_end > out_end ? tables_end : out_end) + 15) & ~15;
    total = st + (int)sizeof(DeepShared);
  }
};

// Given a digit histogram hist[0..nb) (nb a multiple of kNT), finds the bin
// holding the remaining-th key: sets bin, the count before it and its size.
__device__ void deep_find_bin(const int *hist, int nb, int remaining, DeepShared &ds, int &bin,
                              int &before, int &count) {
  const int tid = threadIdx.x;
  const int per = nb / kNT;
  int local = 0;
  for (int b = 0; b < per; b++) local += hist[tid * per + b];
  int total;
  const int excl = block_excl_scan(local, ds.sh, total);
  if (excl < remaining && remaining <= excl + local) {
    int c = excl;
    for (int b = 0; b < per; b++) {
      const int h = hist[tid * per + b];
      if (c + h >= remaining) {
        ds.sh[8] = tid * per + b;
        ds.sh[9] = c;
        ds.sh[10] = h;
        break;
      }
      c += h;
    }
  }
  __syncthreads();
  bin = ds.sh[8];
  before = ds.sh[9];
  count = ds.sh[10];
  __syncthreads();
}


__global__ void reset_counter_kernel(unsigned *counter) { *counter = 0u; }

__global__ void fill_zero_kernel(int *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Query order of the persistent deep-pool scans: queries grouped by their
// nearest list (a counting sort on probes[q][0]) and handed out from
// kQueryRanges contiguous ranges, one per XCD (consecutive workgroups are
// placed on consecutive XCDs), so the queries in flight on one XCD probe
// mostly the same lists and find their codes in its L2.
constexpr int kQueryRanges = 8;

// Bucket of a query: its nearest list, or that list's rank (nullable).
__device__ __forceinline__ int order_bucket(const int32_t *__restrict__ probes, int64_t q, int maxp,
                                            int nlist, const int32_t *__restrict__ rank) {
  int l = probes[q * maxp];
  if (l < 0 || l >= nlist) l = 0;
  return rank ? rank[l] : l;
}

__global__ void order_count_kernel(const int32_t *__restrict__ probes, int64_t nq, int maxp,
                                   int nlist, int *__restrict__ cnt, const int32_t *__restrict__ rank) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  atomicAdd(&cnt[order_bucket(probes, q, maxp, nlist, rank)], 1);
}

// One workgroup: exclusive offsets of the per-list counts (the scatter
// advances them as cursors); clears the range counters.
__global__ __launch_bounds__(kNT) void order_offsets_kernel(int *__restrict__ cnt, int nlist,
                                                            int *__restrict__ off,
                                                            unsigned *__restrict__ ranges) {
  __shared__ int sh[kNW + 1];
  const int tid = threadIdx.x;
  if (tid < kQueryRanges) ranges[tid] = 0u;
  int carry = 0;
  for (int base = 0; base < nlist; base += kNT) {
    const int l = base + tid;
    const int c = l < nlist ? cnt[l] : 0;
    int total;
    const int ex = block_excl_scan(c, sh, total);
    if (l < nlist) off[l] = carry + ex;
    carry += total;
  }
}

__global__ void order_scatter_kernel(const int32_t *__restrict__ probes, int64_t nq, int maxp,
                                     int nlist, int *__restrict__ off, int32_t *__restrict__ order,
                                     const int32_t *__restrict__ rank) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  order[atomicAdd(&off[order_bucket(probes, q, maxp, nlist, rank)], 1)] = (int32_t)q;
}

// Next query of a persistent workgroup (nq when its range is drained).
__device__ __forceinline__ int64_t next_query(const int32_t *__restrict__ order, int64_t nq,
                                              unsigned *__restrict__ ranges) {
  const int r = blockIdx.x % kQueryRanges;
  const int64_t beg = nq * r / kQueryRanges, end = nq * (r + 1) / kQueryRanges;
  const int64_t t = beg + (int64_t)atomicAdd(&ranges[r], 1u);
  return t < end ? (int64_t)order[t] : nq;
}

// Appends v to out[] where `take`, through one LDS counter with one atomic
// per wavefront (called by every active lane); slots at or past cap are
// dropped.
__device__ __forceinline__ void wave_append(bool take, uint64_t v, int *counter, uint64_t *out,
                                            int cap) {
  const uint64_t m = __ballot(take);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int at = 0;
  if (lane == leader) at = atomicAdd(counter, (int)__popcll(m));
  at = __shfl(at, leader, 64);
  if (take) {
    const int slot = at + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (slot < cap) out[slot] = v;
  }
}

// Upper bound on the top digit (kTopBits) of the kk-th smallest key, from one
// wavefront's read of the running digit histogram while other waves still
// add to it: counts only grow (and are exact in every bin at or below the
// current bound), so the first bin at which the read's running sum reaches
// kk is at or above the final bin of the kk-th key. kNoBound while fewer than
// kk keys are counted.
__device__ uint32_t wave_digit_bound(const int *hist, int kk) {
  constexpr int PER = kTopBins / 64;  // bins per lane
  const int lane = threadIdx.x & 63;
  const int4 *h4 = reinterpret_cast<const int4 *>(hist + lane * PER);
  int local = 0;
#pragma unroll
  for (int i = 0; i < PER / 4; i++) {
    const int4 v = h4[i];
    local += (v.x + v.y) + (v.z + v.w);
  }
  int incl = local;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (__shfl(incl, 63, 64) < kk) return kNoBound;
  const int L = __ffsll((unsigned long long)__ballot(incl >= kk)) - 1;
  // The bins of lane L, one per lane, on top of the count before them;
  // re-read counts can only be larger, which keeps the bin a valid bound.
  const int before = __shfl(incl - local, L, 64);
  int c = lane < PER ? hist[L * PER + lane] : 0;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(c, o, 64);
    if (lane >= o) c += y;
  }
  const uint64_t hit = __ballot(lane < PER && before + c >= kk);
  return hit ? (uint32_t)(L * PER + __ffsll((unsigned long long)hit) - 1) : (uint32_t)(L * PER + PER - 1);
}

// One MSD digit pass over a compacted segment of (key << 32 | entry) words:
// histogram of the key digit at `shift` (nbits wide) of the keys whose bits
// under `mask` equal `prefix`. The segment is 16-byte aligned.
__device__ void deep_key_pass64(const uint64_t *__restrict__ seg, int64_t n, int shift, int nbits,
                                uint32_t prefix, uint32_t mask, int *hist) {
  const int tid = threadIdx.x, nb = 1 << nbits;
  const uint32_t dm = (uint32_t)nb - 1u;
  for (int b = tid; b < nb; b += kNT) hist[b] = 0;
  __syncthreads();
  const uint4 *s4 = reinterpret_cast<const uint4 *>(seg);
  const int64_t n2 = n >> 1;
  constexpr int U = 4;
  for (int64_t i0 = tid; i0 < n2; i0 += (int64_t)kNT * U) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + (int64_t)u * kNT;
      v[u] = i < n2 ? s4[i] : make_uint4(0u, kSentinel, 0u, kSentinel);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t kv[2] = {v[u].y, v[u].w};  // high words: the keys
#pragma unroll
      for (int c = 0; c < 2; c++)
        if (kv[c] != kSentinel && (kv[c] & mask) == prefix) atomicAdd(&hist[(kv[c] >> shift) & dm], 1);
    }
  }
  if ((n & 1) && tid == 0) {
    const uint32_t key = (uint32_t)(seg[n - 1] >> 32);
    if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & dm], 1);
  }
  __syncthreads();
}

// Selects the kk smallest (key, id) pairs -- ids taken from ids[entry] --
// of a compacted segment of nseg (key << 32 | entry) words, given the digit
// histogram of its keys, into out[] in no particular order; ds.cnt ends as
// the number found (out past min(ds.cnt, kk) is not written). Every key of
// the query at or below the top-digit bin of the kk-th key must be in the
// segment: the MSD radix select (digit passes over the segment, id passes
// for ties at the kk-th key) then runs as over the full key set.
__device__ void deep_select_segment(const uint64_t *__restrict__ seg, int64_t nseg,
                                    const int32_t *__restrict__ ids, int kk, int *hist,
                                    DeepShared &ds, uint64_t *out) {
  const int tid = threadIdx.x;
  {
    constexpr int PER = kTopBins / kNT;
    int local = 0;
    for (int b = tid * PER; b < tid * PER + PER; b++) local += hist[b];
    const int valid = block_sum(local, ds.sh);
    if (tid == 0) ds.valid = valid;
  }
  __syncthreads();
  const bool take_all = ds.valid <= kk;
  uint32_t T = 0u, Tid = 0xFFFFFFFFu;
  if (!take_all) {
    int bin, before, count, remaining = kk;
    deep_find_bin(hist, kTopBins, remaining, ds, bin, before, count);
    remaining -= before;
    uint32_t prefix = (uint32_t)bin << kTopShift, mask = ~0u << kTopShift;
    deep_key_pass64(seg, nseg, 10, 10, prefix, mask, hist);
    deep_find_bin(hist, 1024, remaining, ds, bin, before, count);
    remaining -= before;
    prefix |= (uint32_t)bin << 10;
    mask |= 0x000FFC00u;
    deep_key_pass64(seg, nseg, 0, 10, prefix, mask, hist);
    deep_find_bin(hist, 1024, remaining, ds, bin, before, count);
    remaining -= before;
    T = prefix | (uint32_t)bin;
    if (remaining < count) {  // ties at T: the lowest ids win
      uint32_t iprefix = 0u, imask = 0u;
      const int shifts[3] = {21, 10, 0};
      const int nbits[3] = {11, 11, 10};
      for (int pass = 0; pass < 3; pass++) {
        const int nb = 1 << nbits[pass];
        const uint32_t dm = (uint32_t)nb - 1u;
        for (int b = tid; b < nb; b += kNT) hist[b] = 0;
        __syncthreads();
        for (int64_t j = tid; j < nseg; j += kNT) {
          const uint64_t w = seg[j];
          if ((uint32_t)(w >> 32) != T) continue;
          const uint32_t id = (uint32_t)ids[(uint32_t)w];
          if ((id & imask) == iprefix) atomicAdd(&hist[(id >> shifts[pass]) & dm], 1);
        }
        __syncthreads();
        int ibin, ibefore, icount;
        deep_find_bin(hist, nb, remaining, ds, ibin, ibefore, icount);
        remaining -= ibefore;
        iprefix |= (uint32_t)ibin << shifts[pass];
        imask |= dm << shifts[pass];
      }
      Tid = iprefix;
    }
  }
  if (tid == 0) ds.cnt = 0;
  __syncthreads();
  {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(seg);
    const int64_t n2 = nseg >> 1;
    constexpr int UC = 4;
    for (int64_t i0 = tid; i0 < n2 + (nseg & 1); i0 += (int64_t)kNT * UC) {
      uint4 v[UC];
#pragma unroll
      for (int u = 0; u < UC; u++) {
        const int64_t i = i0 + (int64_t)u * kNT;
        v[u] = i < n2 ? s4[i] : make_uint4(0u, kSentinel, 0u, kSentinel);
        if (i == n2 && (nseg & 1)) {  // the odd last word
          const uint64_t w = seg[nseg - 1];
          v[u] = make_uint4((uint32_t)w, (uint32_t)(w >> 32), 0u, kSentinel);
        }
      }
#pragma unroll
      for (int u = 0; u < UC; u++) {
        const uint32_t kv[2] = {v[u].y, v[u].w}, ev2[2] = {v[u].x, v[u].z};
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const uint32_t key = kv[c];
          bool take = false;
          uint32_t id = 0u;
          if (key != kSentinel && (take_all || key <= T)) {
            id = (uint32_t)ids[ev2[c]];
            take = take_all || key < T || id <= Tid;
          }
          wave_append(take, ((uint64_t)key << 32) | id, &ds.cnt, out, kk);
        }
      }
    }
  }
  __syncthreads();
}

template <int MF, bool IP, bool MIXED, int DF>
__global__ __launch_bounds__(kNT) void ivfpq_deep_scan_kernel(
    const float *__restrict__ q, int64_t nq, int d, int M, const float *__restrict__ coarse,
    const float *__restrict__ pq, const float *__restrict__ ptab,
    const int64_t *__restrict__ list_off, const uint8_t *__restrict__ codes,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ probes, int maxp,
    const int32_t *__restrict__ nprobe_q, int nprobe_all, int kk, int k,
    const uint8_t *__restrict__ excl, int64_t excl_len, uint64_t *__restrict__ out_keys,
    float *__restrict__ D, int64_t *__restrict__ I, float pad_dist, uint32_t *__restrict__ seg_base,
    int64_t seg_cap, unsigned *__restrict__ counter, const int32_t *__restrict__ order) {
  extern __shared__ __align__(16) unsigned char smem[];
  const DeepLayout lay(M, d, kk, MIXED && !IP, out_keys == nullptr);
  float *lut = reinterpret_cast<float *>(smem + lay.lut);
  float *qtab = reinterpret_cast<float *>(smem + lay.qtab);
  float *rv = reinterpret_cast<float *>(smem + lay.rv);
  uint64_t *out = reinterpret_cast<uint64_t *>(smem + lay.out);
  int *hist = reinterpret_cast<int *>(smem + lay.hist);
  DeepShared &ds = *reinterpret_cast<DeepShared *>(smem + lay.st);
  // compacted segment of (key << 32 | entry) words, as in the stream scan
  uint64_t *seg = reinterpret_cast<uint64_t *>(seg_base) + (int64_t)blockIdx.x * seg_cap;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int np2 = next_pow2(kk);
  float cbr[LutShape<MF, IP, MIXED, DF>::kRegs];
  load_reg_codebook<MF, IP, MIXED, DF>(cbr, pq);

  for (;;) {
    __syncthreads();
    if (tid == 0) ds.q = next_query(order, nq, counter);
    __syncthreads();
    const int64_t qi = ds.q;
    if (qi >= nq) break;
    const float *qq = q + qi * d;
    int np = nprobe_q ? nprobe_q[qi] : nprobe_all;
    np = min(max(np, 1), maxp);
    for (int b = tid; b < kTopBins; b += kNT) hist[b] = 0;
    if (tid == 0) {
      ds.nseg = 0;
      ds.bound = kNoBound;
    }
    build_query_tables<IP, MIXED>(qq, d, M, pq, rv, lut, qtab);

    // 1. scan: keys whose top digit is at most the running bound go to the
    //    compacted segment and the digit histogram (see the stream scan)
    int64_t n = 0;
    for (int p = 0; p < np; p++) {
      const int l = probes[qi * maxp + p];
      if (l < 0) continue;
      const int64_t off = list_off[l];
      const int64_t L = min<int64_t>(list_off[l + 1] - off, seg_cap - n);
      if (L <= 0) continue;
      const float base = build_list_lut<MF, IP, MIXED, DF>(qq, coarse, l, d, M, pq, ptab, rv, lut,
                                                           qtab, cbr, &ds.base);
      const uint32_t bound = ds.bound;  // set before the LUT build's barrier
      if constexpr (MF > 0) {
        constexpr int U = kDeepBatch;
        for (int64_t p0 = tid; p0 < L; p0 += (int64_t)kNT * U) {
          Code<MF> cw[U];
          float initv[U], distv[U];
#pragma unroll
          for (int u = 0; u < U; u++) {  // every load of the batch in flight first
            const int64_t pos = p0 + (int64_t)u * kNT;
            cw[u] = Code<MF>{};
            initv[u] = base;
            if (pos < L) cw[u] = load_code<MF>(codes, off + pos);
          }
          adc_words_multi<MF, U>(cw, lut, initv, distv);  // positions past L are dropped
#pragma unroll
          for (int u = 0; u < U; u++) {
            const int64_t pos = p0 + (int64_t)u * kNT;
            bool put = false;
            uint32_t key = kSentinel;
            if (pos < L) {
              bool skip = false;
              if (excl) {  // :328-334
                const int32_t id = ids[off + pos];
                skip = id >= 0 && id < excl_len && excl[id] != 0;
              }
              if (!skip) {
                key = f2key(distv[u]);
                put = (key >> kTopShift) <= bound;
                if (put) atomicAdd(&hist[key >> kTopShift], 1);
              }
            }
            wave_append(put, ((uint64_t)key << 32) | (uint32_t)(off + pos), &ds.nseg, seg, INT32_MAX);
          }
        }
      } else {
        for (int64_t pos = tid; pos < L; pos += kNT) {
          bool skip = false;
          if (excl) {  // :328-334
            const int32_t id = ids[off + pos];
            skip = id >= 0 && id < excl_len && excl[id] != 0;
          }
          bool put = false;
          uint32_t key = kSentinel;
          if (!skip) {
            key = f2key(adc_sum<MF>(codes + (off + pos) * M, lut, M, base));
            put = (key >> kTopShift) <= bound;
            if (put) atomicAdd(&hist[key >> kTopShift], 1);
          }
          wave_append(put, ((uint64_t)key << 32) | (uint32_t)(off + pos), &ds.nseg, seg, INT32_MAX);
        }
      }
      n += L;
      __syncthreads();  // LUT and base are rebuilt for the next list
      if (wave == 0) {  // tighten the bound for the next list
        const uint32_t b = wave_digit_bound(hist, kk);
        if (lane == 0) ds.bound = b;
      }
    }
    __syncthreads();
    const int64_t nseg = ds.nseg;

    // 2.-4. kk-th smallest key, ties by id, emit (shared with the stream scan)
    deep_select_segment(seg, nseg, ids, kk, hist, ds, out);
    const int cnt = min(ds.cnt, kk);
    // the rerank that reads out_keys selects over all of them: no sort
    const int fill_to = out_keys ? kk : np2;
    for (int i = cnt + tid; i < fill_to; i += kNT) out[i] = kSentinel64;
    __syncthreads();
    if (!out_keys) bitonic_sort_u64(out, np2);
    if (out_keys) {
      for (int t = tid; t < kk; t += kNT) out_keys[qi * kk + t] = out[t];
    } else {
      for (int t = tid; t < k; t += kNT) {
        const uint64_t kv = out[t];
        float dv = pad_dist;
        int64_t iv = -1;
        if (kv != kSentinel64) {
          dv = key2f((uint32_t)(kv >> 32));
          if (IP) dv = -dv;  // :357-360
          iv = (int64_t)(uint32_t)(kv & 0xFFFFFFFFu);
        }
        D[qi * k + t] = dv;
        I[qi * k + t] = iv;
      }
    }
  }
}


// ---------------------------------------------------------------------------
// Stream scan: a query's probed lists are scanned as one stream of entries
// with a single LUT per query, so no per-list LUT is built:
//   dist(q, e) = base[p] + pre[e] + sum_m lut[m][code_m]   (m in order)
// Plain IP is the reference's own form (index_ivf_pq.mojo:288-293, 316-341):
// base = -<q, c_l>, lut = -<q_m, c_mk>, no pre term. Mixed L2 uses the
// residual decomposition
//   |q - c_l - r|^2 = |q - c_l|^2 + sum_m T[l][m][k_m] - 2 sum_m <q_m, c_mk>
// with pre[e] = sum_m T[l][m][code_m] computed once per stored entry, so each
// entry costs its code, id and pre reads and M LDS lookups. DEEP selects the
// top-kk once per query with the digit passes of ivfpq_deep_scan_kernel, but
// over a compacted segment: only keys whose top digit is at most a running
// bound on the kk-th key's digit are appended to the workgroup's HBM segment
// as (key << 32 | entry) words and counted in the LDS digit histogram. The
// bound only tightens (the waves take turns recomputing it from the
// histogram), so every key at or below it was appended and its bins hold
// exact counts: the bin of the kk-th key is found from them as from a full
// histogram. Otherwise a running top-kk with a threshold filter (as
// ivfpq_query_scan_kernel).
constexpr int kStreamMaxProbes = 1024;

constexpr int kThreshAppend = 2048;  // appended keys between threshold flushes (mode 2)
constexpr int kThreshJ = (kMaxCandidates + kThreshAppend + kNT - 1) / kNT;

struct StreamLayout {
  int lut, rv, poff, loff, base, cand, tmp, scr, hist, sh, st, qt, total;
  // q8 (deep pool only): 8-bit tables of M * 256 bytes at 0; the FP32
  // tables they are quantized from are staged in the digit histogram's
  // space (M <= 16), and qt holds each subquantizer's minimum and range.
  __host__ __device__ StreamLayout(int M, int d, int kk, int maxp, int mode, bool sorted = true,
                                   bool q8 = false) {
    const bool deep = mode == 1;
    lut = 0;
    qt = 0;
    rv = (lut + M * kKsub * (q8 ? 1 : 4) + 15) & ~15;
    int tables_end = (rv + 4 * d + 15) & ~15;
    if (deep) {
      // the digit histogram follows the tables; the winners (cand: kk keys,
      // next_pow2(kk) when sorted in place) overlay both once the selection
      // has found its cut (as DeepLayout)
      hist = tables_end;
      tables_end = hist + 4 * kTopBins;
      cand = 0;
      const int out_end = 8 * (sorted ? next_pow2(kk) : kk);
      if (out_end > tables_end) tables_end = (out_end + 15) & ~15;
    }
    poff = tables_end;
    loff = (poff + 4 * (maxp + 1) + 15) & ~15;
    base = loff + 8 * maxp;
    const int after = (base + 4 * maxp + 15) & ~15;
    sh = 0;
    if (deep) {
      tmp = cand;
      scr = cand;
      st = after;
      total = st + (int)sizeof(DeepShared);
      if (q8) {
        qt = (total + 15) & ~15;
        total = qt + 2 * 16 * 4;
      }
    } else if (mode == 2) {  // running candidate set with threshold flushes
      cand = after;
      tmp = cand;
      scr = cand;
      hist = cand + 8 * (kk + kThreshAppend);
      sh = hist + 4 * 2048;
      st = sh + 64;
      total = st + (int)sizeof(QueryScanShared);
    } else {
      cand = after;
      tmp = cand + 8 * (kk + kCandCap);
      scr = tmp + 8 * next_pow2(kk);
      hist = scr;
      st = (scr + (int)sizeof(SelectScratch64) + 15) & ~15;
      total = st + (int)sizeof(QueryScanShared);
    }
  }
};

// Bin of a digit histogram hist[0..nb) (nb a multiple of kNT) holding the
// remaining-th key: returns it with the count before it and its size.
// sh: 16 ints of scratch (block_excl_scan uses 0..kNW, this 8..10).
__device__ int find_bin_sh(const int *hist, int nb, int remaining, int *sh, int &before, int &count) {
  const int tid = threadIdx.x;
  const int per = nb / kNT;
  int local = 0;
  for (int b = 0; b < per; b++) local += hist[tid * per + b];
  int total;
  const int excl = block_excl_scan(local, sh, total);
  if (excl < remaining && remaining <= excl + local) {
    int c = excl;
    for (int b = 0; b < per; b++) {
      const int h = hist[tid * per + b];
      if (c + h >= remaining) {
        sh[8] = tid * per + b;
        sh[9] = c;
        sh[10] = h;
        break;
      }
      c += h;
    }
  }
  __syncthreads();
  const int bin = sh[8];
  before = sh[9];
  count = sh[10];
  __syncthreads();
  return bin;
}

// Running candidate set for large k * k_factor: keeps the kk smallest of the
// unique keys cand[0..n) (n <= kMaxCandidates + kThreshAppend, sentinels
// allowed) in cand[0..kk), unordered and sentinel padded, and returns the
// kk-th smallest key (kSentinel64 while fewer than kk are valid). Radix
// select on the distance half (11/11/10-bit digits), then, when the kk-th
// distance is shared across the cut, on the ids carrying it. No sort.
__device__ uint64_t threshold_compact(uint64_t *cand, int n, int kk, int *hist, int *sh) {
  const int tid = threadIdx.x;
  uint64_t r[kThreshJ];
  int mine = 0;
#pragma unroll
  for (int j = 0; j < kThreshJ; j++) {
    const int i = tid + j * kNT;
    r[j] = i < n ? cand[i] : kSentinel64;
    mine += r[j] != kSentinel64;
  }
  const int valid = block_sum(mine, sh);
  uint64_t thr = kSentinel64;
  if (valid > kk) {
    const int shifts[3] = {21, 10, 0};
    const int nbits[3] = {11, 11, 10};
    int remaining = kk, before = 0, count = 0;
    uint32_t prefix = 0u, mask = 0u;
#pragma unroll
    for (int pass = 0; pass < 3; pass++) {
      const int nb = 1 << nbits[pass];
      const uint32_t dm = (uint32_t)nb - 1u;
      for (int b = tid; b < nb; b += kNT) hist[b] = 0;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kThreshJ; j++) {
        const uint32_t hi = (uint32_t)(r[j] >> 32);
        if (r[j] != kSentinel64 && (hi & mask) == prefix) atomicAdd(&hist[(hi >> shifts[pass]) & dm], 1);
      }
      __syncthreads();
      const int bin = find_bin_sh(hist, nb, remaining, sh, before, count);
      remaining -= before;
      prefix |= (uint32_t)bin << shifts[pass];
      mask |= dm << shifts[pass];
    }
    uint32_t tid_cut = 0xFFFFFFFFu;
    if (remaining < count) {  // ties at the kk-th distance: lower ids win
      uint32_t iprefix = 0u, imask = 0u;
#pragma unroll
      for (int pass = 0; pass < 3; pass++) {
        const int nb = 1 << nbits[pass];
        const uint32_t dm = (uint32_t)nb - 1u;
        for (int b = tid; b < nb; b += kNT) hist[b] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kThreshJ; j++) {
          const uint32_t hi = (uint32_t)(r[j] >> 32), lo = (uint32_t)r[j];
          if (r[j] != kSentinel64 && hi == prefix && (lo & imask) == iprefix)
            atomicAdd(&hist[(lo >> shifts[pass]) & dm], 1);
        }
        __syncthreads();
        const int bin = find_bin_sh(hist, nb, remaining, sh, before, count);
        remaining -= before;
        iprefix |= (uint32_t)bin << shifts[pass];
        imask |= dm << shifts[pass];
      }
      tid_cut = iprefix;
    }
    thr = ((uint64_t)prefix << 32) | tid_cut;
  }
  if (tid == 0) sh[11] = 0;
  __syncthreads();  // every read of cand is in registers
#pragma unroll
  for (int j = 0; j < kThreshJ; j++) {
    if (r[j] != kSentinel64 && r[j] <= thr) {
      const int slot = atomicAdd(&sh[11], 1);
      if (slot < kk) cand[slot] = r[j];
    }
  }
  __syncthreads();
  const int cnt = min(sh[11], kk);
  for (int i = cnt + tid; i < kk; i += kNT) cand[i] = kSentinel64;
  __syncthreads();
  return thr;
}

// Four waves per SIMD for the L2 deep-pool form (its few spills are outside
// the scan loop); three for the others, whose scan loops would spill at four.
// Q8 (L2 deep pool, fixed M, exact rerank after the scan): 8-bit tables.
// OFF32 (with Q8): code bytes, entry terms and pool segments addressed by
// 32-bit offsets from their base (the host checks that they fit).
#define DEEP_MODE_L2(IP, MODE) (!(IP) && (MODE) == 1)
template <int MF, bool IP, int MODE, bool Q8 = false, bool OFF32 = false>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(!IP && MODE == 1 ? 4 : 3))) void ivfpq_stream_scan_kernel(
    const float *__restrict__ q, int64_t nq, int d, int M, const float *__restrict__ coarse,
    const float *__restrict__ pq, const float *__restrict__ pre,
    const int64_t *__restrict__ list_off, const uint8_t *__restrict__ codes,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ probes, int maxp,
    const int32_t *__restrict__ nprobe_q, int nprobe_all, int kk, int k,
    const uint8_t *__restrict__ excl, int64_t excl_len, uint64_t *__restrict__ out_keys,
    float *__restrict__ D, int64_t *__restrict__ I, float pad_dist, uint32_t *__restrict__ seg_base,
    int64_t seg_cap, unsigned *__restrict__ counter, unsigned long long *__restrict__ pool_writes,
    const int32_t *__restrict__ order, const float *__restrict__ mu,
    const unsigned *__restrict__ nq_dev) {
  extern __shared__ __align__(16) unsigned char smem[];
  // nq_dev (DEEP, nullable): only order[0..*nq_dev) is scanned.
  // Q8: pre holds the centred entry terms and the tables are taken over
  // q - mu (mu: column means of the coarse centroids).
  // MODE 0: running top-kk, rank / radix merge per flush; 1: deep pool in
  // HBM (DEEP); 2: running candidate set with threshold flushes (THRESH).
  constexpr bool DEEP = MODE == 1;
  constexpr bool THRESH = MODE == 2;
  static_assert(!Q8 || (DEEP_MODE_L2(IP, MODE) && MF > 0), "8-bit tables: L2 deep pool, fixed M");
  static_assert(!OFF32 || Q8, "32-bit offsets: the 8-bit table form");
  const StreamLayout lay(M, d, kk, maxp, MODE, out_keys == nullptr, Q8);
  float *lut = reinterpret_cast<float *>(smem + lay.lut);
  float *rv = reinterpret_cast<float *>(smem + lay.rv);
  int *poff = reinterpret_cast<int *>(smem + lay.poff);
  int64_t *loff = reinterpret_cast<int64_t *>(smem + lay.loff);
  float *basev = reinterpret_cast<float *>(smem + lay.base);
  uint64_t *cand = reinterpret_cast<uint64_t *>(smem + lay.cand);
  uint64_t *tmp = reinterpret_cast<uint64_t *>(smem + lay.tmp);
  SelectScratch64 *scr = reinterpret_cast<SelectScratch64 *>(smem + lay.scr);
  int *hist = reinterpret_cast<int *>(smem + lay.hist);
  DeepShared &ds = *reinterpret_cast<DeepShared *>(smem + lay.st);
  QueryScanShared *st = reinterpret_cast<QueryScanShared *>(smem + lay.st);
  int *tsh = reinterpret_cast<int *>(smem + lay.sh);
  const int tid = threadIdx.x, nlut = M * kKsub, dsub = d / M;
  constexpr bool kMixed = !IP;
  // DEEP: this workgroup's compacted segment of (key << 32 | entry) words
  uint64_t *seg = DEEP ? reinterpret_cast<uint64_t *>(seg_base) + (int64_t)blockIdx.x * seg_cap : nullptr;
  const int wave = tid >> 6, lane = tid & 63;

  for (;;) {
    int64_t qi;
    if constexpr (DEEP) {
      __syncthreads();
      if (tid == 0) {
        if (nq_dev) {  // a listed subset: order[0..*nq_dev) in the same ranges
          const int64_t nl = min<int64_t>(nq, (int64_t)*nq_dev);
          const int r = blockIdx.x % kQueryRanges;
          const int64_t beg = nl * r / kQueryRanges, end = nl * (r + 1) / kQueryRanges;
          const int64_t t = beg + (int64_t)atomicAdd(&counter[r], 1u);
          ds.q = t < end ? (int64_t)order[t] : nq;
        } else {
          ds.q = next_query(order, nq, counter);
        }
      }
      __syncthreads();
      qi = ds.q;
      if (qi >= nq) break;
    } else {
      qi = blockIdx.x;
    }
    const float *qq = q + qi * d;
    int np = nprobe_q ? nprobe_q[qi] : nprobe_all;
    np = min(max(np, 1), maxp);

    // per-query LUT
    for (int j = tid; j < d; j += kNT) rv[j] = qq[j];
    if constexpr (DEEP) {
      if constexpr (!Q8)  // Q8: the histogram's space stages the FP32 tables first
        for (int b = tid; b < kTopBins; b += kNT) hist[b] = 0;
      if (tid == 0) {
        ds.nseg = 0;
        ds.bound = kNoBound;
      }
    } else {
      for (int i = tid; i < kk; i += kNT) cand[i] = kSentinel64;
      if (tid == 0) {
        st->cnt = 0;
        st->thr = kSentinel64;
      }
    }
    __syncthreads();
    float *ftab = Q8 ? reinterpret_cast<float *>(hist) : lut;
    for (int e = tid; e < nlut; e += kNT) {
      const int m = e >> 8;
      const float *c = pq + (int64_t)e * dsub;
      if (IP) {
        ftab[e] = -ip_pq(rv + m * dsub, c, dsub);  // pq.mojo:285-292
      } else if (Q8) {
        float dot = 0.0f;
        for (int j = 0; j < dsub; j++) dot = fma_(rv[m * dsub + j] - mu[m * dsub + j], c[j], dot);
        ftab[e] = -2.0f * dot;
      } else {
        float dot = 0.0f;
        for (int j = 0; j < dsub; j++) dot = fma_(rv[m * dsub + j], c[j], dot);
        ftab[e] = -2.0f * dot;
      }
    }
    // Q8: each subquantizer's table minus its minimum, in steps of one
    // scale per query (the largest range over 255), rounded to bytes. A sum
    // of M table values is then q8off + q8scale * (sum of the bytes), within
    // M * q8scale / 2; the exact rerank recomputes the distances it returns.
    float q8scale = 1.0f, q8off = 0.0f;
    if constexpr (Q8) {
      __syncthreads();
      float *tmin = reinterpret_cast<float *>(smem + lay.qt), *trng = tmin + 16;
      for (int m = wave; m < MF; m += kNW) {
        float mn = ftab[m * kKsub + lane], mx = mn;
#pragma unroll
        for (int c = lane + 64; c < kKsub; c += 64) {
          const float v = ftab[m * kKsub + c];
          mn = fminf(mn, v);
          mx = fmaxf(mx, v);
        }
        mn = wave_minf(mn);
        mx = wave_maxf(mx);
        if (lane == 0) {
          tmin[m] = mn;
          trng[m] = mx - mn;
        }
      }
      __syncthreads();
      float big = 0.0f;
#pragma unroll
      for (int m = 0; m < MF; m++) {
        big = fmaxf(big, trng[m]);
        q8off += tmin[m];
      }
      const float inv = big > 0.0f ? 255.0f / big : 0.0f;
      q8scale = big > 0.0f ? big / 255.0f : 1.0f;
      uint8_t *tab8 = smem + lay.lut;
      for (int e = tid; e < nlut; e += kNT)
        tab8[e] = (uint8_t)min(255, (int)rintf((ftab[e] - tmin[e >> 8]) * inv));
      __syncthreads();
      for (int b = tid; b < kTopBins; b += kNT) hist[b] = 0;
    }
    // probe table: stream offsets, CSR starts, per-list base
    int carry = 0;
    for (int p0 = 0; p0 < np; p0 += kNT) {
      const int p = p0 + tid;
      int len = 0;
      if (p < np) {
        const int l = probes[qi * maxp + p];
        float b = 0.0f;
        int64_t off = 0;
        if (l >= 0) {
          off = list_off[l];
          len = (int)min<int64_t>(list_off[l + 1] - off, seg_cap);
          const float *cl = coarse + (int64_t)l * d;
          b = IP ? -ip_wn<8>(rv, cl, d) : l2_w8(rv, cl, d);  // :322-324 / |q - c_l|^2
          if (Q8) b += q8off;
        }
        loff[p] = off;
        basev[p] = b;
      }
      int total;
      const int ex = block_excl_scan(len, DEEP ? ds.sh : (THRESH ? tsh : reinterpret_cast<int *>(scr->sh)), total);
      if (p < np) poff[p] = carry + ex;
      carry += total;
    }
    if (tid == 0) poff[np] = carry;
    __syncthreads();
    // one query's stream positions and CSR entries fit in 32 bits (entries are
    // int32 ids), which keeps the per-entry index arithmetic 32-bit
    const int n = DEEP ? (int)min<int64_t>(carry, seg_cap) : carry;

    // scan the stream. Each thread's positions only grow, so it keeps its
    // current probe in registers -- the stream position where it ends, the
    // CSR offset of its entries and its base term -- and reads the probe
    // table only when it crosses into the next list.
    int cur = 0;
    int cur_end = poff[1], cur_delta = (int)(loff[0] - poff[0]);
    float cur_base = basev[0];
    constexpr int U = (MF > 0) ? 4 : 1;
    int round = 0;
    // One round: U positions per thread. FULL: every position of the round is
    // below n, so nothing is tested per position; EX: the exclusion mask is
    // set. Both are fixed per call of the body, which keeps the position and
    // mask tests out of the rounds that do not need them.
    // DEEP: a round's appends (segment slot, -1 for none) wait in registers
    // until the next round has issued its loads
    int pend_slot[U];
    uint32_t pend_key[U], pend_ev[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      pend_slot[u] = -1;
      pend_key[u] = 0u;
      pend_ev[u] = 0u;
    }
    auto scan_round = [&](const int r0, auto full_c, auto ex_c) __attribute__((always_inline)) {
      constexpr bool FULL = decltype(full_c)::value;
      constexpr bool EX = decltype(ex_c)::value;
      const uint64_t thr = DEEP ? kSentinel64 : st->thr;
      // DEEP: keys whose top digit exceeds this bound cannot be among the kk
      // smallest; only the others are written and counted.
      uint32_t bound = kNoBound;
      if constexpr (DEEP) bound = __hip_atomic_load(&ds.bound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      Code<MF> cw[U];
      int32_t idv[U];
      float initv[U], prev[U];
      uint32_t ev[U];
      // Entry of each of the U positions (the probe cursor moves at list
      // crossings only); positions past n read entry 0, which exists whenever
      // n > 0, and are dropped below. The loads are issued unconditionally
      // and nothing uses them until all U are in flight: a load inside the
      // position test made the compiler wait for each entry before the next.
      // A lane whose last position of the round is still in its current
      // list takes all U from the registers; only a lane that crosses reads
      // the probe table.
      const int jlast = FULL ? r0 + (U - 1) * kNT + tid : min(r0 + (U - 1) * kNT + tid, n - 1);
      if (jlast >= cur_end) {
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int j = r0 + u * kNT + tid;
          ev[u] = 0;
          initv[u] = 0.0f;
          if (FULL || j < n) {
            while (j >= cur_end) {
              cur++;
              cur_end = poff[cur + 1];
              cur_delta = (int)(loff[cur] - poff[cur]);
              cur_base = basev[cur];
            }
            ev[u] = (uint32_t)(j + cur_delta);
            initv[u] = cur_base;
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int j = r0 + u * kNT + tid;
          ev[u] = (FULL || j < n) ? (uint32_t)(j + cur_delta) : 0u;
          initv[u] = (FULL || j < n) ? cur_base : 0.0f;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {  // every load of the batch in flight first
        if constexpr (OFF32) {
          cw[u] = load_code32<MF>(codes, ev[u]);
          idv[u] = EX ? *reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(ids) + (uint32_t)(ev[u] << 2)) : -1;
          prev[u] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(pre) + (uint32_t)(ev[u] << 2));
        } else {
          if (MF > 0) cw[u] = load_code<MF>(codes, ev[u]);
          else cw[u] = Code<MF>{};
          idv[u] = (!DEEP || EX) ? ids[ev[u]] : -1;
          prev[u] = kMixed ? pre[ev[u]] : 0.0f;
        }
      }
      // DEEP: the previous round's appends, issued behind this round's loads.
      // Stored at the end of their own round, their data registers went to
      // the next loads, which then waited for the stores to complete.
      if constexpr (DEEP) {
#pragma unroll
        for (int u = 0; u < U; u++)
          if (pend_slot[u] >= 0) seg_put<OFF32>(seg, pend_slot[u], pend_key[u], pend_ev[u]);
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (kMixed) initv[u] = initv[u] + prev[u];
      // Fixed M: the U entries' sums without branches (entries past n score
      // zero codes and are dropped below), so their lookups interleave.
      float distv[U];
      if constexpr (Q8) {
        uint32_t sv[U];
        adc_q8_multi<MF, U>(cw, sv);
#pragma unroll
        for (int u = 0; u < U; u++) distv[u] = __builtin_fmaf(q8scale, (float)sv[u], initv[u]);
      } else if constexpr (MF > 0) {
        adc_words_multi<MF, U>(cw, lut, initv, distv);
      }
      uint32_t putk[U];  // DEEP: keys to append (sentinel: none)
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int j = r0 + u * kNT + tid;
        bool put = false;
        uint32_t key = kSentinel;
        if (FULL || j < n) {
          const int32_t id = idv[u];
          const bool skip = EX && id >= 0 && id < excl_len && excl[id] != 0;  // :328-334
          float dist = 0.0f;
          if (!skip) {
            if constexpr (MF > 0) dist = distv[u];
            else dist = adc_sum<0>(codes + (int64_t)ev[u] * M, lut, M, initv[u]);
          }
          if constexpr (DEEP) {
            if (!skip) {
              key = f2key(dist);
              put = (key >> kTopShift) <= bound;
              if (put) atomicAdd(&hist[key >> kTopShift], 1);
            }
          } else if (!skip) {
            const uint64_t pk = pack_key(dist, (uint32_t)id);
            if (pk < thr) cand[kk + atomicAdd(&st->cnt, 1)] = pk;
          }
        }
        putk[u] = put ? key : kSentinel;
      }
      if constexpr (DEEP) {
        // appends: one LDS atomic per wavefront and round; the keys of one u
        // land in consecutive words in lane order
        uint64_t mu[U];
        int wtotal = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
          mu[u] = __ballot(putk[u] != kSentinel);
          wtotal += (int)__popcll(mu[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) pend_slot[u] = -1;
        if (wtotal > 0) {
          int at = 0;
          if (lane == 0) at = atomicAdd(&ds.nseg, wtotal);
          at = __shfl(at, 0, 64);
#pragma unroll
          for (int u = 0; u < U; u++) {
            if (putk[u] != kSentinel) {
              const int off = (int)__builtin_amdgcn_mbcnt_hi(
                  (uint32_t)(mu[u] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mu[u], 0u));
              pend_slot[u] = at + off;
              pend_key[u] = putk[u];
              pend_ev[u] = ev[u];
            }
            at += (int)__popcll(mu[u]);
          }
        }
        // the waves take turns tightening the bound, one refresh per four rounds
        ++round;
        if ((round & 3) == 0 && ((round >> 2) & (kNW - 1)) == wave) {
          const uint32_t b = wave_digit_bound(hist, kk);
          if (lane == 0) __hip_atomic_store(&ds.bound, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      if constexpr (THRESH) {
        __syncthreads();
        if (st->cnt > kThreshAppend - kNT * U) {
          const uint64_t t = threshold_compact(cand, kk + st->cnt, kk, hist, tsh);
          if (tid == 0) {
            st->cnt = 0;
            st->thr = t;
          }
          __syncthreads();
        }
      } else if constexpr (!DEEP) {
        __syncthreads();
        if (st->cnt > kCandCap - kNT * U) flush_candidates(cand, tmp, kk, *st, *scr);
      }
    };
    // rounds whose U x kNT positions all lie below n, then the partial one
    const int nfull = n - n % (kNT * U);
    if (excl) {
      for (int r0 = 0; r0 < nfull; r0 += kNT * U) scan_round(r0, BoolC<true>{}, BoolC<true>{});
      if (nfull < n) scan_round(nfull, BoolC<false>{}, BoolC<true>{});
    } else {
      for (int r0 = 0; r0 < nfull; r0 += kNT * U) scan_round(r0, BoolC<true>{}, BoolC<false>{});
      if (nfull < n) scan_round(nfull, BoolC<false>{}, BoolC<false>{});
    }
    if constexpr (DEEP) {  // the last round's appends
#pragma unroll
      for (int u = 0; u < U; u++)
        if (pend_slot[u] >= 0) seg_put<OFF32>(seg, pend_slot[u], pend_key[u], pend_ev[u]);
    }

    if constexpr (!DEEP) {
      if constexpr (THRESH) {
        __syncthreads();
        threshold_compact(cand, kk + st->cnt, kk, hist, tsh);
        const int np2 = next_pow2(kk);  // cand holds kk + kThreshAppend >= np2 keys
        for (int i = kk + tid; i < np2; i += kNT) cand[i] = kSentinel64;
        __syncthreads();
        bitonic_sort_u64(cand, np2);
      } else {
        if (st->cnt > 0) flush_candidates(cand, tmp, kk, *st, *scr);
      }
      __syncthreads();
      if (out_keys) {
        for (int t = tid; t < kk; t += kNT) out_keys[qi * kk + t] = cand[t];
      } else {
        for (int t = tid; t < k; t += kNT) {
          const uint64_t kv = cand[t];
          float dv = pad_dist;
          int64_t iv = -1;
          if (kv != kSentinel64) {
            dv = key2f((uint32_t)(kv >> 32));
            if (IP) dv = -dv;  // :357-360
            iv = (int64_t)(uint32_t)(kv & 0xFFFFFFFFu);
          }
          D[qi * k + t] = dv;
          I[qi * k + t] = iv;
        }
      }
      break;
    } else {
      __syncthreads();
      const int64_t nseg = ds.nseg;  // every key at or below the final bound is here
      if (pool_writes && tid == 0) atomicAdd(pool_writes, (unsigned long long)nseg);
      deep_select_segment(seg, nseg, ids, kk, hist, ds, cand);
      __syncthreads();
      const int cnt = min(ds.cnt, kk);
      const int np2 = next_pow2(kk);
      // the rerank that reads out_keys selects over all of them: no sort
      const int fill_to = out_keys ? kk : np2;
      for (int i = cnt + tid; i < fill_to; i += kNT) cand[i] = kSentinel64;
      __syncthreads();
      if (!out_keys) bitonic_sort_u64(cand, np2);
      if (out_keys) {
        for (int t = tid; t < kk; t += kNT) out_keys[qi * kk + t] = cand[t];
      } else {
        for (int t = tid; t < k; t += kNT) {
          const uint64_t kv = cand[t];
          float dv = pad_dist;
          int64_t iv = -1;
          if (kv != kSentinel64) {
            dv = key2f((uint32_t)(kv >> 32));
            if (IP) dv = -dv;
            iv = (int64_t)(uint32_t)(kv & 0xFFFFFFFFu);
          }
          D[qi * k + t] = dv;
          I[qi * k + t] = iv;
        }
      }
    }
  }
}

// pre[e] = sum_m T[l][m][code_m] for every stored entry e of list l (mixed
// L2 stream scan), summed in m order.
__global__ void entry_terms_kernel(const int64_t *__restrict__ list_off, int nlist,
                                   const uint8_t *__restrict__ codes, int M,
                                   const float *__restrict__ ptab, int64_t ntotal,
                                   float *__restrict__ pre) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntotal) return;
  int lo = 0, hi = nlist;  // list l with list_off[l] <= e < list_off[l + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (list_off[mid] <= e) lo = mid; else hi = mid;
  }
  const float *t = ptab + (int64_t)lo * M * kKsub;
  const uint8_t *c = codes + e * M;
  float acc = 0.0f;
  for (int m = 0; m < M; m++) acc += t[m * kKsub + c[m]];
  pre[e] = acc;
}

// Column means of the coarse centroids: the 8-bit tables' centre.
__global__ void column_mean_kernel(const float *__restrict__ x, int n, int d, float *__restrict__ mu) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d) return;
  double acc = 0.0;
  for (int i = 0; i < n; i++) acc += x[(int64_t)i * d + j];
  mu[j] = n > 0 ? (float)(acc / n) : 0.0f;
}

// Entry terms of the 8-bit tables: pre[e] - 2 <mu, r_e> with r_e the
// entry's PQ reconstruction, so that pre8[e] + sum_m -2 <q_m - mu_m, r_e,m>
// is the same distance with tables over q - mu (about half the range of
// tables over q for data away from the origin).
__global__ void entry_terms_centered_kernel(const uint8_t *__restrict__ codes, int M,
                                            const float *__restrict__ pq, int dsub,
                                            const float *__restrict__ mu,
                                            const float *__restrict__ pre, int64_t n,
                                            float *__restrict__ pre8) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint8_t *c = codes + e * M;
  float acc = 0.0f;
  for (int m = 0; m < M; m++) {
    const float *cw = pq + ((int64_t)m * kKsub + c[m]) * dsub;
    for (int j = 0; j < dsub; j++) acc = fma_(mu[m * dsub + j], cw[j], acc);
  }
  pre8[e] = pre[e] - 2.0f * acc;
}

// T[l][m][k] = |c_mk|^2 + 2 <c_l,m, c_mk> (mixed L2 tables).
__global__ void precompute_tables_kernel(const float *__restrict__ coarse, const float *__restrict__ pq,
                                         int d, int M, int nlist, float *__restrict__ ptab) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nlut = (int64_t)M * kKsub;
  if (t >= (int64_t)nlist * nlut) return;
  const int l = (int)(t / nlut);
  const int e = (int)(t - (int64_t)l * nlut);
  const int m = e >> 8;
  const int dsub = d / M;
  const float *c = pq + (int64_t)e * dsub;
  const float *cl = coarse + (int64_t)l * d + (int64_t)m * dsub;
  float nn = 0.0f, dot = 0.0f;
  for (int j = 0; j < dsub; j++) {
    nn = fma_(c[j], c[j], nn);
    dot = fma_(cl[j], c[j], dot);
  }
  ptab[t] = nn + 2.0f * dot;
}

__global__ void scanned_entries_kernel(const int32_t *__restrict__ probes, int64_t nq, int maxp,
                                       const int32_t *__restrict__ nprobe_q, int nprobe_all,
                                       const int64_t *__restrict__ list_off,
                                       unsigned long long *__restrict__ total) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (t < nq * maxp) {
    const int64_t qi = t / maxp;
    const int p = (int)(t - qi * maxp);
    int np = nprobe_q ? nprobe_q[qi] : nprobe_all;
    np = min(max(np, 1), maxp);
    const int l = probes[t];
    if (p < np && l >= 0) v = (unsigned long long)(list_off[l + 1] - list_off[l]);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(total, v);
}

template <int MF, bool IP>
void launch_scan(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk, uint64_t *part,
                 hipStream_t s) {
  const size_t sm = sizeof(float) * a.M * kKsub + sizeof(float) * ((a.d + 4) & ~3) +
                    sizeof(uint32_t) * (kk + kScanChunk) + sizeof(uint32_t) * kk +
                    sizeof(uint32_t) * ((kk + 3) & ~3) + sizeof(SelectScratch);
  if (sm > 150 * 1024) throw Error(kInvalid, "IVF-PQ scan workspace exceeds LDS");
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute((const void *)ivfpq_scan_kernel<MF, IP>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  const int64_t nblocks = a.nq * maxp;
  hipLaunchKernelGGL((ivfpq_scan_kernel<MF, IP>), dim3((unsigned)nblocks), dim3(kNT), sm, s, a.q,
                     a.d, a.M, a.coarse, a.pq, a.list_off, a.codes, a.ids, probes, maxp,
                     a.nprobe_per_query, a.nprobe, kk, a.excl, a.excl_len, part);
  MVG_LAUNCH_CHECK("ivfpq_scan_kernel");
}

template <int MF, bool IP, bool MIXED, int DF, bool WQ>
void launch_query_scan_wq(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                          uint64_t *out_keys, const WorkQueue &wq, hipStream_t s) {
  const QueryScanLayout lay(a.M, a.d, kk, MIXED && !IP);
  const size_t sm = (size_t)lay.total;
  if (sm > 150 * 1024) throw Error(kInvalid, "IVF-PQ scan workspace exceeds LDS");
  const void *fn = (const void *)ivfpq_query_scan_kernel<MF, IP, MIXED, DF, WQ>;
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  unsigned grid = (unsigned)a.nq;
  if (WQ) {  // persistent: enough workgroups to fill every CU
    int per_cu = 0;
    MVG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kNT, sm));
    grid = (unsigned)std::max(1, per_cu) * (unsigned)device_cus();
  }
  hipLaunchKernelGGL((ivfpq_query_scan_kernel<MF, IP, MIXED, DF, WQ>), dim3(grid), dim3(kNT), sm, s,
                     a.q, a.d, a.M, a.coarse, a.pq, a.ptab, a.list_off, a.codes, a.ids, probes, maxp,
                     a.nprobe_per_query, a.nprobe, kk, a.k, a.excl, a.excl_len, out_keys, a.D, a.I,
                     kIvfPadDist, wq);
  MVG_LAUNCH_CHECK("ivfpq_query_scan_kernel");
}

template <int MF, bool IP, bool MIXED, int DF>
void launch_query_scan_df(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                          uint64_t *out_keys, const WorkQueue *wq, hipStream_t s) {
  if (wq) launch_query_scan_wq<MF, IP, MIXED, DF, true>(a, maxp, probes, kk, out_keys, *wq, s);
  else launch_query_scan_wq<MF, IP, MIXED, DF, false>(a, maxp, probes, kk, out_keys, WorkQueue{}, s);
}

// Register-resident codebooks for the plain L2 shapes (M, d) in
// {(16, 128), (12, 96), (8, 64), (16, 64), (32, 128)}; other shapes read the
// codebook through the cache.
template <int MF, bool IP, bool MIXED>
void launch_query_scan(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                       uint64_t *out_keys, const WorkQueue *wq, hipStream_t s) {
  if constexpr (!IP && !MIXED && MF == 16) {
    if (a.d == 128) return launch_query_scan_df<MF, IP, MIXED, 128>(a, maxp, probes, kk, out_keys, wq, s);
    if (a.d == 64) return launch_query_scan_df<MF, IP, MIXED, 64>(a, maxp, probes, kk, out_keys, wq, s);
  }
  if constexpr (!IP && !MIXED && MF == 12) {
    if (a.d == 96) return launch_query_scan_df<MF, IP, MIXED, 96>(a, maxp, probes, kk, out_keys, wq, s);
  }
  if constexpr (!IP && !MIXED && MF == 8) {
    if (a.d == 64) return launch_query_scan_df<MF, IP, MIXED, 64>(a, maxp, probes, kk, out_keys, wq, s);
  }
  launch_query_scan_df<MF, IP, MIXED, 0>(a, maxp, probes, kk, out_keys, wq, s);
}

// Builds the query order and clears the range counters of a persistent
// deep-pool scan; returns the order.
const int32_t *build_query_order(const IvfSearchArgs &a, int maxp, const int32_t *probes,
                                 IvfWork &w, hipStream_t s, const int32_t *rank = nullptr) {
  w.ord_cnt.reserve(sizeof(int) * (size_t)a.nlist);
  w.ord_off.reserve(sizeof(int) * (size_t)a.nlist);
  w.ord.reserve(sizeof(int32_t) * (size_t)a.nq);
  w.deep_counter.reserve(sizeof(unsigned) * kQueryRanges);
  w.ord_ready = true;
  hipLaunchKernelGGL(fill_zero_kernel, dim3(ceil_div(a.nlist, 256)), dim3(256), 0, s,
                     w.ord_cnt.as<int>(), (int64_t)a.nlist);
  MVG_LAUNCH_CHECK("fill_zero_kernel");
  hipLaunchKernelGGL(order_count_kernel, dim3(ceil_div(a.nq, 256)), dim3(256), 0, s, probes, a.nq,
                     maxp, a.nlist, w.ord_cnt.as<int>(), rank);
  MVG_LAUNCH_CHECK("order_count_kernel");
  hipLaunchKernelGGL(order_offsets_kernel, dim3(1), dim3(kNT), 0, s, w.ord_cnt.as<int>(), a.nlist,
                     w.ord_off.as<int>(), w.deep_counter.as<unsigned>());
  MVG_LAUNCH_CHECK("order_offsets_kernel");
  hipLaunchKernelGGL(order_scatter_kernel, dim3(ceil_div(a.nq, 256)), dim3(256), 0, s, probes, a.nq,
                     maxp, a.nlist, w.ord_off.as<int>(), w.ord.as<int32_t>(), rank);
  MVG_LAUNCH_CHECK("order_scatter_kernel");
  return w.ord.as<int32_t>();
}

// Persistent grid: every XCD range needs workgroups, so a multiple of
// kQueryRanges (at least one per range).
int64_t persistent_grid(int per_cu, int64_t limit) {
  int64_t g = std::min<int64_t>((int64_t)std::max(1, per_cu) * device_cus(), limit);
  return std::max<int64_t>(kQueryRanges, g / kQueryRanges * kQueryRanges);
}

// Most persistent workgroups whose pool segments of `cap` words fit: 4 GiB
// of segments, or half of what is left of a fixed workspace (the query
// order and the rerank take their buffers after the segments). A workgroup
// pulls queries from the shared order, so fewer of them change only the
// schedule, not the results.
int64_t segment_grid_limit(const DevBuf &seg, int64_t cap) {
  int64_t limit = (int64_t(4) << 30) / (8 * cap);
  if (seg.arena) limit = std::min<int64_t>(limit, (int64_t)(seg.arena->left() / 2) / (8 * cap));
  return limit;
}

template <int MF, bool IP, bool MIXED, int DF>
void launch_deep_df(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                    uint64_t *out_keys, IvfWork &w, hipStream_t s) {
  const DeepLayout lay(a.M, a.d, kk, MIXED && !IP, out_keys == nullptr);
  const size_t sm = (size_t)lay.total;
  if (sm > 150 * 1024) throw Error(kInvalid, "IVF-PQ deep scan workspace exceeds LDS");
  const void *fn = (const void *)ivfpq_deep_scan_kernel<MF, IP, MIXED, DF>;
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  int per_cu = 0;
  MVG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kNT, sm));
  const int64_t cap = (std::max<int64_t>(a.deep_cap, 1) + 63) & ~int64_t(63);  // 16-byte aligned segments
  const int64_t grid = persistent_grid(per_cu, segment_grid_limit(w.deep_seg, cap));
  w.deep_seg.reserve(sizeof(uint64_t) * (size_t)(grid * cap));
  const int32_t *order = build_query_order(a, maxp, probes, w, s);
  unsigned *counter = w.deep_counter.as<unsigned>();
  hipLaunchKernelGGL((ivfpq_deep_scan_kernel<MF, IP, MIXED, DF>), dim3((unsigned)grid), dim3(kNT), sm,
                     s, a.q, a.nq, a.d, a.M, a.coarse, a.pq, a.ptab, a.list_off, a.codes, a.ids,
                     probes, maxp, a.nprobe_per_query, a.nprobe, kk, a.k, a.excl, a.excl_len,
                     out_keys, a.D, a.I, kIvfPadDist, w.deep_seg.as<uint32_t>(), cap, counter, order);
  MVG_LAUNCH_CHECK("ivfpq_deep_scan_kernel");
}

// Stream scan (plain IP, mixed L2). Returns false when the shape does not
// fit its LDS layout; the caller then takes the per-list kernels.
template <int MF, bool IP, int MODE, bool Q8, bool OFF32 = false>
bool launch_stream_t(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                     uint64_t *out_keys, IvfWork &w, hipStream_t s) {
  constexpr bool DEEP = MODE == 1;
  const StreamLayout lay(a.M, a.d, kk, maxp, MODE, out_keys == nullptr, Q8);
  const size_t sm = (size_t)lay.total;
  if (sm > 150 * 1024 || maxp > kStreamMaxProbes) return false;
  const void *fn = (const void *)ivfpq_stream_scan_kernel<MF, IP, MODE, Q8, OFF32>;
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  int64_t grid = a.nq, cap = INT32_MAX;
  uint32_t *seg = nullptr;
  unsigned *counter = nullptr;
  const int32_t *order = nullptr;
  if (DEEP) {
    int per_cu = 0;
    MVG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kNT, sm));
    cap = (std::max<int64_t>(a.deep_cap, 1) + 63) & ~int64_t(63);  // 16-byte aligned segments
    grid = persistent_grid(per_cu, segment_grid_limit(w.deep_seg, cap));
    w.deep_seg.reserve(sizeof(uint64_t) * (size_t)(grid * cap));
    seg = w.deep_seg.as<uint32_t>();
    order = build_query_order(a, maxp, probes, w, s);
    counter = w.deep_counter.as<unsigned>();
  }
  hipLaunchKernelGGL((ivfpq_stream_scan_kernel<MF, IP, MODE, Q8, OFF32>), dim3((unsigned)grid), dim3(kNT), sm, s,
                     a.q, a.nq, a.d, a.M, a.coarse, a.pq, Q8 ? a.pre8 : a.pre, a.list_off, a.codes, a.ids, probes,
                     maxp, a.nprobe_per_query, a.nprobe, kk, a.k, a.excl, a.excl_len, out_keys, a.D,
                     a.I, kIvfPadDist, seg, cap, counter,
                     DEEP && w.prof.on ? w.prof.counter.as<unsigned long long>() + 1 : nullptr, order,
                     Q8 ? a.mu : nullptr, nullptr);
  MVG_LAUNCH_CHECK("ivfpq_stream_scan_kernel");
  return true;
}

// The FP32-table deep-pool stream scan (mixed L2) over list[0..*count).
template <int MF>
void launch_deep_subset_t(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                          uint64_t *out_keys, const int32_t *list, const unsigned *count, IvfWork &w,
                          hipStream_t s) {
  const StreamLayout lay(a.M, a.d, kk, maxp, 1, false, false);
  const size_t sm = (size_t)lay.total;
  if (sm > 150 * 1024 || maxp > kStreamMaxProbes)
    throw Error(kInvalid, "IVF-PQ deep scan workspace exceeds LDS");
  const void *fn = (const void *)ivfpq_stream_scan_kernel<MF, false, 1, false, false>;
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  const int64_t cap = (std::max<int64_t>(a.deep_cap, 1) + 63) & ~int64_t(63);
  const int64_t grid = std::max<int64_t>(kQueryRanges, std::min<int64_t>(64, segment_grid_limit(w.deep_seg, cap)) /
                                                           kQueryRanges * kQueryRanges);
  w.deep_seg.reserve(sizeof(uint64_t) * (size_t)(grid * cap));
  w.deep_counter.reserve(sizeof(unsigned) * kQueryRanges);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<int *>(w.deep_counter.ptr),
                     (int64_t)kQueryRanges);
  MVG_LAUNCH_CHECK("fill_zero_kernel");
  hipLaunchKernelGGL((ivfpq_stream_scan_kernel<MF, false, 1, false, false>), dim3((unsigned)grid), dim3(kNT), sm,
                     s, a.q, a.nq, a.d, a.M, a.coarse, a.pq, a.pre, a.list_off, a.codes, a.ids, probes, maxp,
                     a.nprobe_per_query, a.nprobe, kk, a.k, a.excl, a.excl_len, out_keys, a.D, a.I,
                     kIvfPadDist, w.deep_seg.as<uint32_t>(), cap, w.deep_counter.as<unsigned>(),
                     w.prof.on ? w.prof.counter.as<unsigned long long>() + 1 : nullptr, list, nullptr, count);
  MVG_LAUNCH_CHECK("ivfpq_stream_scan_kernel");
}

// 8-bit tables only where an exact rerank follows the scan (k_factor > 1):
// the scan's distances then only choose the candidates.
template <int MF, bool IP, int MODE>
bool launch_stream(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                   uint64_t *out_keys, IvfWork &w, hipStream_t s) {
  if constexpr (DEEP_MODE_L2(IP, MODE) && MF > 0) {
    if ((a.table_bits == 8 || a.table_bits == 16) && a.pre8 && a.mu && a.k_factor > 1 && out_keys) {
      // 32-bit offsets when the codes (the largest per-entry array) and a
      // pool segment stay below 4 GiB
      const int64_t cap = (std::max<int64_t>(a.deep_cap, 1) + 63) & ~int64_t(63);
      if ((a.ncsr + 1) * MF < (int64_t(1) << 32) && cap * 8 < (int64_t(1) << 32))
        return launch_stream_t<MF, IP, MODE, true, true>(a, maxp, probes, kk, out_keys, w, s);
      return launch_stream_t<MF, IP, MODE, true>(a, maxp, probes, kk, out_keys, w, s);
    }
  }
  return launch_stream_t<MF, IP, MODE, false>(a, maxp, probes, kk, out_keys, w, s);
}

template <int MF, bool IP>
bool launch_stream_any(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk, int mode,
                       uint64_t *out_keys, IvfWork &w, hipStream_t s) {
  if (mode == 1) return launch_stream<MF, IP, 1>(a, maxp, probes, kk, out_keys, w, s);
  if (mode == 2) return launch_stream<MF, IP, 2>(a, maxp, probes, kk, out_keys, w, s);
  return launch_stream<MF, IP, 0>(a, maxp, probes, kk, out_keys, w, s);
}

template <int MF, bool IP, bool MIXED>
void launch_deep(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk, uint64_t *out_keys,
                 IvfWork &w, hipStream_t s) {
  if constexpr (!IP && !MIXED && MF == 16) {
    if (a.d == 128) return launch_deep_df<MF, IP, MIXED, 128>(a, maxp, probes, kk, out_keys, w, s);
    if (a.d == 64) return launch_deep_df<MF, IP, MIXED, 64>(a, maxp, probes, kk, out_keys, w, s);
  }
  if constexpr (!IP && !MIXED && MF == 12) {
    if (a.d == 96) return launch_deep_df<MF, IP, MIXED, 96>(a, maxp, probes, kk, out_keys, w, s);
  }
  if constexpr (!IP && !MIXED && MF == 8) {
    if (a.d == 64) return launch_deep_df<MF, IP, MIXED, 64>(a, maxp, probes, kk, out_keys, w, s);
  }
  launch_deep_df<MF, IP, MIXED, 0>(a, maxp, probes, kk, out_keys, w, s);
}

}  // namespace

const int32_t *build_tile_order(const IvfSearchArgs &a, int maxp, const int32_t *probes,
                                const int32_t *rank, IvfWork &w, hipStream_t s) {
  return build_query_order(a, maxp, probes, w, s, rank);
}

void launch_deep_subset(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                        uint64_t *out_keys, const int32_t *list, const unsigned *count, IvfWork &w,
                        hipStream_t s) {
  switch (a.M) {
    case 8: launch_deep_subset_t<8>(a, maxp, probes, kk, out_keys, list, count, w, s); break;
    case 12: launch_deep_subset_t<12>(a, maxp, probes, kk, out_keys, list, count, w, s); break;
    case 16: launch_deep_subset_t<16>(a, maxp, probes, kk, out_keys, list, count, w, s); break;
    default: launch_deep_subset_t<0>(a, maxp, probes, kk, out_keys, list, count, w, s); break;
  }
}

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    MVG_HIP(hipGetDevice(&dev));
    MVG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  return cus;
}


void precompute_ivf_tables(const float *coarse, const float *pq, int d, int M, int nlist,
                           float *ptab, hipStream_t s) {
  const int64_t total = (int64_t)nlist * M * kKsub;
  hipLaunchKernelGGL(precompute_tables_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, s, coarse, pq,
                     d, M, nlist, ptab);
  MVG_LAUNCH_CHECK("precompute_tables_kernel");
}

void precompute_centered_terms(const float *coarse, int nlist, int d, const uint8_t *codes, int M,
                               const float *pq, const float *pre, int64_t n, float *mu, float *pre8,
                               hipStream_t s) {
  hipLaunchKernelGGL(column_mean_kernel, dim3(ceil_div(d, 64)), dim3(64), 0, s, coarse, nlist, d, mu);
  MVG_LAUNCH_CHECK("column_mean_kernel");
  if (n == 0) return;
  hipLaunchKernelGGL(entry_terms_centered_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, codes, M, pq,
                     d / M, mu, pre, n, pre8);
  MVG_LAUNCH_CHECK("entry_terms_centered_kernel");
}

void precompute_entry_terms(const int64_t *list_off, int nlist, const uint8_t *codes, int M,
                            const float *ptab, int64_t ntotal, float *pre, hipStream_t s) {
  if (ntotal == 0) return;
  hipLaunchKernelGGL(entry_terms_kernel, dim3(ceil_div(ntotal, 256)), dim3(256), 0, s, list_off, nlist,
                     codes, M, ptab, ntotal, pre);
  MVG_LAUNCH_CHECK("entry_terms_kernel");
}

void rerank(const float *q, int64_t nq, int d, int metric, const float *vectors, int64_t nrows,
            const uint64_t *cand, int kk, int k, float pad_dist, float *D, int64_t *I,
            hipStream_t s, const int32_t *order, RowMap rm, const int32_t *kq) {
  if (nq == 0) return;
  const bool ip = metric == kMetricIP;
  const unsigned grid =
      (unsigned)(order ? kRerankRanges * ((nq + kRerankRanges - 1) / kRerankRanges) : nq);
  if (d >= 64 && d % 64 == 0) {
    const size_t sm = sizeof(float) * d + sizeof(uint64_t) * (size_t)(((kk + 1) & ~1) + next_pow2(k)) +
                      sizeof(SelectScratch64);
    if (sm <= 150 * 1024) {
      const bool small = d < 256;
#define MVG_RR(IPV, SV, MV)                                                                      \
  {                                                                                              \
    if (sm > 64 * 1024)                                                                          \
      MVG_HIP(hipFuncSetAttribute((const void *)rerank_group_kernel<IPV, SV, MV>,                \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));         \
    hipLaunchKernelGGL((rerank_group_kernel<IPV, SV, MV>), dim3(grid), dim3(kNT), sm, s, q, d,   \
                       vectors, nrows, cand, kk, k, pad_dist, D, I, nq, order, rm, kq);          \
  }
#define MVG_RR2(IPV, SV)                                                                         \
  if (rm.row_of) MVG_RR(IPV, SV, true) else MVG_RR(IPV, SV, false)
      if (ip) {
        if (small) MVG_RR2(true, true) else MVG_RR2(true, false)
      } else {
        if (small) MVG_RR2(false, true) else MVG_RR2(false, false)
      }
#undef MVG_RR2
#undef MVG_RR
      MVG_LAUNCH_CHECK("rerank_group_kernel");
      return;
    }
  }
  int buf_n = 1;
  while (buf_n < kk || buf_n < k) buf_n <<= 1;
  const size_t sm = sizeof(float) * ((d + 1) & ~1) + 2 * sizeof(uint64_t) * buf_n + sizeof(SelectScratch);
  if (sm > 150 * 1024) throw Error(kInvalid, "rerank workspace exceeds LDS");
#define MVG_RK(IPV, MV)                                                                          \
  {                                                                                              \
    if (sm > 64 * 1024)                                                                          \
      MVG_HIP(hipFuncSetAttribute((const void *)rerank_kernel<IPV, MV>,                          \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));         \
    hipLaunchKernelGGL((rerank_kernel<IPV, MV>), dim3(grid), dim3(kNT), sm, s, q, d, vectors,    \
                       nrows, cand, kk, k, buf_n, pad_dist, D, I, nq, order, rm, kq);            \
  }
  if (ip) {
    if (rm.row_of) MVG_RK(true, true) else MVG_RK(true, false)
  } else {
    if (rm.row_of) MVG_RK(false, true) else MVG_RK(false, false)
  }
#undef MVG_RK
  MVG_LAUNCH_CHECK("rerank_kernel");
}

int prefilter_width(int k, int kk, int per_k) {
  if (per_k <= 0) return kk;
  return (int)std::min<int64_t>(kk, std::max<int64_t>(64, (int64_t)per_k * k));
}

// Mixed-form rerank of kk candidates per query: with SQ8 row codes and
// kk > prefilter_width, the SQ8 pre-rerank narrows them first.
void rerank_mixed(const IvfSearchArgs &a, IvfWork &w, const uint64_t *keys, int kk,
                  const int32_t *order, hipStream_t s) {
  const int kk2 = prefilter_width(a.k, kk, a.prefilter);
  if (!a.vq_codes || kk2 >= kk) {
    rerank(a.q, a.nq, a.d, a.metric, a.vectors, a.ntotal, keys, kk, a.k, kIvfPadDist, a.D, a.I, s,
           order, a.rows, a.kq);
    return;
  }
  const bool ip = a.metric == kMetricIP;
  const int ld = sq8_row_stride(a.d);
  w.vq_qcodes.reserve((size_t)a.nq * ld);
  w.vq_qnorms.reserve(sizeof(uint32_t) * a.nq);
  w.vq_keys.reserve(sizeof(uint64_t) * (size_t)a.nq * kk2);
  sq8_encode(a.q, a.nq, a.d, ip ? 1 : 0, ip ? 0.0f : a.vq_min, 1.0f / a.vq_scale,
             w.vq_qcodes.as<uint8_t>(), w.vq_qnorms.as<uint32_t>(), nullptr, s);
  const size_t sm = sizeof(uint64_t) * (size_t)(((kk + 1) & ~1) + ((kk2 + 1) & ~1)) + sizeof(BucketScratch);
  if (sm > 150 * 1024) throw Error(kInvalid, "pre-rerank workspace exceeds LDS");
  const unsigned grid =
      (unsigned)(order ? kRerankRanges * ((a.nq + kRerankRanges - 1) / kRerankRanges) : a.nq);
#define MVG_PF(IPV, MV)                                                                          \
  {                                                                                              \
    if (sm > 64 * 1024)                                                                          \
      MVG_HIP(hipFuncSetAttribute((const void *)sq8_prefilter_kernel<IPV, MV>,                   \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));         \
    hipLaunchKernelGGL((sq8_prefilter_kernel<IPV, MV>), dim3(grid), dim3(kNT), sm, s,            \
                       w.vq_qcodes.as<uint8_t>(), w.vq_qnorms.as<uint32_t>(), ld, a.vq_codes,    \
                       a.vq_norms, a.ntotal, keys, kk, kk2, w.vq_keys.as<uint64_t>(), a.nq,      \
                       order, a.rows, a.kq, kk / a.k, a.prefilter);                              \
  }
  if (ip) {
    if (a.rows.row_of) MVG_PF(true, true) else MVG_PF(true, false)
  } else {
    if (a.rows.row_of) MVG_PF(false, true) else MVG_PF(false, false)
  }
#undef MVG_PF
  MVG_LAUNCH_CHECK("sq8_prefilter_kernel");
  rerank(a.q, a.nq, a.d, a.metric, a.vectors, a.ntotal, w.vq_keys.as<uint64_t>(), kk2, a.k, kIvfPadDist,
         a.D, a.I, s, order, a.rows, a.kq);
}

void ivfpq_search(const IvfSearchArgs &a, IvfWork &w, hipStream_t s) {
  if (a.nq == 0) return;
  w.ord_ready = false;  // set when a deep-pool scan builds the clustered query order
  const int kf = a.k_factor > 1 ? a.k_factor : 1;
  const int64_t kk64 = (int64_t)a.k * kf;
  if (kk64 > kMaxCandidates) throw Error(kInvalid, "k * k_factor exceeds the device limit of 4096");
  const int kk = (int)kk64;
  if (kf > 1 && !a.vectors) throw Error(kState, "k_factor rerank needs the stored vectors");
  // Largest probe count across the batch (the host clamps to nlist).
  const int maxp = min(a.nprobe, a.nlist);
  IvfProfile &pf = w.prof;
  if (pf.on) {
    for (auto &e : pf.ev)
      if (!e) MVG_HIP(hipEventCreate(&e));
    // [0] list entries the scan visits, [1] keys written to deep pools
    pf.counter.reserve(2 * sizeof(unsigned long long));
    MVG_HIP(hipMemsetAsync(pf.counter.ptr, 0, 2 * sizeof(unsigned long long), s));
    MVG_HIP(hipEventRecord(pf.ev[0], s));
  }
  // 1. coarse top-maxp (IndexFlat.search over the centroids, :264).
  w.probes.reserve(sizeof(int32_t) * (size_t)a.nq * maxp);
  int32_t *probes = w.probes.as<int32_t>();
  if (a.precision != 0 && a.coarse_bf16 && a.nlist > kMfmaCoarseMinLists) {
    // mixed: bf16 GEMM-form coarse distances on MFMA, FP32 accumulate. Below
    // kMfmaCoarseMinLists centroids the contraction is negligible and the
    // exact FP32 kernel's per-chunk radix select is the faster path.
    const int rb = gemm_row_bytes(kGemmBf16L2, a.d);
    w.qbf.reserve((size_t)a.nq * rb);
    w.qnorm.reserve(sizeof(float) * a.nq);
    to_bf16(a.q, a.nq, a.d, w.qbf.as<uint16_t>(), w.qnorm.as<float>(), s);
    int64_t ncand = 0;
    gemm_topk(a.metric == kMetricIP ? kGemmBf16IP : kGemmBf16L2, w.qbf.as<uint8_t>(),
              reinterpret_cast<const uint8_t *>(a.coarse_bf16), a.nq, a.nlist, a.d, w.qnorm.as<float>(),
              a.coarse_norm, nullptr, nullptr, nullptr, nullptr, 0.0f, maxp, nullptr, 0, w.coarse_part,
              ncand, s);
    launch_merge(w.coarse_part.as<uint64_t>(), a.nq, ncand, maxp, 0, 0.0f, nullptr,
                 nullptr, probes, nullptr, s);
  } else {
    flat_topk(a.coarse, a.nlist, a.d, a.metric, a.q, a.nq, maxp, nullptr, 0, w.coarse_part, 0.0f,
              nullptr, nullptr, probes, nullptr, s);
  }
  const bool ip = a.metric == kMetricIP;
  if (pf.on) MVG_HIP(hipEventRecord(pf.ev[1], s));
  // Per-query k (a.kq): every query is scanned for the batch's k * k_factor
  // candidates, then cut to its own kq * k_factor best (the candidate set of
  // a search with its k alone) before the rerank or the output.
  const bool per_k = a.kq != nullptr;
  if (!a.split_scan) {
    // 2'. one workgroup per query: LUT + ADC scan + running top-kk, then the
    // optional rerank.
    uint64_t *keys = nullptr;
    if (kf > 1 || per_k) {
      w.keys.reserve(sizeof(uint64_t) * (size_t)a.nq * kk);
      keys = w.keys.as<uint64_t>();
    }
    const bool mixed = a.precision != 0 && !ip;
    if (mixed && !a.ptab) throw Error(kState, "mixed-precision tables are missing");
    const bool deep = kk >= (a.deep_min > 0 ? a.deep_min : kDeepMinCandidates);
    WorkQueue wq{};
    const WorkQueue *wqp = nullptr;
    bool streamed = false;
    // FP16 MFMA tile scan: mixed L2 candidates that the exact rerank follows
    if (mixed && a.table_bits == 16 && kf > 1 && keys) streamed = launch_tile_scan(a, maxp, probes, kk, keys, w, s);
    if (!streamed && (ip || (mixed && a.pre)) && (deep || a.wq_entries == 0)) {
      // large pools: deep pool in HBM or running set with threshold flushes
      const int smode = !deep ? 0 : (a.deep_pool ? 1 : 2);
#define MVG_STREAM(MFV)                                                                       \
  streamed = ip ? launch_stream_any<MFV, true>(a, maxp, probes, kk, smode, keys, w, s)        \
                : launch_stream_any<MFV, false>(a, maxp, probes, kk, smode, keys, w, s);
      switch (a.M) {
        case 8: MVG_STREAM(8) break;
        case 12: MVG_STREAM(12) break;
        case 16: MVG_STREAM(16) break;
        default: MVG_STREAM(0) break;
      }
#undef MVG_STREAM
    }
    if (streamed) {
    } else if (deep) {
#define MVG_DEEP(MFV)                                                                \
  if (ip) launch_deep<MFV, true, false>(a, maxp, probes, kk, keys, w, s);           \
  else if (mixed) launch_deep<MFV, false, true>(a, maxp, probes, kk, keys, w, s);   \
  else launch_deep<MFV, false, false>(a, maxp, probes, kk, keys, w, s);
      switch (a.M) {
        case 8: MVG_DEEP(8) break;
        case 12: MVG_DEEP(12) break;
        case 16: MVG_DEEP(16) break;
        default: MVG_DEEP(0) break;
      }
#undef MVG_DEEP
    } else if (a.wq_entries > 0) {
      // (query, entry-range) work queue built on the device: no host sync.
      w.wq_poff.reserve(sizeof(int64_t) * (size_t)a.nq * (maxp + 1));
      w.wq_nitems.reserve(sizeof(int) * (size_t)a.nq);
      w.wq_rq.reserve(sizeof(int64_t) * (size_t)a.nq);
      w.wq_off.reserve(sizeof(int64_t) * (size_t)(a.nq + 1));
      w.wq_gthr.reserve(sizeof(uint64_t) * (size_t)a.nq);
      w.wq_counter.reserve(sizeof(unsigned));
      w.part.reserve(sizeof(uint64_t) * (size_t)a.nq * kMaxItemsPerQuery * kk);
      hipLaunchKernelGGL(wq_setup_kernel, dim3(ceil_div(a.nq, 256)), dim3(256), 0, s, probes, a.nq,
                         maxp, a.nprobe_per_query, a.nprobe, a.list_off, (int64_t)a.wq_entries,
                         w.wq_poff.as<int64_t>(), w.wq_nitems.as<int>(), w.wq_rq.as<int64_t>(),
                         w.wq_gthr.as<uint64_t>());
      MVG_LAUNCH_CHECK("wq_setup_kernel");
      hipLaunchKernelGGL(wq_offsets_kernel, dim3(1), dim3(kNT), 0, s, w.wq_nitems.as<int>(), a.nq,
                         w.wq_off.as<int64_t>(), w.wq_counter.as<unsigned>());
      MVG_LAUNCH_CHECK("wq_offsets_kernel");
      wq.nq = a.nq;
      wq.item_off = w.wq_off.as<int64_t>();
      wq.poff = w.wq_poff.as<int64_t>();
      wq.rq = w.wq_rq.as<int64_t>();
      wq.counter = w.wq_counter.as<unsigned>();
      wq.gthr = w.wq_gthr.as<uint64_t>();
      wq.part = w.part.as<uint64_t>();
      wqp = &wq;
    }
#define MVG_QSCAN(MFV)                                                                   \
  if (ip) launch_query_scan<MFV, true, false>(a, maxp, probes, kk, keys, wqp, s);       \
  else if (mixed) launch_query_scan<MFV, false, true>(a, maxp, probes, kk, keys, wqp, s); \
  else launch_query_scan<MFV, false, false>(a, maxp, probes, kk, keys, wqp, s);
    if (!deep && !streamed) {
      switch (a.M) {
        case 8: MVG_QSCAN(8) break;
        case 12: MVG_QSCAN(12) break;
        case 16: MVG_QSCAN(16) break;
        default: MVG_QSCAN(0) break;
      }
    }
#undef MVG_QSCAN
    if (wqp) {  // segmented merge of each query's items
      if (kf == 1 && !per_k)
        launch_merge(wq.part, a.nq, kk, a.k, ip ? 1 : 0, kIvfPadDist, a.D, a.I, nullptr, nullptr, s,
                     wq.item_off);
      else
        launch_merge(wq.part, a.nq, kk, kk, 0, kIvfPadDist, nullptr, nullptr, nullptr, keys, s,
                     wq.item_off);
    }
    if (per_k) {  // each query's own kq * k_factor best, sorted
      w.keys2.reserve(sizeof(uint64_t) * (size_t)a.nq * kk);
      if (kf == 1) {
        launch_merge(keys, a.nq, kk, kk, ip ? 1 : 0, kIvfPadDist, a.D, a.I, nullptr, nullptr, s, nullptr,
                     a.kq, 1);
      } else {
        launch_merge(keys, a.nq, kk, kk, 0, kIvfPadDist, nullptr, nullptr, nullptr,
                     w.keys2.as<uint64_t>(), s, nullptr, a.kq, kf);
        keys = w.keys2.as<uint64_t>();
      }
    }
    if (pf.on) MVG_HIP(hipEventRecord(pf.ev[2], s));
    if (kf > 1) {
      const int32_t *ord = w.ord_ready ? w.ord.as<int32_t>() : nullptr;
      if (a.precision != 0)
        rerank_mixed(a, w, keys, kk, ord, s);
      else
        rerank(a.q, a.nq, a.d, a.metric, a.vectors, a.ntotal, keys, kk, a.k, kIvfPadDist, a.D, a.I, s,
               ord, a.rows, a.kq);
    }
  } else {
  // 2. per-(query, probe) LUT + ADC scan + list-local top-kk.
  w.part.reserve(sizeof(uint64_t) * (size_t)a.nq * maxp * kk);
  uint64_t *part = w.part.as<uint64_t>();
#define MVG_SCAN(MFV)                                                          \
  if (ip) launch_scan<MFV, true>(a, maxp, probes, kk, part, s);                \
  else launch_scan<MFV, false>(a, maxp, probes, kk, part, s);
  switch (a.M) {
    case 8: MVG_SCAN(8) break;
    case 12: MVG_SCAN(12) break;
    case 16: MVG_SCAN(16) break;
    default: MVG_SCAN(0) break;
  }
#undef MVG_SCAN
  if (pf.on) MVG_HIP(hipEventRecord(pf.ev[2], s));
  // 3. per-query merge (and rerank); with per-query k the merge keeps each
  // query's own kq (or kq * k_factor) best.
  if (kf == 1) {
    launch_merge(part, a.nq, (int64_t)maxp * kk, a.k, ip ? 1 : 0, kIvfPadDist, a.D, a.I, nullptr,
                 nullptr, s, nullptr, a.kq, 1);
  } else {
    w.keys.reserve(sizeof(uint64_t) * (size_t)a.nq * kk);
    launch_merge(part, a.nq, (int64_t)maxp * kk, kk, 0, kIvfPadDist, nullptr, nullptr, nullptr,
                 w.keys.as<uint64_t>(), s, nullptr, a.kq, kf);
    if (a.precision != 0)
      rerank_mixed(a, w, w.keys.as<uint64_t>(), kk, nullptr, s);
    else
      rerank(a.q, a.nq, a.d, a.metric, a.vectors, a.ntotal, w.keys.as<uint64_t>(), kk, a.k, kIvfPadDist,
             a.D, a.I, s, nullptr, a.rows, a.kq);
  }
  }
  if (pf.on) {
    MVG_HIP(hipEventRecord(pf.ev[3], s));
    const int64_t nt = a.nq * maxp;
    hipLaunchKernelGGL(scanned_entries_kernel, dim3(ceil_div(nt, 256)), dim3(256), 0, s, probes, a.nq,
                       maxp, a.nprobe_per_query, a.nprobe, a.list_off,
                       pf.counter.as<unsigned long long>());
    MVG_LAUNCH_CHECK("scanned_entries_kernel");
    pf.stream = s;
    pf.recorded = true;
  }
}

}  // namespace mvg
