// Mixed-form IVF-PQ L2 scan on the matrix cores (FP16 MFMA), for searches
// whose candidates are reranked exactly afterwards (k_factor > 1).
//
// The mixed L2 distance of a query q to entry e of list l is
//   |q - c_l|^2 + pre[e] - 2 <q, r_e>,      r_e = PQ reconstruction of e
// (index_ivf_pq.mojo:316-347 sums the same decomposition through per-list
// residual tables; pre[e] = sum_m T[l][m][code_m] is stored per entry). The
// query-dependent part is a dense contraction of queries with reconstructed
// entries, so a workgroup takes a tile of 16 queries that probe mostly the
// same lists, walks the union of their probed lists, decodes each list's PQ
// codes into FP16 MFMA fragments from an FP16 copy of the codebooks held in
// LDS, and multiplies them with the query tile (held in registers) on
// v_mfma_f32_16x16x32_f16. A query that does not probe a list gets an
// infinite base term there, so only its own lists can yield candidates.
//
// Selection (the k * k_factor smallest (distance, id) pairs of each query,
// exactly, for these distances): two passes over the tile's lists.
//   1. a sample of the keys (every eighth block of 16 entries) is counted in
//      a 512-bin histogram per query (1/64 octave per bin, over a window of
//      eight octaves placed at the query's nearest-centroid distance) under a
//      running bound aimed at an eighth of 1.25 kk keys; the bound only
//      tightens, so the counts are exact at and below it;
//   2. every key below the bin edge pass 1 found is collected into a
//      per-query buffer and counted in an exact histogram; if at least kk
//      were collected (or pass 1 set no bound), the kk smallest of them are
//      the kk smallest overall: keys below the kk-th key's bin are written
//      out and a wavefront radix select on (key, id) takes the rest of that
//      bin.
// A query whose collected keys fall short of kk, or overflow the buffers, is
// listed for the table-based deep-pool scan, which then writes its
// candidates (the host always enqueues it; it exits at once when the list is
// empty).
#include <cmath>

#include "dist.h"
#include "kernels.h"
#include "select.h"

namespace mvg {

namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int kTQ = 16;             // queries per tile
constexpr int kTBins = 512;         // per-query key bins (32-bit counts)
constexpr int kTShift = 17;         // key bits below a bin: 1/64 octave
constexpr int kTBelow = 192;        // window start: 3 octaves below the nearest centroid
constexpr int kTItem = 512;         // list entries per work item
constexpr int kTItem1 = 4096;       // list entries per pass-1 work item (where its prefix fits in LDS)
constexpr int kTGroup = 4;          // blocks of 16 entries loaded together, one group ahead
constexpr int kTStage = 128;        // per-wave staging of pass-2 keys
constexpr int kTRanges = 8;         // tile ranges, one per XCD
constexpr int kTPerQuery = 8;       // per-query words in LDS
constexpr int kTSample = 8;         // pass 1 visits every kTSample-th block of 16 entries
constexpr int kTNT = 512;           // threads per workgroup: two waves per SIMD
constexpr int kTNW = kTNT / 64;
constexpr int kTSelW = 4;           // waves of the second selection round (the codebook's LDS holds their buffers)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float t_wave_maxf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ int lane_rank(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// 2^e as a float (e clamped to the normal range).
__device__ __forceinline__ float pow2f(int e) {
  e = min(max(e, -126), 127);
  return __uint_as_float((uint32_t)(e + 127) << 23);
}

// Bin of a key under a window starting at `anc` (keys below and above the
// window fall into the first and last bins).
__device__ __forceinline__ int key_bin(uint32_t key, int anc) {
  return min(max((int)(key >> kTShift) - anc, 0), kTBins - 1);
}

// Smallest distance whose key is at or above the lower edge of window bin b
// (b in [1, kTBins]); +inf past the window or the finite range.
__device__ __forceinline__ float bin_edge(int anc, int b) {
  const int64_t top = (int64_t)anc + b;
  if (b >= kTBins || top >= (int64_t)1 << (32 - kTShift)) return __builtin_inff();
  const uint32_t edge = (uint32_t)top << kTShift;
  if (edge >= 0xFF800000u) return __builtin_inff();  // key of +inf
  return key2f(edge);
}

struct TileLayout {
  int cb, hist, run, ua, ub, ua1, pq, wb, stg, misc, total;
  __host__ __device__ TileLayout(int MF, int rc, bool items1) {
    cb = 0;
    hist = cb + MF * kKsub * 16;        // codebook rows of 8 halves
    run = hist + kTQ * kTBins * 4;      // (query rows during set-up)
    ua = run + rc * 8;
    ub = ua + rc * 8;
    ua1 = ub + rc * 8;                  // pass-1 item prefix (items1)
    pq = ua1 + (items1 ? rc * 4 : 0);
    wb = pq + kTPerQuery * kTQ * 4;
    stg = wb + kTNW * kTQ * 4;
    misc = stg + kTNW * kTStage * 12;
    total = misc + 64 * 4;
  }
};

struct TileScanArgs {
  const float *q;
  int64_t nq;
  const float *coarse;
  const uint16_t *cb16;  // [M][256][8] FP16 of -2 * pq * 2^cb_exp
  float cb_inv;          // 2^-cb_exp
  const float *pre;
  const int64_t *list_off;
  const uint8_t *codes;
  const int32_t *ids;
  int64_t ncsr;
  const int32_t *probes;
  int maxp;
  const int32_t *nprobe_q;
  int nprobe_all, kk;
  int kk1;  // sampled keys pass 1 aims at
  int items1;  // pass 1 runs items of kTItem1 entries (their prefix in LDS)
  uint64_t *out_keys;
  uint64_t *cbuf;  // per workgroup: kTQ collect buffers of cmax words
  int cmax, bufcap, rc;
  const int32_t *order;
  int64_t ntiles;
  unsigned *ranges;
  int32_t *ovf_list;
  unsigned *ovf_count;
  unsigned long long *pool_writes;
};

// Next tile of a persistent workgroup: its XCD's range first, then the others.
__device__ int64_t next_tile(int64_t ntiles, unsigned *ranges) {
  for (int a = 0; a < kTRanges; a++) {
    const int r = (int)((blockIdx.x + a) % kTRanges);
    const int64_t beg = ntiles * r / kTRanges, end = ntiles * (r + 1) / kTRanges;
    if (beg >= end) continue;
    const int64_t t = beg + (int64_t)atomicAdd(&ranges[r], 1u);
    if (t < end) return t;
  }
  return -1;
}

// First bin of a 512-bin histogram at which the running count reaches
// `need`; -1 if the total is below it. before: the count ahead of that bin;
// count: the bin's own; total (nullable): every bin's.
__device__ int wave_find_bin(const uint32_t *h, int need, int &before, int &count, int *total = nullptr) {
  const int lane = threadIdx.x & 63;
  const uint4 w0 = reinterpret_cast<const uint4 *>(h)[2 * lane];  // bins 8 * lane .. + 8
  const uint4 w1 = reinterpret_cast<const uint4 *>(h)[2 * lane + 1];
  const int hs[8] = {(int)w0.x, (int)w0.y, (int)w0.z, (int)w0.w, (int)w1.x, (int)w1.y, (int)w1.z, (int)w1.w};
  int local = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) local += hs[i];
  // inclusive scan over the wave: DPP row shifts by 1, 2, 4, 8 within each
  // 16-lane row (a lane with no source in its row adds 0), then the totals
  // of the rows below (lanes 15, 31, 47 read as scalars)
  int incl = local;
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xF, 0xF, false);
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xF, 0xF, false);
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xF, 0xF, false);
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xF, 0xF, false);
  const int r0 = __builtin_amdgcn_readlane(incl, 15), r1 = __builtin_amdgcn_readlane(incl, 31),
            r2 = __builtin_amdgcn_readlane(incl, 47), r3 = __builtin_amdgcn_readlane(incl, 63);
  incl += (lane >= 16 ? r0 : 0) + (lane >= 32 ? r1 : 0) + (lane >= 48 ? r2 : 0);
  if (total) *total = r0 + r1 + r2 + r3;
  const uint64_t hit = __ballot(incl >= need);
  if (hit == 0) return -1;
  const int L = __ffsll((unsigned long long)hit) - 1;
  int c = __builtin_amdgcn_readlane(incl - local, L);
  int bin = -1, cnt = 0, bef = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    if (bin < 0 && c + hs[i] >= need) {
      bin = 8 * L + i;
      cnt = hs[i];
      bef = c;
    }
    c += hs[i];
  }
  // lane L's findings
  before = __builtin_amdgcn_readlane(bef, L);
  count = __builtin_amdgcn_readlane(cnt, L);
  return __builtin_amdgcn_readlane(bin, L);
}

// The `need` smallest of the unique words buf[0..n) (every one of them shares
// the bits above `shift + 8`) appended to out[] from slot `at` (slots at or
// past `limit` are never written), by one wavefront; h: 256 ints of scratch.
__device__ void wave_select_words(const uint64_t *buf, int n, int need, int shift, int *h,
                                  uint64_t *out, int at, int limit) {
  const int lane = threadIdx.x & 63;
  uint64_t prefix = 0, pmask = 0;
  for (;;) {
#pragma unroll
    for (int i = 0; i < 4; i++) h[lane + 64 * i] = 0;
    wave_sync();
    for (int j = lane; j < n; j += 64) {
      const uint64_t v = buf[j];
      if ((v & pmask) == prefix) atomicAdd(&h[(int)((v >> shift) & 0xFF)], 1);
    }
    wave_sync();
    const int4 c4 = reinterpret_cast<const int4 *>(h)[lane];  // bins 4 * lane .. + 4
    const int cs[4] = {c4.x, c4.y, c4.z, c4.w};
    const int local = cs[0] + cs[1] + cs[2] + cs[3];
    int incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    const int L = __ffsll((unsigned long long)__ballot(incl >= need)) - 1;
    int c = __shfl(incl - local, L, 64), bin = 0, bef = 0, cnt = 0;
    if (lane == L) {
      bool found = false;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (!found && c + cs[i] >= need) {
          found = true;
          bin = 4 * L + i;
          bef = c;
          cnt = cs[i];
        }
        c += cs[i];
      }
    }
    bin = __shfl(bin, L, 64);
    bef = __shfl(bef, L, 64);
    cnt = __shfl(cnt, L, 64);
    wave_sync();
    prefix |= (uint64_t)bin << shift;
    pmask |= (uint64_t)0xFF << shift;
    need -= bef;
    if (cnt == need || shift == 0) break;  // the whole bin (unique words: one at shift 0)
    shift = shift >= 8 ? shift - 8 : 0;
  }
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    const uint64_t v = j < n ? buf[j] : ~0ull;
    const bool take = j < n && (v & pmask) <= prefix;
    const uint64_t m = __ballot(take);
    if (take && at + lane_rank(m) < limit) out[at + lane_rank(m)] = v;
    at += (int)__popcll(m);
  }
}

// Exclusive scan of one int per thread over the workgroup (kTNT threads);
// sh needs kTNW + 1 ints.
__device__ __forceinline__ int tile_excl_scan(int v, int *sh, int &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kTNW; w++) {
    const int t = sh[w];
    if (w < wave) off += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

// Ascending bitonic sort of keys[0..n) in LDS by kTNT threads, n a power of two.
__device__ void tile_sort_u64(uint64_t *keys, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += kTNT) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t x = keys[i], y = keys[ixj];
          if ((x > y) == ((i & k) == 0)) {
            keys[i] = y;
            keys[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Per-entry arrays as buffer resources (32-bit byte offsets; reads past the
// end return zero). The host checks that they stay below 2 GiB.
struct TileSrc {
  __amdgpu_buffer_rsrc_t codes, pre, ids;
};

// One group of kTGroup blocks of 16 list entries as one lane loads them:
// its entry's PQ code, entry term and (pass 2) id in each block; blocks
// `stride` blocks apart (pass 1 samples every kTSample-th block).
template <int MF>
struct TileGrp {
  uint32_t cw[kTGroup][MF / 4];
  float pre[kTGroup];
  int32_t id[kTGroup];
};

template <int MF, bool IDS, int STRIDE>
__device__ __forceinline__ void tile_load(const TileSrc &src, uint32_t e0, TileGrp<MF> &g) {
#pragma unroll
  for (int b = 0; b < kTGroup; b++) {
    const uint32_t e = e0 + 16u * STRIDE * b;
    if constexpr (MF == 16) {
      const uint4 v = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(src.codes, (int)(e * 16u), 0, 0));
      g.cw[b][0] = v.x;
      g.cw[b][1] = v.y;
      g.cw[b][2] = v.z;
      g.cw[b][3] = v.w;
    } else if constexpr (MF == 8) {
      const uint2 v = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(src.codes, (int)(e * 8u), 0, 0));
      g.cw[b][0] = v.x;
      g.cw[b][1] = v.y;
    } else {
#pragma unroll
      for (int i = 0; i < MF / 4; i++)
        g.cw[b][i] = __builtin_amdgcn_raw_buffer_load_b32(src.codes, (int)(e * (uint32_t)MF), 4 * i, 0);
    }
    g.pre[b] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(src.pre, (int)(e * 4u), 0, 0));
    g.id[b] = IDS ? (int32_t)__builtin_amdgcn_raw_buffer_load_b32(src.ids, (int)(e * 4u), 0, 0) : 0;
  }
}

// Per-item registers of a lane: the list's base term and the bounds of the
// queries whose distances the lane receives (4 * lgrp + j).
struct TileItem {
  float base[4], E[4];
};

// Pass-2 staging of one wave: (key << 32 | id) words and their query slots.
struct TileStage {
  uint64_t *w;
  uint32_t *s;
};

// Flush of a wave's staged pass-2 keys, all lanes at once: each key takes a
// slot of its query's collect buffer and is counted in its histogram bin.
__device__ __forceinline__ void stage_flush(const TileScanArgs &a, TileStage st, int n, int *cnt_s, const int *anc_s,
                                            uint32_t *hist, uint64_t *cbw) {
  const int lane = threadIdx.x & 63;
  wave_sync();
  for (int i = lane; i < n; i += 64) {
    const uint64_t w = st.w[i];
    const int s = (int)st.s[i];
    const int slot = atomicAdd(&cnt_s[s], 1);
    if (slot < a.cmax) cbw[(int64_t)s * a.cmax + slot] = w;
    atomicAdd(&hist[s * kTBins + key_bin((uint32_t)(w >> 32), anc_s[s])], 1u);
  }
  wave_sync();
}

// One group of blocks: decode, MFMA, and the pass's epilogue, two blocks at
// a time (their eight codebook reads issued before either MFMA chain, the
// two chains interleaved).
// MFMA layout: A = the tile's query fragments (lane: query lrow, the 8
// dimensions of subquantizer 4 * ks + lgrp), B = the decoded entries (lane:
// entry lrow of the block, same dimensions); the lane then holds the
// distances of its entry to queries 4 * lgrp + j (j < 4).
template <int MF, int PASS, int STRIDE>
__device__ __forceinline__ void tile_group(const TileScanArgs &a, const TileGrp<MF> &g, uint32_t e0, uint32_t ie,
                                           int nleft, const half8_t (&qf)[MF / 4], const float (&inv)[4],
                                           const int (&anc)[4], const TileItem &ti, const char *cbb,
                                           uint32_t cb_lane, uint32_t *hist, TileStage st, int &nst, int *cnt_s,
                                           const int *anc_s, uint64_t *cbw) {
  constexpr int KS = MF / 4;
  const int lane = threadIdx.x & 63;
  const int lgrp = lane >> 4;
#pragma unroll
  for (int b = 0; b < kTGroup; b += 2) {
    if (b >= nleft) break;
    half8_t ev[2][KS];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const uint32_t c = __builtin_amdgcn_ubfe(g.cw[b + h][ks], 8 * lgrp, 8);
        ev[h][ks] = *reinterpret_cast<const half8_t *>(cbb + cb_lane + (uint32_t)ks * (4 * kKsub * 16) + (c << 4));
      }
    floatx4_t acc[2] = {floatx4_t{0.0f, 0.0f, 0.0f, 0.0f}, floatx4_t{0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
      for (int h = 0; h < 2; h++) acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_f16(qf[ks], ev[h][ks], acc[h], 0, 0, 0);
    float dist[2][4];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t e = e0 + 16u * STRIDE * (b + h);  // this lane's entry (past the item's end: +inf)
      const float pv = (e < ie && b + h < nleft) ? g.pre[b + h] : __builtin_inff();
#pragma unroll
      for (int j = 0; j < 4; j++) dist[h][j] = __builtin_fmaf(acc[h][j], inv[j], pv) + ti.base[j];
    }
    if constexpr (PASS == 1) {
      // count sampled keys under the running bound (no return: nothing waits)
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (dist[h][j] < ti.E[j]) {
            const int bn = key_bin(f2key(dist[h][j]), anc[j]);
            atomicAdd(&hist[(4 * lgrp + j) * kTBins + bn], 1u);
          }
        }
    } else {
      // staged: positions from the ballots; slots and counts at the flush
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const bool in = dist[h][j] < ti.E[j];
          const uint64_t m = __ballot(in);
          if (m == 0) continue;
          if (nst + 64 > kTStage) {  // a combination stages at most 64 keys
            stage_flush(a, st, nst, cnt_s, anc_s, hist, cbw);
            nst = 0;
          }
          if (in) {
            const int pos = nst + lane_rank(m);
            st.w[pos] = ((uint64_t)f2key(dist[h][j]) << 32) | (uint32_t)g.id[b + h];
            st.s[pos] = (uint32_t)(4 * lgrp + j);
          }
          nst += (int)__popcll(m);
        }
    }
  }
}

// One sweep of a workgroup's waves over the tile's work items (list chunks
// of kTItem entries). PASS 1 visits every kTSample-th block of 16 entries and
// counts the keys below the running bound (aimed at kk1 sampled keys); PASS 2
// visits every block and collects the keys below the bound pass 1 left.
template <int MF, int PASS>
__device__ __forceinline__ void tile_sweep(const TileScanArgs &a, const TileSrc &src, const half8_t (&qf)[MF / 4],
                                           const float (&inv)[4], const int (&anc)[4], const _Float16 *cb,
                                           uint32_t *hist, const uint64_t *run, const uint64_t *ua,
                                           const uint64_t *ub, const uint32_t *ua1, int U, int NI, float *E_s,
                                           int *cnt_s,
                                           const int *anc_s, float *wb, TileStage st, int *itc, uint64_t *cbw) {
  constexpr bool IDS = PASS == 2;
  constexpr int STRIDE = PASS == 1 ? kTSample : 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lrow = lane & 15, lgrp = lane >> 4;
  float *wbw = wb + wave * kTQ;
  // LDS byte offset of this lane's codebook rows: subquantizer 4 * ks + lgrp
  const uint32_t cb_lane = (uint32_t)lgrp * kKsub * 16;
  const char *cbb = reinterpret_cast<const char *>(cb);
  int nst = 0;  // staged pass-2 keys of this wave
  int u = 0;    // union cursor: items are claimed in increasing order
  int nupd = 0;  // pass-1 items of this wave
  const uint32_t item = ua1 ? (uint32_t)kTItem1 : (uint32_t)kTItem;
  // An item: its first entry of this lane, its end, the blocks visited, its
  // list's base-term run; past the last item, entries from 0 with no block.
  struct Item {
    int it, nblk, rs, rl;
    uint32_t ib, ie;
  };
  auto claim = [&]() __attribute__((always_inline)) {
    int it = 0;
    if (lane == 0) it = atomicAdd(itc, 1);
    it = __shfl(it, 0, 64);
    Item m{it, 0, 0, 0, (uint32_t)lrow, 0u};
    if (it < NI) {
      // item starts: the pass-1 prefix ua1 (items of kTItem1), else ua's
      while (u + 1 < U && (ua1 ? (int)ua1[u + 1] : (int)(ua[u + 1] >> 32)) <= it) u++;
      const uint64_t ea = ua[u], eb = ub[u];
      const uint32_t chunk = (uint32_t)(it - (ua1 ? (int)ua1[u] : (int)(ea >> 32)));
      m.rs = (int)((ea >> 8) & 0xFFFFFFu);
      m.rl = (int)(ea & 0xFFu);
      const uint32_t lbeg = (uint32_t)(eb >> 32), llen = (uint32_t)eb;
      m.ib = lbeg + chunk * item + lrow;
      m.ie = lbeg + min(llen, (chunk + 1) * item);
      m.nblk = (int)((m.ie - (m.ib - lrow) + 16 * STRIDE - 1) / (16 * STRIDE));  // blocks visited
    }
    return m;
  };
  // two load buffers in turn: a group's loads are issued one group ahead,
  // unconditionally (a conditional load leaves the compiler unable to wait
  // for the older group alone; past the item the entries are unused and
  // past the arrays the buffer reads return zero). An item runs an even
  // number of groups (a last empty one costs its loads only), so the next
  // item's first group is always loaded into the first buffer, during the
  // current item's last group.
  constexpr uint32_t kStep = 16u * STRIDE * kTGroup;
  Item cur = claim();
  TileGrp<MF> ga, gb;
  tile_load<MF, IDS, STRIDE>(src, cur.ib, ga);
  while (cur.it < NI) {
    const Item nxt = claim();
    // this list's base term per tile query (+inf where the query does not probe it)
    if (lane < kTQ) wbw[lane] = __builtin_inff();
    wave_sync();
    if (lane < cur.rl) {
      const uint64_t v = run[cur.rs + lane];
      wbw[(int)(v >> 32)] = __uint_as_float((uint32_t)v);
    }
    wave_sync();
    TileItem ti;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int s = 4 * lgrp + j;
      ti.base[j] = wbw[s];
      ti.E[j] = E_s[s];
    }
    const int ngrp = (((cur.nblk + kTGroup - 1) / kTGroup) + 1) & ~1;
    for (int gi = 0; gi < ngrp; gi += 2) {
      tile_load<MF, IDS, STRIDE>(src, cur.ib + kStep * (gi + 1), gb);
      tile_group<MF, PASS, STRIDE>(a, ga, cur.ib + kStep * gi, cur.ie, cur.nblk - gi * kTGroup, qf, inv, anc, ti,
                                   cbb, cb_lane, hist, st, nst, cnt_s, anc_s, cbw);
      tile_load<MF, IDS, STRIDE>(src, gi + 2 < ngrp ? cur.ib + kStep * (gi + 2) : nxt.ib, ga);
      tile_group<MF, PASS, STRIDE>(a, gb, cur.ib + kStep * (gi + 1), cur.ie, cur.nblk - (gi + 1) * kTGroup, qf, inv,
                                   anc, ti, cbb, cb_lane, hist, st, nst, cnt_s, anc_s, cbw);
    }
    if constexpr (PASS == 1) {
      // tighten the bounds of this wave's queries after every large item or
      // every fourth small one (counts only grow, so a bound read while
      // others add is still an upper bound; a looser one only lets more
      // sampled keys be counted)
      if (ua1 || (++nupd & 3) == 0)
        for (int s = wave; s < kTQ; s += kTNW) {
          int bef, cnt;
          const int bn = wave_find_bin(hist + s * kTBins, a.kk1, bef, cnt);
          if (lane == 0 && bn >= 0) E_s[s] = bin_edge(anc_s[s], bn + 1);
        }
    }
    cur = nxt;
  }
  if constexpr (PASS == 2) {
    if (nst > 0) stage_flush(a, st, nst, cnt_s, anc_s, hist, cbw);
  }
}

template <int MF>
__global__ __launch_bounds__(kTNT) void ivfpq_tile_scan_kernel(TileScanArgs a) {
  constexpr int KS = MF / 4;
  constexpr int D = 8 * MF;
  constexpr int T = kTQ;
  extern __shared__ __align__(16) unsigned char smem[];
  const TileLayout lay(MF, a.rc, a.items1 != 0);
  _Float16 *cb = reinterpret_cast<_Float16 *>(smem + lay.cb);
  uint32_t *hist = reinterpret_cast<uint32_t *>(smem + lay.hist);
  float *qv = reinterpret_cast<float *>(smem + lay.hist);
  uint64_t *run = reinterpret_cast<uint64_t *>(smem + lay.run);
  uint64_t *ua = reinterpret_cast<uint64_t *>(smem + lay.ua);
  uint64_t *ub = reinterpret_cast<uint64_t *>(smem + lay.ub);
  uint32_t *ua1 = a.items1 ? reinterpret_cast<uint32_t *>(smem + lay.ua1) : nullptr;
  int *qid_s = reinterpret_cast<int *>(smem + lay.pq);
  float *E_s = reinterpret_cast<float *>(qid_s + T);    // pass-1 running bound, then the collect bound
  int *cnt_s = reinterpret_cast<int *>(E_s + T);        // keys collected
  int *anc_s = cnt_s + T;
  int *np_s = anc_s + T;
  unsigned *bmin_s = reinterpret_cast<unsigned *>(np_s + T);
  float *inv_s = reinterpret_cast<float *>(bmin_s + T);
  float *sq_s = inv_s + T;
  float *wb = reinterpret_cast<float *>(smem + lay.wb);
  int *misc = reinterpret_cast<int *>(smem + lay.misc);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lgrp = lane >> 4;
  uint64_t *cbw = a.cbuf + (int64_t)blockIdx.x * T * a.cmax;
  TileStage st;
  st.w = reinterpret_cast<uint64_t *>(smem + lay.stg) + wave * kTStage;
  st.s = reinterpret_cast<uint32_t *>(smem + lay.stg + kTNW * kTStage * 8) + wave * kTStage;
  TileSrc src;
  src.codes = __builtin_amdgcn_make_buffer_rsrc((void *)a.codes, (short)0, (int)(a.ncsr * MF), 0x00020000);
  src.pre = __builtin_amdgcn_make_buffer_rsrc((void *)a.pre, (short)0, (int)(a.ncsr * 4), 0x00020000);
  src.ids = __builtin_amdgcn_make_buffer_rsrc((void *)a.ids, (short)0, (int)(a.ncsr * 4), 0x00020000);

  for (;;) {
    __syncthreads();
    if (tid == 0) misc[0] = (int)next_tile(a.ntiles, a.ranges);
    __syncthreads();
    const int tile = misc[0];
    if (tile < 0) break;
    // ---- set-up: codebook, tile queries, their probes, the union of lists
    for (int i = tid; i < MF * kKsub; i += kTNT)
      reinterpret_cast<uint4 *>(cb)[i] = reinterpret_cast<const uint4 *>(a.cb16)[i];
    if (tid < T) {
      const int64_t t = (int64_t)tile * T + tid;
      const int q = t < a.nq ? a.order[t] : -1;
      qid_s[tid] = q;
      int np = 0;
      if (q >= 0) np = min(max(a.nprobe_q ? a.nprobe_q[q] : a.nprobe_all, 1), a.maxp);
      np_s[tid] = np;
      bmin_s[tid] = 0x7F800000u;  // +inf
      cnt_s[tid] = 0;
      E_s[tid] = __builtin_inff();
    }
    if (tid == 0) misc[4] = 0;
    __syncthreads();
    for (int i = tid; i < T * D; i += kTNT) {
      const int s = i / D;
      const int q = qid_s[s];
      qv[i] = q >= 0 ? a.q[(int64_t)q * D + (i - s * D)] : 0.0f;
    }
    const int nel = T * a.maxp;
    for (int i = tid; i < a.rc; i += kTNT) {
      uint64_t key = kSentinel64;
      if (i < nel) {
        const int s = i / a.maxp, p = i - s * a.maxp;
        const int q = qid_s[s];
        if (q >= 0 && p < np_s[s]) {
          const int l = a.probes[(int64_t)q * a.maxp + p];
          if (l >= 0) key = ((uint64_t)l << 13) | ((uint64_t)p << 6) | (uint64_t)s;
        }
      }
      run[i] = key;
    }
    __syncthreads();
    tile_sort_u64(run, a.rc);
    // union entries (one per run of equal lists), keyed by the smallest probe
    // rank among the tile's queries so that near lists come first
    const int per = a.rc / kTNT;
    const int r0 = tid * per;
    int nstart = 0;
    for (int i = r0; i < r0 + per; i++) {
      const uint64_t k = run[i];
      nstart += k != kSentinel64 && (i == 0 || (run[i - 1] >> 13) != (k >> 13));
    }
    int U;
    int at = tile_excl_scan(nstart, misc + 8, U);
    for (int i = r0; i < r0 + per; i++) {
      const uint64_t k = run[i];
      if (k != kSentinel64 && (i == 0 || (run[i - 1] >> 13) != (k >> 13))) {
        int len = 1;
        while (i + len < a.rc && run[i + len] != kSentinel64 && (run[i + len] >> 13) == (k >> 13)) len++;
        ua[at++] = (((k >> 6) & 127u) << 32) | ((uint64_t)i << 8) | (uint64_t)len;
      }
    }
    const int np2 = next_pow2(max(U, 1));
    for (int i = U + tid; i < np2; i += kTNT) ua[i] = kSentinel64;
    __syncthreads();
    tile_sort_u64(ua, np2);
    // list bounds and work items (kTItem entries each), exclusive item prefix
    const int perU = (U + kTNT - 1) / kTNT;
    const int u0 = min(tid * perU, U), u1 = min(u0 + perU, U);
    int nit = 0;
    for (int u = u0; u < u1; u++) {
      const int rs = (int)((ua[u] >> 8) & 0xFFFFFFu);
      const int64_t l = (int64_t)(run[rs] >> 13);
      const int64_t lb = a.list_off[l], le = a.list_off[l + 1];
      const uint32_t len = (uint32_t)(le - lb);
      ub[u] = ((uint64_t)(uint32_t)lb << 32) | len;
      nit += (int)((len + kTItem - 1) / kTItem);
    }
    int NI;
    int ip = tile_excl_scan(nit, misc + 8, NI);
    for (int u = u0; u < u1; u++) {
      const uint64_t e = ua[u];
      ua[u] = ((uint64_t)(uint32_t)ip << 32) | (e & 0xFFFFFFFFull);
      ip += (int)(((uint32_t)ub[u] + kTItem - 1) / kTItem);
    }
    // pass-1 items of kTItem1 entries (the sampled pass's per-item costs
    // over more blocks), or the pass-2 items where their prefix does not fit
    int NI1 = NI;
    if (ua1) {
      int nit1 = 0;
      for (int u = u0; u < u1; u++) nit1 += (int)(((uint32_t)ub[u] + kTItem1 - 1) / kTItem1);
      int ip1 = tile_excl_scan(nit1, misc + 8, NI1);
      for (int u = u0; u < u1; u++) {
        ua1[u] = (uint32_t)ip1;
        ip1 += (int)(((uint32_t)ub[u] + kTItem1 - 1) / kTItem1);
      }
    }
    __syncthreads();  // every list id read from run[] before it is rewritten
    // base term |q - c_l|^2 of every (query, probed list), in place of the key
    for (int i = tid; i < a.rc; i += kTNT) {
      const uint64_t k = run[i];
      if (k == kSentinel64) continue;
      const int s = (int)(k & 63u);
      const int64_t l = (int64_t)(k >> 13);
      const float b = l2_w8(qv + s * D, a.coarse + l * D, D);
      atomicMin(&bmin_s[s], __float_as_uint(b));
      run[i] = ((uint64_t)s << 32) | __float_as_uint(b);
    }
    // per-query scale 2^e so that the largest |q| lies in [2^14, 2^15)
    for (int s = wave; s < T; s += kTNW) {
      float m = 0.0f;
      for (int j = lane; j < D; j += 64) m = fmaxf(m, fabsf(qv[s * D + j]));
      m = t_wave_maxf(m);
      if (lane == 0) {
        const int e = m > 0.0f ? (int)((__float_as_uint(m) >> 23) & 0xFF) - 127 : 0;
        sq_s[s] = pow2f(14 - e);
        inv_s[s] = a.cb_inv * pow2f(e - 14);
      }
    }
    __syncthreads();
    if (tid < T) anc_s[tid] = max(0, (int)(f2key(__uint_as_float(bmin_s[tid])) >> kTShift) - kTBelow);
    __syncthreads();
    // query fragments: lane holds query lrow, dimensions of subquantizer
    // 4 * ks + lgrp; the scales and windows of the queries whose distances
    // the lane receives (4 * lgrp + j)
    half8_t qf[KS];
    float inv[4];
    int anc[4];
    {
      const float sc = sq_s[lrow];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        inv[j] = inv_s[4 * lgrp + j];
        anc[j] = anc_s[4 * lgrp + j];
      }
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const float *q8 = qv + lrow * D + 8 * (4 * ks + lgrp);
#pragma unroll
        for (int j = 0; j < 8; j++) qf[ks][j] = (_Float16)(q8[j] * sc);
      }
    }
    __syncthreads();  // query rows consumed
    for (int i = tid; i < T * kTBins; i += kTNT) hist[i] = 0u;
    __syncthreads();

    // ---- pass 1: histograms of a sample of the keys under running bounds
    tile_sweep<MF, 1>(a, src, qf, inv, anc, cb, hist, run, ua, ub, ua1, U, NI1, E_s, cnt_s, anc_s, wb, st, misc + 4,
                      cbw);
    __syncthreads();
    // collect bound of each query: the top edge of the bin where its sampled
    // keys reach kk1 (+inf when fewer were sampled: every key)
    for (int s = wave; s < T; s += kTNW) {
      int bef, cnt;
      const int b = wave_find_bin(hist + s * kTBins, a.kk1, bef, cnt);
      if (lane == 0) E_s[s] = qid_s[s] < 0 ? -__builtin_inff() : (b >= 0 ? bin_edge(anc_s[s], b + 1) : __builtin_inff());
    }
    __syncthreads();
    for (int i = tid; i < T * kTBins; i += kTNT) hist[i] = 0u;
    if (tid == 0) misc[4] = 0;
    __syncthreads();

    // ---- pass 2: every key below the bound into the collect buffers, counted
    tile_sweep<MF, 2>(a, src, qf, inv, anc, cb, hist, run, ua, ub, nullptr, U, NI, E_s, cnt_s, anc_s, wb, st,
                      misc + 4, cbw);
    __syncthreads();

    // ---- per query (one wavefront each): the kk smallest of the collected
    // keys by their exact histogram; the codebook's LDS holds the boundary
    // bins. Round 1: every wavefront, bufcap / 2 keys of boundary buffer
    // each; a query whose boundary bin holds more waits for round 2 (kTSelW
    // wavefronts, bufcap keys each; np_s, free after the set-up, lists them).
    if (tid == 0) misc[5] = 0;
    __syncthreads();
    auto select_query = [&](int s, uint64_t *wbuf, int cap, bool defer) __attribute__((always_inline)) {
      const int q = qid_s[s];
      if (q < 0) return;
      const int n = cnt_s[s];
      int bef = 0, cnt = 0;
      const int b = wave_find_bin(hist + s * kTBins, a.kk, bef, cnt);
      if (defer && b >= 0 && cnt > cap && n <= a.cmax) {
        if (lane == 0) np_s[atomicAdd(&misc[5], 1)] = s;
        return;
      }
      // fewer keys below a finite bound than kk (the sample's margin missed),
      // more than the buffers hold: the deep-pool scan takes the query
      if (n > a.cmax || (b < 0 && E_s[s] != __builtin_inff()) || (b >= 0 && cnt > cap)) {
        if (lane == 0) a.ovf_list[atomicAdd(a.ovf_count, 1u)] = q;
        return;
      }
      uint64_t *out = a.out_keys + (int64_t)q * a.kk;
      const uint64_t *cq = cbw + (int64_t)s * a.cmax;
      int nout = 0, nb = 0;
      const int anc_q = anc_s[s];
      for (int j0 = 0; j0 < n; j0 += 4 * 64) {
        uint64_t v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int j = j0 + r * 64 + lane;
          v[r] = j < n ? cq[j] : ~0ull;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int j = j0 + r * 64 + lane;
          const int bn = key_bin((uint32_t)(v[r] >> 32), anc_q);
          const bool lo = j < n && (b < 0 || bn < b), eq = j < n && b >= 0 && bn == b;
          const uint64_t ml = __ballot(lo), me = __ballot(eq);
          if (lo && nout + lane_rank(ml) < a.kk) out[nout + lane_rank(ml)] = v[r];
          if (eq && nb + lane_rank(me) < cap) wbuf[nb + lane_rank(me)] = v[r];
          nout += (int)__popcll(ml);
          nb += (int)__popcll(me);
        }
      }
      wave_sync();
      int taken = 0;
      if (b >= 0) {
        taken = min(a.kk - nout, nb);
        if (taken > 0) {
          // the buffer's keys share their bin; interior bins fix the key bits
          // above kTShift, the two edge bins do not
          const int shift = (b == 0 || b == kTBins - 1) ? 56 : 32 + kTShift - 8;
          wave_select_words(wbuf, nb, taken, shift, reinterpret_cast<int *>(hist + s * kTBins), out, nout, a.kk);
        }
        taken = max(taken, 0);
      }
      for (int j = nout + taken + lane; j < a.kk; j += 64) out[j] = kSentinel64;
      if (a.pool_writes && lane == 0) atomicAdd(a.pool_writes, (unsigned long long)n);
    };
    {
      const int cap = a.bufcap / 2;
      uint64_t *wbuf = reinterpret_cast<uint64_t *>(smem + lay.cb) + wave * cap;
      for (int s = wave; s < T; s += kTNW) select_query(s, wbuf, cap, true);
    }
    __syncthreads();
    {
      const int nd = misc[5];
      uint64_t *wbuf = reinterpret_cast<uint64_t *>(smem + lay.cb) + wave * a.bufcap;
      for (int i = wave; i < nd && wave < kTSelW; i += kTSelW) select_query(np_s[i], wbuf, a.bufcap, false);
    }
  }
}

__global__ void tile_codebook_kernel(const float *__restrict__ pq, int64_t n, float scale,
                                     _Float16 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (_Float16)(-2.0f * pq[i] * scale);
}

// pre with the entries whose id the exclusion mask sets made infinite
// (index_ivf_pq.mojo:328-334).
__global__ void tile_mask_terms_kernel(const float *__restrict__ pre, const int32_t *__restrict__ ids, int64_t n,
                                       const uint8_t *__restrict__ excl, int64_t excl_len, float *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int32_t id = ids[e];
  out[e] = (id >= 0 && id < excl_len && excl[id] != 0) ? __builtin_inff() : pre[e];
}

__global__ void tile_zero_kernel(unsigned *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace

float tile_codebook(const float *pq, int M, int dsub, float max_abs, uint16_t *out, hipStream_t s) {
  // 2^e with the largest |2 pq| * 2^e in [2^14, 2^15): FP16 keeps its full
  // precision and cannot overflow
  int e = 0;
  const float m = 2.0f * max_abs;
  if (m > 0.0f) {
    int ex = 0;
    (void)std::frexp(m, &ex);  // m = f * 2^ex, f in [0.5, 1)
    e = 15 - ex;
  }
  const float scale = std::ldexp(1.0f, e);
  const int64_t n = (int64_t)M * kKsub * dsub;
  hipLaunchKernelGGL(tile_codebook_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, pq, n, scale,
                     reinterpret_cast<_Float16 *>(out));
  MVG_LAUNCH_CHECK("tile_codebook_kernel");
  return std::ldexp(1.0f, -e);
}

bool tile_scan_supported(const IvfSearchArgs &a, int maxp) {
  // per-entry arrays addressed by 32-bit byte offsets below 2 GiB
  return a.cb16 && a.list_rank && a.pre && (a.M == 8 || a.M == 12 || a.M == 16) && a.d == 8 * a.M &&
         maxp <= 128 && a.ncsr * a.M < (int64_t(1) << 31) && a.ncsr * 4 < (int64_t(1) << 31) &&
         a.metric == kMetricL2;
}

template <int MF>
static void launch_tile_t(const TileScanArgs &ta, size_t sm, int64_t grid, hipStream_t s) {
  const void *fn = (const void *)ivfpq_tile_scan_kernel<MF>;
  MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  hipLaunchKernelGGL((ivfpq_tile_scan_kernel<MF>), dim3((unsigned)grid), dim3(kTNT), sm, s, ta);
  MVG_LAUNCH_CHECK("ivfpq_tile_scan_kernel");
}

bool launch_tile_scan(const IvfSearchArgs &a, int maxp, const int32_t *probes, int kk,
                      uint64_t *out_keys, IvfWork &w, hipStream_t s) {
  if (!tile_scan_supported(a, maxp) || !out_keys) return false;
  // a batch of fewer tiles than CUs runs each tile's union on one CU: the
  // per-query deep-pool scan is faster there
  const int64_t min_batch = a.tile_min_batch > 0 ? a.tile_min_batch : (int64_t)kTQ * device_cus();
  if (a.nq < min_batch) return false;
  const int T = kTQ;
  const int rc = std::max(kTNT, next_pow2(T * maxp));
  const bool items1 = TileLayout(a.M, rc, true).total <= 160 * 1024;
  const TileLayout lay(a.M, rc, items1);
  const size_t sm = (size_t)lay.total;
  if (sm > 160 * 1024) return false;
  // query order: by the super-cluster rank of each query's nearest list;
  // tiles of T consecutive queries
  const int32_t *order = build_tile_order(a, maxp, probes, a.list_rank, w, s);
  const int64_t ntiles = (a.nq + T - 1) / T;
  const int bufcap = a.M * kKsub * 16 / kTSelW / 8;
  // pass 1 aims its sample at 1.25 kk / kTSample + 16 keys: the bound then
  // lies near the 1.25 kk-th key, five standard deviations of the sample
  // count above the kk-th at kk = 2880; a query whose collected keys still
  // fall short of kk goes to the deep-pool scan
  const int kk1 = (5 * kk + 4 * kTSample - 1) / (4 * kTSample) + 16;
  const int cmax = 2 * kk + 512;  // collect buffer per query
  int64_t grid = std::min<int64_t>((int64_t)device_cus(), ntiles);
  grid = std::max<int64_t>(kTRanges, (grid + kTRanges - 1) / kTRanges * kTRanges);
  if (w.tile_bnd.arena) {  // a fixed workspace: as many workgroups as a quarter of it holds
    const int64_t per_wg = (int64_t)sizeof(uint64_t) * T * cmax;
    grid = std::min<int64_t>(grid, (int64_t)(w.tile_bnd.arena->left() / 4) / per_wg / kTRanges * kTRanges);
    if (grid < kTRanges) return false;
  }
  w.tile_bnd.reserve(sizeof(uint64_t) * (size_t)(grid * T * cmax));
  w.tile_ctr.reserve(sizeof(unsigned) * (kTRanges + 1));
  w.tile_ovf.reserve(sizeof(int32_t) * (size_t)a.nq);
  hipLaunchKernelGGL(tile_zero_kernel, dim3(1), dim3(64), 0, s, w.tile_ctr.as<unsigned>(), kTRanges + 1);
  MVG_LAUNCH_CHECK("tile_zero_kernel");
  TileScanArgs ta{};
  ta.q = a.q;
  ta.nq = a.nq;
  ta.coarse = a.coarse;
  ta.cb16 = a.cb16;
  ta.cb_inv = a.cb16_inv;
  ta.pre = a.pre;
  ta.list_off = a.list_off;
  ta.codes = a.codes;
  ta.ids = a.ids;
  ta.ncsr = a.ncsr;
  ta.probes = probes;
  ta.maxp = maxp;
  ta.nprobe_q = a.nprobe_per_query;
  ta.nprobe_all = a.nprobe;
  ta.kk = kk;
  ta.out_keys = out_keys;
  ta.kk1 = kk1;
  ta.items1 = items1 ? 1 : 0;
  ta.cbuf = w.tile_bnd.as<uint64_t>();
  ta.cmax = cmax;
  ta.bufcap = bufcap;
  ta.rc = rc;
  ta.order = order;
  ta.ntiles = ntiles;
  ta.ranges = w.tile_ctr.as<unsigned>();
  ta.ovf_count = w.tile_ctr.as<unsigned>() + kTRanges;
  ta.ovf_list = w.tile_ovf.as<int32_t>();
  ta.pool_writes = w.prof.on ? w.prof.counter.as<unsigned long long>() + 1 : nullptr;
  if (a.excl) {  // excluded entries: an infinite entry term (a masked copy)
    w.tile_pre.reserve(sizeof(float) * (size_t)std::max<int64_t>(a.ncsr, 1));
    hipLaunchKernelGGL(tile_mask_terms_kernel, dim3(ceil_div(std::max<int64_t>(a.ncsr, 1), 256)), dim3(256), 0, s,
                       a.pre, a.ids, a.ncsr, a.excl, a.excl_len, w.tile_pre.as<float>());
    MVG_LAUNCH_CHECK("tile_mask_terms_kernel");
    ta.pre = w.tile_pre.as<float>();
  }
  switch (a.M) {
    case 8: launch_tile_t<8>(ta, sm, grid, s); break;
    case 12: launch_tile_t<12>(ta, sm, grid, s); break;
    default: launch_tile_t<16>(ta, sm, grid, s); break;
  }
  // queries the tiles could not finish: the FP32-table deep-pool scan
  launch_deep_subset(a, maxp, probes, kk, out_keys, ta.ovf_list, ta.ovf_count, w, s);
  return true;
}

}  // namespace mvg
