// IVF-Flat search over FP32 inverted lists.
//
// Plain FP32 form of IndexIVFFlat.search (mojovec/index/index_ivf_flat.mojo:
// 139-244): the query's nprobe nearest centroids (IndexFlat top-nprobe), then
// every row of those lists scored with l2_distance_simd[8] or
// -inner_product_simd[4] in the reference's operation order (:222-225), and a
// top-k in the canonical (distance, lower id) order.
//
// Work is organised by list rather than by query so that a list's rows are
// read from HBM once per group of queries that probe it:
//   1. coarse top-nprobe per query (the Flat kernel)          -> probes[q][p]
//   2. probe inversion on the device: per-list pair counts, offsets and the
//      (query, probe) pairs of each list                        -> pairs
//   3. one work item per (list, group of kFQ pairs): chunk by chunk of kFC
//      rows, each thread keeps one row in registers and scores it against
//      the group's queries held in LDS; for k <= 64 each wavefront keeps the
//      running top-k of two queries in registers (bitonic networks, no
//      workgroup barrier), above that a workgroup radix select keeps each
//      chunk's top-k (then an item covers one chunk)          -> part
//   4. per-query merge of its probes' (and chunks') winners     -> D, I
// Work items are taken from per-XCD counters over contiguous ranges, so the
// items that share a chunk of rows run on one XCD and share its L2.
// Rows of each list are stored in id order on the device, which makes the
// position order inside a list (the selects' tie rule) the id order.
#include "dist.h"
#include "kernels.h"
#include "select.h"

namespace mvg {

namespace {

constexpr int kFQ = 8;      // (query, probe) pairs per work item
constexpr int kFC = 1024;   // list rows per chunk
constexpr int kFPer = kFC / kNT;
constexpr int kXcds = 8;    // work ranges (one per XCD of the MI355X)
constexpr size_t kPartBudget = size_t(1) << 30;  // bytes of per-probe winners per batch

template <int D, bool IP>
__device__ __forceinline__ float list_dist(const float *xr, const float *xg, const float *qs, int d) {
  if constexpr (D > 0) {
    return IP ? -ip_w4_fixed<D>(xr, qs) : l2_w8_fixed<D>(xr, qs);
  } else {
    return IP ? -ip_wn<4>(qs, xg, d) : l2_w8(qs, xg, d);
  }
}

__global__ void fill_i32_kernel(int *p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Per (query, probe) pair: counts the pair on its list and writes sentinel
// winners to the slots no work item will fill: with per-chunk slots (`wave`
// false) the chunks its list does not have, with one slot per pair only an
// empty or invalid list's.
__global__ void ivfflat_count_kernel(const int32_t *__restrict__ probes, int64_t npairs, int nlist,
                                     const int64_t *__restrict__ list_off, int *__restrict__ cnt,
                                     uint64_t *__restrict__ part, int nslots, int k, int wave) {
  const int64_t pr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= npairs) return;
  const int l = probes[pr];
  int filled = 0;
  if (l >= 0 && l < nlist) {
    const int64_t len = list_off[l + 1] - list_off[l];
    const int nch = (int)((len + kFC - 1) / kFC);
    filled = wave ? (nch > 0 ? 1 : 0) : nch;
    if (nch > 0) atomicAdd(&cnt[l], 1);
  }
  uint64_t *dst = part + (pr * nslots + filled) * (int64_t)k;
  const int64_t m = (int64_t)(nslots - filled) * k;
  for (int64_t i = 0; i < m; i++) dst[i] = kSentinel64;
}

// One workgroup: exclusive offsets of the pairs and work items of every list
// (one item per group of kFQ pairs, times the list's chunks unless the item
// walks every chunk itself), totals at index nlist; clears the fill cursors
// and the range counters.
__global__ __launch_bounds__(kNT) void ivfflat_offsets_kernel(
    const int *__restrict__ cnt, const int64_t *__restrict__ list_off, int nlist,
    int *__restrict__ pair_off, int *__restrict__ tile_off, int *__restrict__ cursor,
    unsigned *__restrict__ range_ctr, int wave) {
  __shared__ int sh[kNW + 1];
  const int tid = threadIdx.x;
  if (tid < kXcds) range_ctr[tid] = 0u;
  int pcarry = 0, tcarry = 0;
  for (int base = 0; base < nlist; base += kNT * 4) {
    int pc[4], tc[4], ps = 0, ts = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int l = base + tid * 4 + j;
      pc[j] = tc[j] = 0;
      if (l < nlist) {
        const int c = cnt[l];
        const int64_t len = list_off[l + 1] - list_off[l];
        const int nch = (int)((len + kFC - 1) / kFC);
        pc[j] = c;
        tc[j] = ((c + kFQ - 1) / kFQ) * (wave ? (nch > 0 ? 1 : 0) : nch);
        cursor[l] = 0;
      }
      ps += pc[j];
      ts += tc[j];
    }
    int ptot, ttot;
    int pex = block_excl_scan(ps, sh, ptot);
    int tex = block_excl_scan(ts, sh, ttot);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int l = base + tid * 4 + j;
      if (l < nlist) {
        pair_off[l] = pcarry + pex;
        tile_off[l] = tcarry + tex;
      }
      pex += pc[j];
      tex += tc[j];
    }
    pcarry += ptot;
    tcarry += ttot;
  }
  if (tid == 0) {
    pair_off[nlist] = pcarry;
    tile_off[nlist] = tcarry;
  }
}

// Places every valid pair in its list's segment (order within a list is
// free: each pair's winners land in its own slot).
__global__ void ivfflat_fill_kernel(const int32_t *__restrict__ probes, int64_t npairs, int nlist,
                                    const int64_t *__restrict__ list_off,
                                    const int *__restrict__ pair_off, int *__restrict__ cursor,
                                    int32_t *__restrict__ pairs) {
  const int64_t pr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= npairs) return;
  const int l = probes[pr];
  if (l < 0 || l >= nlist || list_off[l + 1] == list_off[l]) return;
  pairs[pair_off[l] + atomicAdd(&cursor[l], 1)] = (int32_t)pr;
}

// Scores rows r0 .. r0+clen of one list against the item's queries: one row
// per thread in registers, the queries broadcast from LDS; u32 keys to
// keys[qi][local] (sentinel for rows past clen, excluded ids, absent queries).
template <int D, bool IP>
__device__ __forceinline__ void score_chunk(const float *__restrict__ rows,
                                            const int32_t *__restrict__ ids, int64_t r0, int clen,
                                            int d, const float *qs, int qcount,
                                            const uint8_t *__restrict__ excl, int64_t excl_len,
                                            uint32_t *keys) {
  const int tid = threadIdx.x;
  for (int r = 0; r < kFPer; r++) {
    const int local = r * kNT + tid;
    const int64_t j = r0 + local;
    bool live = local < clen;
    if (live && excl) {
      const int32_t id = ids[j];
      live = !(id < excl_len && excl[id] != 0);
    }
    float xr[D > 0 ? D : 1];
    const float *xg = rows + (live ? j : r0) * d;
    if constexpr (D > 0) {
      if (live) {
#pragma unroll
        for (int e = 0; e < D; e += 4) {
          const float4 v = *reinterpret_cast<const float4 *>(xg + e);
          xr[e] = v.x; xr[e + 1] = v.y; xr[e + 2] = v.z; xr[e + 3] = v.w;
        }
      }
    }
#pragma unroll
    for (int qi = 0; qi < kFQ; qi++) {
      uint32_t key = kSentinel;
      if (live && qi < qcount) key = f2key(list_dist<D, IP>(xr, xg, qs + qi * d, d));
      keys[qi * kFC + local] = key;
    }
  }
}

// Work items from per-XCD ranges. WAVE (k <= 64): an item is (list, group of
// kFQ pairs) and walks every chunk of the list; wave w keeps the running top
// k of queries w and w + 4 in registers (wave_fold_keys) and the item writes
// one slot of k winners per pair. Otherwise an item is (list, chunk, group)
// and a workgroup radix select writes the chunk's k winners per pair.
template <int D, bool IP, bool WAVE>
__global__ __launch_bounds__(kNT) void ivfflat_scan_kernel(
    const float *__restrict__ q, int d, int np, int nlist, const int64_t *__restrict__ list_off,
    const float *__restrict__ rows, const int32_t *__restrict__ ids, const int *__restrict__ cnt,
    const int *__restrict__ pair_off, const int *__restrict__ tile_off,
    const int32_t *__restrict__ pairs, unsigned *__restrict__ range_ctr, int k,
    const uint8_t *__restrict__ excl, int64_t excl_len, uint64_t *__restrict__ part, int nslots) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *qs = reinterpret_cast<float *>(smem);                  // kFQ * d
  uint32_t *keys = reinterpret_cast<uint32_t *>(qs + kFQ * d);   // kFQ * kFC
  uint32_t *sel = keys + kFQ * kFC;                              // k (radix form)
  SelectScratch *scr = reinterpret_cast<SelectScratch *>(sel + ((k + 3) & ~3));
  __shared__ int s_pair[kFQ];
  __shared__ int s_item;

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int range = blockIdx.x % kXcds;  // consecutive workgroups go to consecutive XCDs
  const int total = tile_off[nlist];
  const int beg = (int)((int64_t)total * range / kXcds);
  const int end = (int)((int64_t)total * (range + 1) / kXcds);
  for (;;) {
    if (tid == 0) s_item = beg + (int)atomicAdd(&range_ctr[range], 1u);
    __syncthreads();
    const int t = s_item;
    if (t >= end) break;
    // The list of item t: the last l with tile_off[l] <= t (lists without
    // items share the offset of the next list and are skipped this way).
    int lo = 0, hi = nlist - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tile_off[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int l = lo;
    const int c = cnt[l];
    const int npt = (c + kFQ - 1) / kFQ;
    const int tl = t - tile_off[l];
    const int ch = WAVE ? 0 : tl / npt;  // chunk-major: neighbouring items share rows
    const int pt = tl - ch * npt;
    const int qcount = min(kFQ, c - pt * kFQ);
    const int64_t lb = list_off[l], len = list_off[l + 1] - lb;
    if (tid < kFQ) s_pair[tid] = tid < qcount ? pairs[pair_off[l] + pt * kFQ + tid] : -1;
    __syncthreads();
    for (int i = tid; i < kFQ * d; i += kNT) {
      const int qi = i / d;
      const int pr = s_pair[qi];
      qs[i] = pr >= 0 ? q[(int64_t)(pr / np) * d + (i - qi * d)] : 0.0f;
    }
    __syncthreads();
    if constexpr (WAVE) {
      uint64_t best0 = kSentinel64, best1 = kSentinel64;  // queries wave, wave + kNW
      for (int64_t c0 = 0; c0 < len; c0 += kFC) {
        const int clen = (int)min<int64_t>(kFC, len - c0);
        score_chunk<D, IP>(rows, ids, lb + c0, clen, d, qs, qcount, excl, excl_len, keys);
        __syncthreads();
        if (wave < qcount)
          best0 = wave_fold_keys<kFC>(keys + wave * kFC, clen, (uint32_t)c0, best0, k);
        if (wave + kNW < qcount)
          best1 = wave_fold_keys<kFC>(keys + (wave + kNW) * kFC, clen, (uint32_t)c0, best1, k);
        __syncthreads();
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int qi = wave + h * kNW;
        const uint64_t v = h ? best1 : best0;
        if (qi < qcount && lane < k) {
          part[(int64_t)s_pair[qi] * k + lane] =
              v == kSentinel64 ? kSentinel64
                               : ((v >> 32) << 32) | (uint32_t)ids[lb + (uint32_t)(v & 0xFFFFFFFFu)];
        }
      }
    } else {
      const int64_t r0 = lb + (int64_t)ch * kFC;
      const int clen = (int)min<int64_t>(kFC, len - (int64_t)ch * kFC);
      score_chunk<D, IP>(rows, ids, r0, clen, d, qs, qcount, excl, excl_len, keys);
      __syncthreads();
      for (int qi = 0; qi < qcount; qi++) {
        const int got = select_topk_u32(keys + qi * kFC, clen, k, *scr, sel);
        uint64_t *dst = part + ((int64_t)s_pair[qi] * nslots + ch) * k;
        for (int u = tid; u < k; u += kNT) {
          uint64_t v = kSentinel64;
          if (u < got) {
            const uint32_t idx = sel[u];
            v = ((uint64_t)keys[qi * kFC + idx] << 32) | (uint32_t)ids[r0 + idx];
          }
          dst[u] = v;
        }
        __syncthreads();
      }
    }
  }
}

size_t scan_smem(int d, int k) {
  return sizeof(float) * kFQ * d + sizeof(uint32_t) * kFQ * kFC + sizeof(uint32_t) * ((k + 3) & ~3) +
         sizeof(SelectScratch);
}

template <int D, bool IP, bool WAVE>
void launch_scan(const IvfFlatArgs &a, const float *q, int np, const IvfFlatWork &w, int nslots,
                 hipStream_t s) {
  const size_t sm = scan_smem(a.d, a.k);
  const void *fn = (const void *)ivfflat_scan_kernel<D, IP, WAVE>;
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  int per_cu = 0;
  MVG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kNT, sm));
  const int grid = std::max(1, per_cu) * device_cus() / kXcds * kXcds;
  hipLaunchKernelGGL((ivfflat_scan_kernel<D, IP, WAVE>), dim3((unsigned)std::max(grid, kXcds)),
                     dim3(kNT), sm, s, q, a.d, np, a.nlist, a.list_off, a.rows, a.ids,
                     w.cnt.as<int>(), w.pair_off.as<int>(), w.tile_off.as<int>(),
                     w.pairs.as<int32_t>(), w.range_ctr.as<unsigned>(), a.k, a.excl, a.excl_len,
                     w.part.as<uint64_t>(), nslots);
  MVG_LAUNCH_CHECK("ivfflat_scan_kernel");
}

}  // namespace

void ivfflat_search(const IvfFlatArgs &a, IvfFlatWork &w, hipStream_t s) {
  if (a.nq == 0) return;
  if (a.k > kMaxCandidates) throw Error(kInvalid, "k exceeds the device top-k limit of 4096");
  const int np = std::min(a.nprobe, a.nlist);
  if (np < 1) throw Error(kInvalid, "nprobe must be at least 1");
  if (np > kMaxCandidates) throw Error(kInvalid, "nprobe above 4096 is not supported");
  if (scan_smem(a.d, a.k) > 150 * 1024) throw Error(kInvalid, "dimension too large for IVF-Flat");
  const int nch_max = std::max(1, (int)((a.max_list + kFC - 1) / kFC));
  // k <= 64: one item walks a list's chunks keeping a wavefront top-k, one
  // slot of winners per (query, probe); above, one item and slot per chunk.
  const bool wave = a.k <= 64;
  const int nslots = wave ? 1 : nch_max;
  // Query batches: the per-probe winners fit the budget, pair indices and
  // the work-item count stay within int32.
  const int64_t per_query = (int64_t)np * nslots * a.k * (int64_t)sizeof(uint64_t);
  int64_t qb = std::max<int64_t>(1, (int64_t)kPartBudget / per_query);
  qb = std::min<int64_t>(qb, (INT32_MAX / 2) / np);
  const int64_t fixed_items = a.nlist + a.ntotal / kFC + 1;
  const int64_t item_room = (INT32_MAX / 2 - fixed_items) / nslots - 1;
  if (item_room < 1) throw Error(kInvalid, "IVF-Flat lists too large for one work range");
  qb = std::max<int64_t>(1, std::min<int64_t>(qb, item_room * kFQ / np));
  qb = std::min<int64_t>(qb, a.nq);

  const bool ip = a.metric == kMetricIP;
  w.probes.reserve(sizeof(int32_t) * (size_t)qb * np);
  w.cnt.reserve(sizeof(int) * (size_t)a.nlist);
  w.cursor.reserve(sizeof(int) * (size_t)a.nlist);
  w.pair_off.reserve(sizeof(int) * (size_t)(a.nlist + 1));
  w.tile_off.reserve(sizeof(int) * (size_t)(a.nlist + 1));
  w.range_ctr.reserve(sizeof(unsigned) * kXcds);
  w.pairs.reserve(sizeof(int32_t) * (size_t)qb * np);
  w.part.reserve((size_t)(qb * per_query));
  int32_t *probes = w.probes.as<int32_t>();
  for (int64_t b = 0; b < a.nq; b += qb) {
    const int64_t nb = std::min<int64_t>(qb, a.nq - b);
    const float *qb_ptr = a.q + b * a.d;
    const int64_t npairs = nb * np;
    flat_topk(a.coarse, a.nlist, a.d, a.coarse_metric, qb_ptr, nb, np, nullptr, 0, w.coarse_part, 0.0f,
              nullptr, nullptr, probes, nullptr, s);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(a.nlist, 256)), dim3(256), 0, s,
                       w.cnt.as<int>(), (int64_t)a.nlist, 0);
    MVG_LAUNCH_CHECK("fill_i32_kernel");
    hipLaunchKernelGGL(ivfflat_count_kernel, dim3(ceil_div(npairs, 256)), dim3(256), 0, s, probes,
                       npairs, a.nlist, a.list_off, w.cnt.as<int>(), w.part.as<uint64_t>(), nslots,
                       a.k, wave ? 1 : 0);
    MVG_LAUNCH_CHECK("ivfflat_count_kernel");
    hipLaunchKernelGGL(ivfflat_offsets_kernel, dim3(1), dim3(kNT), 0, s, w.cnt.as<int>(),
                       a.list_off, a.nlist, w.pair_off.as<int>(), w.tile_off.as<int>(),
                       w.cursor.as<int>(), w.range_ctr.as<unsigned>(), wave ? 1 : 0);
    MVG_LAUNCH_CHECK("ivfflat_offsets_kernel");
    hipLaunchKernelGGL(ivfflat_fill_kernel, dim3(ceil_div(npairs, 256)), dim3(256), 0, s, probes,
                       npairs, a.nlist, a.list_off, w.pair_off.as<int>(), w.cursor.as<int>(),
                       w.pairs.as<int32_t>());
    MVG_LAUNCH_CHECK("ivfflat_fill_kernel");
#define MVG_IVFFLAT_CASE(DD)                                               \
  if (ip && wave) launch_scan<DD, true, true>(a, qb_ptr, np, w, nslots, s);        \
  else if (ip) launch_scan<DD, true, false>(a, qb_ptr, np, w, nslots, s);          \
  else if (wave) launch_scan<DD, false, true>(a, qb_ptr, np, w, nslots, s);        \
  else launch_scan<DD, false, false>(a, qb_ptr, np, w, nslots, s);
    switch (a.d) {
      case 32: MVG_IVFFLAT_CASE(32) break;
      case 64: MVG_IVFFLAT_CASE(64) break;
      case 96: MVG_IVFFLAT_CASE(96) break;
      case 128: MVG_IVFFLAT_CASE(128) break;
      default: MVG_IVFFLAT_CASE(0) break;
    }
#undef MVG_IVFFLAT_CASE
    launch_merge(w.part.as<uint64_t>(), nb, (int64_t)np * nslots * a.k, a.k, ip ? 1 : 0,
                 kIvfPadDist, a.D + b * a.k, a.I + b * a.k, nullptr, nullptr, s);
  }
}

}  // namespace mvg
