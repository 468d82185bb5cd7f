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
//   3. one work item per (list, chunk of kFC rows, group of kFQ pairs): each
//      thread keeps one row in registers and scores it against the group's
//      queries held in LDS; a radix select keeps each query's top-k of the
//      chunk                                                    -> part
//   4. per-query merge of its probes' and chunks' winners       -> D, I
// Work items are taken from per-XCD counters over contiguous ranges, so the
// items that share a chunk of rows run on one XCD and share its L2.
// Rows of each list are stored in id order on the device, which makes the
// select's lowest-position rule for boundary ties the lowest-id rule.
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
// winners for the chunk slots its list does not have (all of them for an
// empty or invalid list).
__global__ void ivfflat_count_kernel(const int32_t *__restrict__ probes, int64_t npairs, int nlist,
                                     const int64_t *__restrict__ list_off, int *__restrict__ cnt,
                                     uint64_t *__restrict__ part, int nch_max, int k) {
  const int64_t pr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= npairs) return;
  const int l = probes[pr];
  int nch = 0;
  if (l >= 0 && l < nlist) {
    const int64_t len = list_off[l + 1] - list_off[l];
    nch = (int)((len + kFC - 1) / kFC);
    if (nch > 0) atomicAdd(&cnt[l], 1);
  }
  uint64_t *dst = part + (pr * nch_max + nch) * (int64_t)k;
  const int64_t m = (int64_t)(nch_max - nch) * k;
  for (int64_t i = 0; i < m; i++) dst[i] = kSentinel64;
}

// One workgroup: exclusive offsets of the pairs and work items of every list,
// totals at index nlist; clears the fill cursors and the range counters.
__global__ __launch_bounds__(kNT) void ivfflat_offsets_kernel(
    const int *__restrict__ cnt, const int64_t *__restrict__ list_off, int nlist,
    int *__restrict__ pair_off, int *__restrict__ tile_off, int *__restrict__ cursor,
    unsigned *__restrict__ range_ctr) {
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
        tc[j] = ((c + kFQ - 1) / kFQ) * nch;
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

template <int D, bool IP>
__global__ __launch_bounds__(kNT) void ivfflat_scan_kernel(
    const float *__restrict__ q, int d, int np, int nlist, const int64_t *__restrict__ list_off,
    const float *__restrict__ rows, const int32_t *__restrict__ ids, const int *__restrict__ cnt,
    const int *__restrict__ pair_off, const int *__restrict__ tile_off,
    const int32_t *__restrict__ pairs, unsigned *__restrict__ range_ctr, int k,
    const uint8_t *__restrict__ excl, int64_t excl_len, uint64_t *__restrict__ part, int nch_max) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *qs = reinterpret_cast<float *>(smem);                  // kFQ * d
  uint32_t *keys = reinterpret_cast<uint32_t *>(qs + kFQ * d);   // kFQ * kFC
  uint32_t *sel = keys + kFQ * kFC;                              // k
  SelectScratch *scr = reinterpret_cast<SelectScratch *>(sel + ((k + 3) & ~3));
  __shared__ int s_pair[kFQ];
  __shared__ int s_item;

  const int tid = threadIdx.x;
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
    const int ch = tl / npt, pt = tl - ch * npt;  // chunk-major: neighbours share rows
    const int qcount = min(kFQ, c - pt * kFQ);
    const int64_t r0 = list_off[l] + (int64_t)ch * kFC;
    const int clen = (int)min<int64_t>(kFC, list_off[l + 1] - r0);
    if (tid < kFQ) s_pair[tid] = tid < qcount ? pairs[pair_off[l] + pt * kFQ + tid] : -1;
    __syncthreads();
    for (int i = tid; i < kFQ * d; i += kNT) {
      const int qi = i / d;
      const int pr = s_pair[qi];
      qs[i] = pr >= 0 ? q[(int64_t)(pr / np) * d + (i - qi * d)] : 0.0f;
    }
    __syncthreads();
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
    __syncthreads();
    for (int qi = 0; qi < qcount; qi++) {
      const int got = select_topk_u32(keys + qi * kFC, clen, k, *scr, sel);
      uint64_t *dst = part + ((int64_t)s_pair[qi] * nch_max + ch) * k;
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

size_t scan_smem(int d, int k) {
  return sizeof(float) * kFQ * d + sizeof(uint32_t) * kFQ * kFC + sizeof(uint32_t) * ((k + 3) & ~3) +
         sizeof(SelectScratch);
}

template <int D, bool IP>
void launch_scan(const IvfFlatArgs &a, const float *q, int np, const IvfFlatWork &w, int nch_max,
                 hipStream_t s) {
  const size_t sm = scan_smem(a.d, a.k);
  const void *fn = (const void *)ivfflat_scan_kernel<D, IP>;
  if (sm > 64 * 1024)
    MVG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  int per_cu = 0;
  MVG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kNT, sm));
  const int grid = std::max(1, per_cu) * device_cus() / kXcds * kXcds;
  hipLaunchKernelGGL((ivfflat_scan_kernel<D, IP>), dim3((unsigned)std::max(grid, kXcds)), dim3(kNT),
                     sm, s, q, a.d, np, a.nlist, a.list_off, a.rows, a.ids, w.cnt.as<int>(),
                     w.pair_off.as<int>(), w.tile_off.as<int>(), w.pairs.as<int32_t>(),
                     w.range_ctr.as<unsigned>(), a.k, a.excl, a.excl_len, w.part.as<uint64_t>(),
                     nch_max);
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
  // Query batches: the per-probe winners fit the budget, pair indices and
  // the work-item count stay within int32.
  const int64_t per_query = (int64_t)np * nch_max * a.k * (int64_t)sizeof(uint64_t);
  int64_t qb = std::max<int64_t>(1, (int64_t)kPartBudget / per_query);
  qb = std::min<int64_t>(qb, (INT32_MAX / 2) / np);
  const int64_t fixed_items = a.nlist + a.ntotal / kFC + 1;
  const int64_t item_room = (INT32_MAX / 2 - fixed_items) / nch_max - 1;
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
    flat_topk(a.coarse, a.nlist, a.d, a.metric, qb_ptr, nb, np, nullptr, 0, w.coarse_part, 0.0f,
              nullptr, nullptr, probes, nullptr, s);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(a.nlist, 256)), dim3(256), 0, s,
                       w.cnt.as<int>(), (int64_t)a.nlist, 0);
    MVG_LAUNCH_CHECK("fill_i32_kernel");
    hipLaunchKernelGGL(ivfflat_count_kernel, dim3(ceil_div(npairs, 256)), dim3(256), 0, s, probes,
                       npairs, a.nlist, a.list_off, w.cnt.as<int>(), w.part.as<uint64_t>(), nch_max,
                       a.k);
    MVG_LAUNCH_CHECK("ivfflat_count_kernel");
    hipLaunchKernelGGL(ivfflat_offsets_kernel, dim3(1), dim3(kNT), 0, s, w.cnt.as<int>(),
                       a.list_off, a.nlist, w.pair_off.as<int>(), w.tile_off.as<int>(),
                       w.cursor.as<int>(), w.range_ctr.as<unsigned>());
    MVG_LAUNCH_CHECK("ivfflat_offsets_kernel");
    hipLaunchKernelGGL(ivfflat_fill_kernel, dim3(ceil_div(npairs, 256)), dim3(256), 0, s, probes,
                       npairs, a.nlist, a.list_off, w.pair_off.as<int>(), w.cursor.as<int>(),
                       w.pairs.as<int32_t>());
    MVG_LAUNCH_CHECK("ivfflat_fill_kernel");
#define MVG_IVFFLAT_CASE(DD)                                   \
  if (ip) launch_scan<DD, true>(a, qb_ptr, np, w, nch_max, s); \
  else launch_scan<DD, false>(a, qb_ptr, np, w, nch_max, s);
    switch (a.d) {
      case 32: MVG_IVFFLAT_CASE(32) break;
      case 64: MVG_IVFFLAT_CASE(64) break;
      case 96: MVG_IVFFLAT_CASE(96) break;
      case 128: MVG_IVFFLAT_CASE(128) break;
      default: MVG_IVFFLAT_CASE(0) break;
    }
#undef MVG_IVFFLAT_CASE
    launch_merge(w.part.as<uint64_t>(), nb, (int64_t)np * nch_max * a.k, a.k, ip ? 1 : 0,
                 kIvfPadDist, a.D + b * a.k, a.I + b * a.k, nullptr, nullptr, s);
  }
}

}  // namespace mvg
