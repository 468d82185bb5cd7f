// Workgroup-level top-k selection in LDS.
//
// Candidates are ordered canonically: by distance, exact ties broken by the
// lower id (the reference's heap, mojovec/utils/heap.mojo, keeps the same
// multiset of distances but is not tie-stable; the canonical order is the
// build's documented tie rule). Distances become order-preserving u32 keys
// (f2key); within one inverted list or one base chunk the position order is
// the id order, so (key, position) is the canonical order there. Across
// lists the ids travel with the key as a u64 (key << 32 | id).
//
// Selection is a most-significant-digit radix select: a few LDS histogram
// passes find the k-th key, then an order-preserving compaction emits the
// winners. Cost is linear in the candidate count; no sorting network runs on
// the full candidate set.
#pragma once

#include "common.h"

namespace mvg {

constexpr int kNT = 256;  // threads per workgroup for every select user
constexpr int kNW = kNT / 64;

// Exclusive scan of one int per thread over the workgroup; `sh` needs kNW+1
// ints. All threads must call it.
__device__ __forceinline__ int block_excl_scan(int v, int *sh, int &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kNW; w++) {
    int t = sh[w];
    if (w < wave) off += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

__device__ __forceinline__ int block_sum(int v, int *sh) {
  int total;
  (void)block_excl_scan(v, sh, total);
  return total;
}

// Scratch for the selects: a 2048-bin histogram and a few scalars.
struct SelectScratch {
  int hist[2048];
  int sh[16];
};

// Scratch for the u64 selects only (8-bit digits).
struct SelectScratch64 {
  int hist[256];
  int sh[16];
};

// k-th smallest valid u32 key (1 <= k <= number of valid keys). Returns the
// key T and how many keys equal to T belong to the k smallest.
__device__ void radix_kth_u32(const uint32_t *keys, int n, int k, SelectScratch &s,
                              uint32_t &T, int &need_eq) {
  const int tid = threadIdx.x;
  uint32_t prefix = 0, mask = 0;
  int remaining = k;
  const int shifts[3] = {21, 10, 0};
  const int nbits[3] = {11, 11, 10};
  for (int p = 0; p < 3; p++) {
    const int nb = 1 << nbits[p];
    const uint32_t dm = (uint32_t)nb - 1u;
    for (int b = tid; b < nb; b += kNT) s.hist[b] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += kNT) {
      uint32_t key = keys[i];
      if (key != kSentinel && (key & mask) == prefix)
        atomicAdd(&s.hist[(key >> shifts[p]) & dm], 1);
    }
    __syncthreads();
    const int per = nb / kNT;
    int local = 0;
    for (int b = 0; b < per; b++) local += s.hist[tid * per + b];
    int total;
    int excl = block_excl_scan(local, s.sh, total);
    if (excl < remaining && remaining <= excl + local) {
      int c = excl;
      for (int b = 0; b < per; b++) {
        int h = s.hist[tid * per + b];
        if (c + h >= remaining) {
          s.sh[8] = tid * per + b;
          s.sh[9] = c;
          break;
        }
        c += h;
      }
    }
    __syncthreads();
    const int bin = s.sh[8];
    remaining -= s.sh[9];
    prefix |= (uint32_t)bin << shifts[p];
    mask |= dm << shifts[p];
    __syncthreads();
  }
  T = prefix;
  need_eq = remaining;
}

// Selects the min(k, valid) smallest keys of keys[0..n) under (key, index)
// order and writes their indices, in index order, to out[]. Returns the
// number written. Sentinel keys are never selected.
__device__ int select_topk_u32(const uint32_t *keys, int n, int k, SelectScratch &s,
                               uint32_t *out) {
  const int tid = threadIdx.x;
  const int per = (n + kNT - 1) / kNT;
  const int beg = min(tid * per, n), end = min(beg + per, n);
  int valid = 0;
  for (int i = beg; i < end; i++) valid += keys[i] != kSentinel;
  const int nvalid = block_sum(valid, s.sh);
  const bool take_all = nvalid <= k;
  uint32_t T = 0;
  int need_eq = 0;
  if (!take_all) radix_kth_u32(keys, n, k, s, T, need_eq);
  int lt = 0, eq = 0;
  for (int i = beg; i < end; i++) {
    uint32_t key = keys[i];
    if (key == kSentinel) continue;
    if (take_all || key < T) lt++;
    else if (key == T) eq++;
  }
  int tot;
  const int eq_before = block_excl_scan(eq, s.sh, tot);
  const int my_eq = max(0, min(eq, need_eq - eq_before));
  int total;
  int off = block_excl_scan(lt + my_eq, s.sh, total);
  int e = eq_before;
  for (int i = beg; i < end; i++) {
    uint32_t key = keys[i];
    if (key == kSentinel) continue;
    if (take_all || key < T) {
      out[off++] = (uint32_t)i;
    } else if (key == T) {
      if (e < need_eq) out[off++] = (uint32_t)i;
      e++;
    }
  }
  __syncthreads();
  return total;
}

// k-th smallest of unique u64 keys (1 <= k <= number of valid keys).
template <typename Scratch>
__device__ uint64_t radix_kth_u64(const uint64_t *keys, int n, int k, Scratch &s) {
  const int tid = threadIdx.x;
  uint64_t prefix = 0, mask = 0;
  int remaining = k;
  for (int p = 0; p < 8; p++) {
    const int shift = 56 - 8 * p;
    s.hist[tid] = 0;  // 256 bins == kNT
    __syncthreads();
    for (int i = tid; i < n; i += kNT) {
      uint64_t key = keys[i];
      if (key != kSentinel64 && (key & mask) == prefix)
        atomicAdd(&s.hist[(int)((key >> shift) & 0xFF)], 1);
    }
    __syncthreads();
    const int local = s.hist[tid];
    int total;
    const int excl = block_excl_scan(local, s.sh, total);
    if (excl < remaining && remaining <= excl + local) {
      s.sh[8] = tid;
      s.sh[9] = excl;
    }
    __syncthreads();
    const int bin = s.sh[8];
    remaining -= s.sh[9];
    prefix |= (uint64_t)bin << shift;
    mask |= (uint64_t)0xFF << shift;
    __syncthreads();
  }
  return prefix;
}

// Ascending bitonic sort of keys[0..n) in LDS, n a power of two.
__device__ void bitonic_sort_u64(uint64_t *keys, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += kNT) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint64_t a = keys[i], b = keys[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Smallest power of two >= v (v >= 1): the output size select_sort_u64 needs.
__host__ __device__ inline int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Writes the min(k, valid) smallest unique u64 keys of in[0..n) to out[] in
// ascending order and sentinels to the rest of out[0..k); out must hold
// next_pow2(k) entries and must not overlap in. Returns the count.
template <typename Scratch>
__device__ int select_sort_u64(const uint64_t *in, int n, int k, Scratch &s,
                               uint64_t *out) {
  const int tid = threadIdx.x;
  int valid = 0;
  for (int i = tid; i < n; i += kNT) valid += in[i] != kSentinel64;
  const int nvalid = block_sum(valid, s.sh);
  const int cnt = min(k, nvalid);
  uint64_t T = kSentinel64 - 1;
  if (nvalid > k) T = radix_kth_u64(in, n, k, s);
  if (tid == 0) s.sh[10] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kNT) {
    const uint64_t key = in[i];
    if (key != kSentinel64 && key <= T) out[atomicAdd(&s.sh[10], 1)] = key;
  }
  __syncthreads();
  int np2 = 1;
  while (np2 < cnt) np2 <<= 1;
  // Callers read out[0..k): every slot past the count holds a sentinel.
  const int fill_to = max(np2, k);
  for (int i = cnt + tid; i < fill_to; i += kNT) out[i] = kSentinel64;
  __syncthreads();
  bitonic_sort_u64(out, np2);
  return cnt;
}

// ---------------------------------------------------------------------------
// Wavefront top-k for k <= 64 over unique u64 composites (one value per lane,
// sentinel = kSentinel64): bitonic networks across the 64 lanes, no LDS and
// no workgroup barrier.

// Value of lane `lane ^ j` (j a power of two below 64, a constant after
// unrolling). Strides 1, 2 and 8 are DPP moves (quad permutes, row rotate by
// 8); stride 4 takes whichever of the two row rotates by 4 and 12 brings lane
// ^ 4 (checked on the lane ids, so either rotate direction is right); 16 and
// 32 go through the LDS permute.
__device__ __forceinline__ uint32_t dpp_xor32(uint32_t v, int j) {
  if (j == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
  if (j == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
  if (j == 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);
  if (j == 4) {
    const int lane = threadIdx.x & 63;
    const int from4 = __builtin_amdgcn_mov_dpp(lane, 0x124, 0xF, 0xF, false);
    const uint32_t a = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, false);
    const uint32_t b = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x12C, 0xF, 0xF, false);
    return from4 == (lane ^ 4) ? a : b;
  }
  return (uint32_t)__shfl_xor((int)v, j, 64);
}
__device__ __forceinline__ uint64_t xor_lane64(uint64_t v, int j) {
  const uint32_t lo = dpp_xor32((uint32_t)v, j), hi = dpp_xor32((uint32_t)(v >> 32), j);
  return ((uint64_t)hi << 32) | lo;
}

// Ascending bitonic sort of the wave's 64 values.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t p = xor_lane64(v, j);
      const bool asc = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      v = (lower == asc) ? (p < v ? p : v) : (p > v ? p : v);
    }
  }
  return v;
}

// The 64 smallest of two ascending sequences, ascending (Batcher: the
// lane-wise minimum of a and reversed b is bitonic; half-cleaners sort it).
__device__ __forceinline__ uint64_t wave_merge64(uint64_t a, uint64_t b) {
  const int lane = threadIdx.x & 63;
  const uint64_t r = __shfl(b, 63 - lane, 64);
  uint64_t v = a < r ? a : r;
#pragma unroll
  for (int j = 32; j > 0; j >>= 1) {
    const uint64_t p = xor_lane64(v, j);
    v = (lane & j) == 0 ? (p < v ? p : v) : (p > v ? p : v);
  }
  return v;
}

// Folds one block of CHUNK (a multiple of 64) u32 keys into a wave's running
// ascending set `best` of composites key << 32 | (pos0 + index); the k
// smallest of the union end in lanes 0..k-1. Keys past `len` and sentinel
// keys are ignored. A key can only enter the top k if it is at most the
// current k-th best, or -- while fewer than k are held -- at most the k-th
// smallest of the lanes' two smallest composites (at least k keys lie at or
// below that, and few more unless a lane holds three of the block's k
// smallest). When the keys under the bound fit a wavefront they are compacted
// into the block's first 64 composites (the block is overwritten: the caller's
// wavefront owns it and reads it no more) and sorted once; otherwise they are
// merged 64 at a time.
template <int CHUNK>
__device__ uint64_t wave_fold_keys(uint32_t *keys, int len, uint32_t pos0, uint64_t best, int k) {
  const int lane = threadIdx.x & 63;
  constexpr int PER = CHUNK / 64;
  uint32_t kr[PER];
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int idx = lane + 64 * j;
    kr[j] = idx < len ? keys[idx] : kSentinel;
  }
  auto comp = [&](int j) { return ((uint64_t)kr[j] << 32) | (pos0 + (uint32_t)(lane + 64 * j)); };
  uint64_t theta = __shfl(best, k - 1, 64);
  if (theta == kSentinel64) {
    uint64_t m1 = kSentinel64, m2 = kSentinel64;
#pragma unroll
    for (int j = 0; j < PER; j++) {
      if (kr[j] == kSentinel) continue;
      const uint64_t c = comp(j);
      if (c < m1) {
        m2 = m1;
        m1 = c;
      } else if (c < m2) {
        m2 = c;
      }
    }
    theta = __shfl(wave_merge64(wave_sort64(m1), wave_sort64(m2)), k - 1, 64);
  }
  uint32_t mask = 0;
#pragma unroll
  for (int j = 0; j < PER; j++)
    if (kr[j] != kSentinel && comp(j) <= theta) mask |= 1u << j;
  const int mine = __popc(mask);
  int incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  const int total = __shfl(incl, 63, 64);
  if (total == 0) return best;
  if (total <= 64) {
    // every lane's keys are in registers: the block's first 64 composites
    // take the candidates (a wavefront's LDS operations complete in order)
    uint64_t *slot = reinterpret_cast<uint64_t *>(keys);
    int at = incl - mine;
#pragma unroll
    for (int j = 0; j < PER; j++)
      if ((mask >> j) & 1u) slot[at++] = comp(j);
    const uint64_t c = lane < total ? slot[lane] : kSentinel64;
    return wave_merge64(best, wave_sort64(c));
  }
  while (__ballot(mask != 0) != 0) {
    uint64_t c = kSentinel64;
    if (mask) {
      const int j = __builtin_ctz(mask);
      mask &= mask - 1;
      c = ((uint64_t)keys[lane + 64 * j] << 32) | (pos0 + (uint32_t)(lane + 64 * j));
    }
    best = wave_merge64(best, wave_sort64(c));
  }
  return best;
}

// ---------------------------------------------------------------------------
// Unordered workgroup selection of the w smallest u64 keys by range buckets.

struct BucketScratch {
  uint64_t bucket[64];
  unsigned long long red[2 * kNW];
  int hist[1024];
  int sh[16];
  int nout, nb, bin, before, cnt;
};

// Writes the min(w, valid) smallest keys of in[0..n) (sentinels excluded;
// equal keys are interchangeable) to out[0..cnt) in no particular order and
// returns cnt; out must hold w keys and must not overlap in. The keys still
// in play lie in [lo, hi], starting at the smallest and largest valid key. A
// pass counts them in 1024 bins of (key - lo) >> sh, sh the least shift that
// fits hi - lo into 10 bits, writes out every key below the bin where the
// running count reaches the number still needed, and keeps that bin in play;
// it ends when the bin holds exactly that number (all written), at most 64
// keys (one wavefront sorts them and writes the smallest) or one value (any
// of them). Each pass narrows the range 1024-fold: at most seven passes.
// All threads must call it.
__device__ int bucket_select_u64(const uint64_t *in, int n, int w, uint64_t *out, BucketScratch &s) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t mn = kSentinel64, mx = 0;
  int valid = 0;
  for (int i = tid; i < n; i += kNT) {
    const uint64_t k = in[i];
    if (k != kSentinel64) {
      mn = k < mn ? k : mn;
      mx = k > mx ? k : mx;
      valid++;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
    valid += __shfl_xor(valid, o, 64);
  }
  if (lane == 0) {
    s.red[wave] = mn;
    s.red[kNW + wave] = mx;
    s.sh[wave] = valid;
  }
  if (tid == 0) s.nout = 0;
  __syncthreads();
  uint64_t lo = kSentinel64, hi = 0;
  int nvalid = 0;
#pragma unroll
  for (int v = 0; v < kNW; v++) {
    lo = s.red[v] < lo ? s.red[v] : lo;
    hi = s.red[kNW + v] > hi ? s.red[kNW + v] : hi;
    nvalid += s.sh[v];
  }
  __syncthreads();
  if (nvalid <= w) {  // every valid key
    for (int i = tid; i < n; i += kNT) {
      const uint64_t k = in[i];
      if (k != kSentinel64) out[atomicAdd(&s.nout, 1)] = k;
    }
    __syncthreads();
    return nvalid;
  }
  int need = w;  // keys still to write; [lo, hi] holds at least that many
  for (;;) {
    const uint64_t width = hi - lo;
    if (width == 0) {  // one value: any `need` of its keys
      if (tid == 0) s.nb = 0;
      __syncthreads();
      const int base = s.nout;
      for (int i = tid; i < n; i += kNT) {
        if (in[i] == lo) {
          const int slot = atomicAdd(&s.nb, 1);
          if (slot < need) out[base + slot] = lo;
        }
      }
      __syncthreads();
      break;
    }
    const int sh = (width >> 10) == 0 ? 0 : 64 - __clzll((long long)width) - 10;
    for (int b = tid; b < 1024; b += kNT) s.hist[b] = 0;
    if (tid == 0) s.nb = 0;
    __syncthreads();
    for (int i = tid; i < n; i += kNT) {
      const uint64_t k = in[i];
      if (k != kSentinel64 && k >= lo && k <= hi) atomicAdd(&s.hist[(int)((k - lo) >> sh)], 1);
    }
    __syncthreads();
    int c4[4], loc = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      c4[j] = s.hist[4 * tid + j];
      loc += c4[j];
    }
    int total;
    const int ex = block_excl_scan(loc, s.sh, total);
    if (ex < need && ex + loc >= need) {
      int run = ex;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (run + c4[j] >= need) {
          s.bin = 4 * tid + j;
          s.before = run;
          s.cnt = c4[j];
          break;
        }
        run += c4[j];
      }
    }
    __syncthreads();
    const int b = s.bin, before = s.before, cb = s.cnt;
    const uint64_t blo = lo + ((uint64_t)b << sh);
    const uint64_t bhi = (b == 1023 || ((uint64_t)(b + 1) << sh) - 1 >= width) ? hi
                                                                              : lo + ((uint64_t)(b + 1) << sh) - 1;
    need -= before;
    const bool all = cb == need, small = cb <= 64;
    for (int i = tid; i < n; i += kNT) {
      const uint64_t k = in[i];
      if (k == kSentinel64 || k < lo || k > bhi) continue;
      if (k < blo || all) out[atomicAdd(&s.nout, 1)] = k;
      else if (small) s.bucket[atomicAdd(&s.nb, 1)] = k;
    }
    __syncthreads();
    if (all) break;
    if (small) {
      if (wave == 0) {
        uint64_t v = lane < s.nb ? s.bucket[lane] : kSentinel64;
        v = wave_sort64(v);
        const int base = s.nout;
        if (lane < need) out[base + lane] = v;
      }
      __syncthreads();
      if (tid == 0) s.nout += need;
      __syncthreads();
      break;
    }
    lo = blo;
    hi = bhi;
  }
  return w;
}

}  // namespace mvg
