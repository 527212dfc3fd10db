// Pippenger multi-scalar multiplication on G1 for gfx950.
//
// Replaces curves/src/msm.rs:460-532 (msm_best) and proofs/src/poly/kzg/msm.rs:174-201
// (msm_specific).  The result is the unique point sum_i s_i P_i; the reference's
// Booth digits and batch-affine buckets are a CPU schedule, re-planned here for the
// GPU:
//
//  1. k_sort_hist     scalars -> canonical form, signed c-bit window digits (one
//                     thread per scalar), per-tile LDS histogram of the top key bits
//  2. scan            exclusive prefix sum over (bin, tile) counts
//  3. k_sort_scatter  (key low bits, point reference) pairs into their top-bit bins,
//                     positions from LDS atomics (no global atomics)
//  4. k_sort_bins     per bin, a counting sort on the low key bits: references
//                     grouped by bucket, and the bucket offsets
//  5. k_msm_accum     each thread adds a fixed chunk of K sorted references into XYZZ
//                     partial bucket sums (mixed additions, 8M + 2S, in the radix-2^28
//                     field form of fp28.hpp): equal work per
//                     thread whatever the bucket sizes
//  6. k_bucket_seg    per segment of L = 4 buckets: merge each bucket's chunk
//                     partials, running sums -> A_t = sum_i (i+1) S, R_t = sum_i S
//  7. k_chan_sum      channels: sum_t A_t, and U_j = sum of the R_t whose index t has
//                     bit j set, so sum_b (b+1) S_b = sum_t A_t + L sum_j 2^j U_j --
//                     plain sums in any order, short dependency chains
//     (up to 4 sets of 2^t buckets instead: k_chan_tree merges the partials of 128
//     buckets per block and reduces them as a binary tree whose nodes carry their
//     per-bit sums, k_chan_top sums each set's block records into the same channels)
//  8. host tail       Horner over the channel bit positions (one host core runs a
//                     dependent chain of point operations ~20x faster than one wave).
//
// Two bucket layouts share the kernels:
//  * per-window sets (any bases): key = w * B + |d| - 1, set weight 2^(c w);
//  * one shared set (resident SRS with window-shifted copies 2^(cw) g_i built once):
//    key = |d| - 1, reference = w * N + offset + i, a single set of weight 1.
#include <pthread.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "g1.hpp"
#include "g1_28.hpp"
#include "host_ec.hpp"
#include "msm.hpp"
#include "srs.hpp"
#include "tune.hpp"

namespace mzk {

static constexpr uint32_t NEG_FLAG = 0x80000000u;
static constexpr int SORT_BLOCK = 256;
static constexpr int CHAN_BLOCK = 256;

// Signed window digit w of a canonical scalar (8 x u32 LE); `carry` threads the
// recoding from window to window: digit = bits + carry, digits above 2^(c-1)
// become digit - 2^c with a carry into the next window.
// The limb pair is picked by selects, not by indexing: an indexed read of a register
// array puts the scalar in scratch memory.
__device__ __forceinline__ int window_digit(const uint32_t s[8], int w, int c, uint32_t& carry) {
  const int bit = w * c;
  uint32_t bits = 0;
  if (bit < 256) {
    const int limb = bit >> 5, sh = bit & 31;
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      lo = limb == k ? s[k] : lo;
      hi = limb + 1 == k ? s[k] : hi;
    }
    const uint64_t two = ((uint64_t)hi << 32) | lo;
    bits = (uint32_t)(two >> sh) & ((1u << c) - 1u);
  }
  uint32_t d = bits + carry;
  const uint32_t half = 1u << (c - 1);
  if (d > half) {
    carry = 1;
    return (int)d - (int)(1u << c);
  }
  carry = 0;
  return (int)d;
}

// One MSM of a batch: its scalars, its first reference, where its canonical
// scalars go, and its first bucket set.
struct JobDesc {
  const uint32_t* coeffs;
  uint32_t n;
  uint32_t ref_offset;
  uint32_t canon_off;
  uint32_t set_base;
};

// A tile of one job's scalars, handled by one sort workgroup.  Jobs are sorted in groups
// of consecutive jobs (a group's buckets are [kbase, kbase + NBg)); the tile's histogram
// column is hist[hbase + bin * hstride] (hbase = D1 * the group's first tile + the tile's
// index in its group, hstride = the group's tile count).
struct TileDesc {
  uint32_t job, lo, hi;
  uint32_t hbase, hstride, kbase;
};

// A sort group: its first tile and tile count (its histogram block starts at D1 * tile0).
struct GroupDesc {
  uint32_t tile0, G;
};

struct SortArgs {
  const JobDesc* jobs;
  const TileDesc* tiles;
  const GroupDesc* groups;
  uint32_t G;            // tiles
  uint32_t NBg;          // buckets per sort group (the last group may have fewer)
  int c, W, per_window;
  uint32_t l2;           // low key bits sorted in k_sort_bins
  uint32_t D1;           // bins of the top key bits
  uint32_t ref_stride;   // table stride between windows (shared-set layout)
  // > 0: key low bits and reference share one word between scatter and bins, as
  // (low key << pack) | reference index | sign << (pack - 1); 0: two arrays
  uint32_t pack;
  // refs per accumulation chunk, written by k_sort_bins to off[NB + 1] once the digit count
  // M is known: ceil(M / kchunks) within [kmin, kmax]
  uint32_t kchunks, kmin, kmax;
};

// Bucket key of digit d of window w of a job.
__device__ __forceinline__ uint32_t bucket_key(const SortArgs& a, const JobDesc& j, int w, int d) {
  const uint32_t B = 1u << (a.c - 1);
  const uint32_t set = j.set_base + (a.per_window ? (uint32_t)w : 0u);
  return set * B + (uint32_t)(d > 0 ? d : -d) - 1u;
}

// Exclusive prefix sum of one value per thread over a SORT_BLOCK workgroup: wave-level
// shuffles, then the wave totals (wtot: SORT_BLOCK / 64 words of LDS).  *total gets the
// sum of all values.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wtot, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wtot[wid] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < SORT_BLOCK / 64; w++) {
    const uint32_t t = wtot[w];
    pre += w < wid ? t : 0u;
    tot += t;
  }
  __syncthreads();   // wtot may be rewritten by the next call
  *total = tot;
  return pre + x - v;
}

// 1. digits + per-tile histogram of key >> l2 (LDS atomics)
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_hist(SortArgs a, uint32_t* __restrict__ canon,
                                                          uint32_t* __restrict__ hist) {
  extern __shared__ uint32_t cnt[];   // D1 counters
  for (uint32_t i = threadIdx.x; i < a.D1; i += SORT_BLOCK) cnt[i] = 0;
  __syncthreads();
  const TileDesc tl = a.tiles[blockIdx.x];
  const JobDesc j = a.jobs[tl.job];
  for (uint32_t i = tl.lo + threadIdx.x; i < tl.hi; i += SORT_BLOCK) {
    fr_t s = from_mont(load<FrP>(j.coeffs + 8 * (size_t)i));
    store<FrP>(canon + 8 * ((size_t)j.canon_off + i), s);
    uint32_t carry = 0;
    for (int w = 0; w < a.W; w++) {
      int d = window_digit(s.v, w, a.c, carry);
      if (d != 0) atomicAdd(&cnt[(bucket_key(a, j, w, d) - tl.kbase) >> a.l2], 1u);
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < a.D1; b += SORT_BLOCK) hist[tl.hbase + (size_t)b * tl.hstride] = cnt[b];
}

// 3. scatter into bins: tmp_key[pos] = key & (2^l2 - 1), tmp_ref[pos] = reference | sign.
//    Each round stages the digits of SORT_BLOCK scalars x SCAT_WG windows in LDS in bin
//    order, then writes each bin's entries of the round as one contiguous run (one
//    scattered 4-byte store per digit would make this kernel bound by partial-line
//    HBM writes).
static constexpr int SCAT_WG = 16;   // 16 windows per round: 2^20 scatter 0.196 -> 0.187 ms against 8
static constexpr uint32_t SCAT_MAX_D1 = 8192;   // LDS: at most (3 D1 + 3 SCAT_SLOTS + SORT_BLOCK) words = 145 KiB
static constexpr int SCAT_SLOTS = SORT_BLOCK * SCAT_WG;
// buckets per sort group (batches of many bucket sets sort a group of jobs at a time:
// 96 shared sets of 2^15 in one group made k_sort_bins 10 ms of a 55 ms batch)
static constexpr size_t SORT_GROUP_BUCKETS = 1u << 18;

// LDS words of the scatter: cur, cnt, offs (D1 each), the partial sums, and the staging.
// Unpacked: rank, key and reference per slot (3 SCAT_SLOTS).  Packed (key low bits,
// reference and sign in one word): a 16-bit rank and bin per slot plus the packed word
// (2 SCAT_SLOTS), 39 KiB at 512 bins, so four workgroups share a CU instead of two.
__host__ __device__ constexpr size_t scatter_lds_words(uint32_t D1, bool packed) {
  return 3 * (size_t)D1 + (packed ? 2 : 3) * (size_t)SCAT_SLOTS + SORT_BLOCK;
}

template <bool PACKED>
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_scatter(const uint32_t* __restrict__ canon, SortArgs a,
                                                             const uint32_t* __restrict__ hoff,
                                                             uint32_t* __restrict__ tmp_key,
                                                             uint32_t* __restrict__ tmp_ref) {
  extern __shared__ uint32_t sm[];
  uint32_t* cur = sm;                   // D1: next output position of each bin
  uint32_t* cnt = cur + a.D1;           // D1: entries of the round per bin
  uint32_t* offs = cnt + a.D1;          // D1: LDS offset of each bin's entries in the round
  uint32_t* psum = offs + a.D1;         // SORT_BLOCK partial sums of the offset scan
  uint32_t* stage = psum + SORT_BLOCK;
  // unpacked: rank, key, reference (SCAT_SLOTS words each); packed: the packed word, then
  // the 16-bit ranks and bins (ranks < SCAT_SLOTS, bins < SCAT_MAX_D1)
  uint32_t* rank = stage;                                   // unpacked
  uint32_t* skey = stage + SCAT_SLOTS;                      // unpacked: staged keys, bin order
  uint32_t* sref = stage + 2 * SCAT_SLOTS;                  // unpacked: staged references
  uint32_t* spk = stage;                                    // packed: staged words, bin order
  uint16_t* rank16 = reinterpret_cast<uint16_t*>(stage + SCAT_SLOTS);
  uint16_t* sbin = rank16 + SCAT_SLOTS;                     // packed: bin of each staged word
  const uint32_t tid = threadIdx.x;
  const TileDesc tl = a.tiles[blockIdx.x];
  for (uint32_t b = tid; b < a.D1; b += SORT_BLOCK) cur[b] = hoff[tl.hbase + (size_t)b * tl.hstride];
  const uint32_t mask2 = (1u << a.l2) - 1u;
  const JobDesc j = a.jobs[tl.job];
  const uint32_t per_t = (a.D1 + SORT_BLOCK - 1) / SORT_BLOCK;   // bins per thread in the scan
  const uint32_t b0 = tid * per_t, b1 = b0 + per_t < a.D1 ? b0 + per_t : a.D1;
  for (uint32_t base = tl.lo; base < tl.hi; base += SORT_BLOCK) {
    const uint32_t i = base + tid;
    const bool live = i < tl.hi;
    fr_t s = fr_t::zero();
    if (live) s = load<FrP>(canon + 8 * ((size_t)j.canon_off + i));
    uint32_t carry0 = 0;   // recoding carry entering window w0
    for (int w0 = 0; w0 < a.W; w0 += SCAT_WG) {
      for (uint32_t b = tid; b < a.D1; b += SORT_BLOCK) cnt[b] = 0;
      __syncthreads();
      // pass 1: each digit's rank within its bin
      uint32_t carry = carry0;
      for (int wl = 0; wl < SCAT_WG; wl++) {
        const int w = w0 + wl;
        if (!live || w >= a.W) break;
        const int d = window_digit(s.v, w, a.c, carry);
        if (d != 0) {
          const uint32_t r = atomicAdd(&cnt[(bucket_key(a, j, w, d) - tl.kbase) >> a.l2], 1u);
          if constexpr (PACKED) rank16[tid * SCAT_WG + wl] = (uint16_t)r;
          else rank[tid * SCAT_WG + wl] = r;
        }
      }
      __syncthreads();
      // exclusive scan of the counts: per-thread ranges, then a block scan of the partials
      uint32_t part = 0;
      for (uint32_t b = b0; b < b1; b++) part += cnt[b];
      uint32_t all;
      uint32_t run = block_excl_scan(part, psum, &all);
      for (uint32_t b = b0; b < b1; b++) {
        offs[b] = run;
        run += cnt[b];
      }
      __syncthreads();
      // pass 2: place every digit at its bin's offset plus its rank
      carry = carry0;
      for (int wl = 0; wl < SCAT_WG; wl++) {
        const int w = w0 + wl;
        if (!live || w >= a.W) break;
        const int d = window_digit(s.v, w, a.c, carry);
        if (d != 0) {
          const uint32_t key = bucket_key(a, j, w, d) - tl.kbase;   // within the sort group
          uint32_t ref = j.ref_offset + i;
          if (!a.per_window) ref += (uint32_t)w * a.ref_stride;
          const uint32_t bin = key >> a.l2;
          if constexpr (PACKED) {
            const uint32_t p = offs[bin] + rank16[tid * SCAT_WG + wl];
            spk[p] = ((key & mask2) << a.pack) | ref | (d < 0 ? 1u << (a.pack - 1) : 0u);
            sbin[p] = (uint16_t)bin;
          } else {
            const uint32_t p = offs[bin] + rank[tid * SCAT_WG + wl];
            skey[p] = key;
            sref[p] = ref | (d < 0 ? NEG_FLAG : 0u);
          }
        }
      }
      carry0 = carry;
      __syncthreads();
      // write out in LDS order: consecutive entries of a bin go to consecutive positions
      const uint32_t staged = offs[a.D1 - 1] + cnt[a.D1 - 1];
      for (uint32_t e = tid; e < staged; e += SORT_BLOCK) {
        if constexpr (PACKED) {
          const uint32_t bin = sbin[e];
          tmp_key[cur[bin] + (e - offs[bin])] = spk[e];
        } else {
          const uint32_t key = skey[e];
          const uint32_t bin = key >> a.l2;
          const uint32_t pos = cur[bin] + (e - offs[bin]);
          tmp_key[pos] = key & mask2;
          tmp_ref[pos] = sref[e];
        }
      }
      __syncthreads();
      for (uint32_t b = tid; b < a.D1; b += SORT_BLOCK) cur[b] += cnt[b];
      __syncthreads();
    }
  }
}

static constexpr int BINS_UNROLL = 16;
static constexpr size_t BINS_LDS_CU = 160u << 10;   // LDS per CU

// 4. per top-bit bin of a sort group (workgroup g D1 + bin): counting sort on the low
//    bits; writes bucket offsets off[] (off[NB] = M) and the references in bucket order.
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_bins(const uint32_t* __restrict__ hoff, SortArgs a, uint32_t NB,
                                                          const uint32_t* __restrict__ tmp_key,
                                                          const uint32_t* __restrict__ tmp_ref,
                                                          uint32_t* __restrict__ refs, uint32_t* __restrict__ off,
                                                          uint32_t* __restrict__ bmax) {
  extern __shared__ uint32_t c2[];   // 2^l2 counters, then scan scratch
  const uint32_t g = blockIdx.x / a.D1, bin = blockIdx.x % a.D1;
  const GroupDesc gr = a.groups[g];
  const uint32_t R = 1u << a.l2;
  // the group's block of hoff starts at D1 tile0 (bin-major); its last bin ends where the
  // next group's block starts (hoff has D1 G + 1 entries)
  const size_t h0 = (size_t)a.D1 * gr.tile0;
  const uint32_t lo = hoff[h0 + (size_t)bin * gr.G];
  const uint32_t hi = hoff[h0 + (size_t)(bin + 1) * gr.G];
  for (uint32_t d = threadIdx.x; d < R; d += SORT_BLOCK) c2[d] = 0;
  __syncthreads();
  const uint32_t sh = a.pack;   // 0: keys alone in tmp_key (a shift by 0 keeps them)
  // BINS_UNROLL loads in flight per thread before their LDS atomics (one at a time, each
  // atomic waited for its load: the kernel was bound by load latency)
  for (uint32_t e = lo + threadIdx.x; e < hi; e += BINS_UNROLL * SORT_BLOCK) {
    uint32_t v[BINS_UNROLL];
#pragma unroll
    for (int u = 0; u < BINS_UNROLL; u++) {
      const uint32_t x = e + (uint32_t)u * SORT_BLOCK;
      v[u] = x < hi ? tmp_key[x] : 0u;
    }
#pragma unroll
    for (int u = 0; u < BINS_UNROLL; u++)
      if (e + (uint32_t)u * SORT_BLOCK < hi) atomicAdd(&c2[v[u] >> sh], 1u);
  }
  __syncthreads();
  {   // exclusive scan of the R counters from lo: per-thread ranges, then a block scan
    __shared__ uint32_t wtot[SORT_BLOCK / 64];
    const uint32_t per = (R + SORT_BLOCK - 1) / SORT_BLOCK;
    const uint32_t d0 = threadIdx.x * per, d1 = d0 + per < R ? d0 + per : R;
    uint32_t part = 0, big = 0;
    for (uint32_t d = d0; d < d1; d++) {
      part += c2[d];
      big = max(big, c2[d]);
    }
    uint32_t all;
    uint32_t run = lo + block_excl_scan(part, wtot, &all);
    for (uint32_t d = d0; d < d1; d++) {
      const uint32_t v = c2[d];
      c2[d] = run;
      run += v;
    }
    // the bin's largest bucket, for k_chunk_size
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) big = max(big, (uint32_t)__shfl_xor((int)big, d));
    if ((threadIdx.x & 63u) == 0) wtot[threadIdx.x >> 6] = big;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t m = 0;
      for (int w = 0; w < SORT_BLOCK / 64; w++) m = max(m, wtot[w]);
      bmax[blockIdx.x] = m;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) off[NB] = hoff[(size_t)a.D1 * a.G];
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < R; d += SORT_BLOCK) {
    const uint32_t lb = bin * R + d;   // bucket within the group
    const size_t b = (size_t)g * a.NBg + lb;
    if (lb < a.NBg && b < NB) off[b] = c2[d];
  }
  __syncthreads();
  if (sh) {
    const uint32_t idx_mask = (1u << (sh - 1)) - 1u;
    for (uint32_t e = lo + threadIdx.x; e < hi; e += BINS_UNROLL * SORT_BLOCK) {
      uint32_t v[BINS_UNROLL];
#pragma unroll
      for (int u = 0; u < BINS_UNROLL; u++) {
        const uint32_t x = e + (uint32_t)u * SORT_BLOCK;
        v[u] = x < hi ? tmp_key[x] : 0u;
      }
#pragma unroll
      for (int u = 0; u < BINS_UNROLL; u++) {
        if (e + (uint32_t)u * SORT_BLOCK < hi) {
          const uint32_t pos = atomicAdd(&c2[v[u] >> sh], 1u);
          refs[pos] = (v[u] & idx_mask) | (((v[u] >> (sh - 1)) & 1u) ? NEG_FLAG : 0u);
        }
      }
    }
  } else {
    for (uint32_t e = lo + threadIdx.x; e < hi; e += BINS_UNROLL * SORT_BLOCK) {
      uint32_t v[BINS_UNROLL], r[BINS_UNROLL];
#pragma unroll
      for (int u = 0; u < BINS_UNROLL; u++) {
        const uint32_t x = e + (uint32_t)u * SORT_BLOCK;
        v[u] = x < hi ? tmp_key[x] : 0u;
        r[u] = x < hi ? tmp_ref[x] : 0u;
      }
#pragma unroll
      for (int u = 0; u < BINS_UNROLL; u++) {
        if (e + (uint32_t)u * SORT_BLOCK < hi) {
          const uint32_t pos = atomicAdd(&c2[v[u]], 1u);
          refs[pos] = r[u];
        }
      }
    }
  }
}

// The refs per accumulation chunk, from the sort's results (off[NB] = M digits, bmax[] the
// bins' largest buckets): ceil(M / kchunks) spreads the digits over every thread (sparse
// scalars give few digits), but a bucket of D digits then spans D / K chunks whose
// partials k_long_merge sums by 64 lanes; on the few waves such inputs leave, the
// accumulation's K mixed additions (~9 us each on a lone wave) and the merge's D / (64 K)
// one-lane additions (~23 us each) balance near K = sqrt(D) / 4.  K = the larger of the
// two, within [kmin, kmax]; written to off[NB + 1], where every later kernel reads it.
__global__ void __launch_bounds__(256) k_chunk_size(uint32_t* __restrict__ off, uint32_t NB,
                                                    const uint32_t* __restrict__ bmax, uint32_t nbins, SortArgs a) {
  __shared__ uint32_t wm[4];
  uint32_t big = 0;
  for (uint32_t i = threadIdx.x; i < nbins; i += 256) big = max(big, bmax[i]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) big = max(big, (uint32_t)__shfl_xor((int)big, d));
  if ((threadIdx.x & 63u) == 0) wm[threadIdx.x >> 6] = big;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t D = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
    const uint32_t M = off[NB];
    const uint32_t spread = (uint32_t)(((uint64_t)M + a.kchunks - 1) / a.kchunks);
    const uint32_t skew = (uint32_t)ceilf(sqrtf((float)D) * 0.25f);
    off[NB + 1] = min(a.kmax, max(a.kmin, max(spread, skew)));
  }
}

// --------------------------------------------------------------------- scan
static constexpr int SCAN_BLOCK = 256;
static constexpr int SCAN_ITEMS = 8;   // elements per thread
static constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < SCAN_BLOCK; d <<= 1) {
    uint32_t x = t >= d ? sh[t - d] : 0u;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  total = sh[SCAN_BLOCK - 1];
  uint32_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

// Phase 1: per-tile exclusive scan, tile totals.
__global__ void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n,
                             uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t sh[SCAN_BLOCK];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < n ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t pre = block_exclusive_scan(sum, sh, total);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) out[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Phase 2: one block scans the tile totals (tiles <= SCAN_TILE), writes out[n] = total.
__global__ void k_scan_sums(uint32_t* __restrict__ tile_sums, int tiles, uint32_t* __restrict__ out, size_t n) {
  __shared__ uint32_t sh[SCAN_BLOCK];
  int base = threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < tiles ? tile_sums[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t pre = block_exclusive_scan(sum, sh, total);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < tiles) tile_sums[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) out[n] = total;
}

// Phase 3: add tile prefixes.
__global__ void k_scan_add(uint32_t* __restrict__ out, size_t n, const uint32_t* __restrict__ tile_sums) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] += tile_sums[i / SCAN_TILE];
}

// --------------------------------------------------------------------- accumulation
// Thread k adds refs[kK, (k+1)K) into running XYZZ sums, one per bucket run, and
// stores the sum for bucket b at partial slot k + b.  The running sums use the
// radix-2^28 field form (g1_28.hpp: one v_mad_u64_u32 per limb product, no carry
// instruction); partials are written in that form.  Slot k + b is written by exactly
// the chunks k that hold entries of bucket b, so no initialisation is needed.  The
// grid covers the largest possible M; M itself is read from off[NB].
//
// The point of reference r: blst affine (24 words, converted here) for caller bases, or
// a 128-byte record of a resident SRS's shifted tables (already in the radix-2^28 form,
// below 2p, identity flagged): one cache line per gathered point instead of 1.5 on
// average, and a lane that opens a bucket takes the record's coordinates as they are.
// The loads are issued one reference ahead (raw words held in registers) so that each
// gather's latency overlaps the previous mixed addition.
template <bool REC>
struct RawPoint {
  static constexpr int W = REC ? 29 : 24;
  uint32_t w[W];
};

template <bool REC>
__device__ __forceinline__ void issue_point(const uint32_t* __restrict__ bases, uint32_t r, RawPoint<REC>& o) {
  const size_t idx = (size_t)(r & ~NEG_FLAG);
  if constexpr (REC) {
    const uint4* q = reinterpret_cast<const uint4*>(bases + MZK_SRS_REC_WORDS * idx);
#pragma unroll
    for (int i = 0; i < 7; i++) {
      const uint4 a = q[i];
      o.w[4 * i] = a.x; o.w[4 * i + 1] = a.y; o.w[4 * i + 2] = a.z; o.w[4 * i + 3] = a.w;
    }
    o.w[28] = bases[MZK_SRS_REC_WORDS * idx + 28];
  } else {
    const uint4* q = reinterpret_cast<const uint4*>(bases + 24 * idx);
#pragma unroll
    for (int i = 0; i < 6; i++) {
      const uint4 a = q[i];
      o.w[4 * i] = a.x; o.w[4 * i + 1] = a.y; o.w[4 * i + 2] = a.z; o.w[4 * i + 3] = a.w;
    }
  }
}

// raw words -> (x, y) in the radix-2^28 form; false for the identity (adds nothing)
template <bool REC>
__device__ __forceinline__ bool decode_point(const RawPoint<REC>& o, f28& x, f28& y) {
  if constexpr (REC) {
#pragma unroll
    for (int i = 0; i < 14; i++) {
      x.l[i] = o.w[i];
      y.l[i] = o.w[14 + i];
    }
    return o.w[28] == 0;
  }
  uint32_t any = 0;
#pragma unroll
  for (int i = 0; i < 24; i++) any |= o.w[i];
  x = f28_from_std(o.w);
  y = f28_from_std(o.w + 12);
  return any != 0;
}

template <bool REC>
__global__ void __launch_bounds__(64) k_msm_accum(const uint32_t* __restrict__ refs,
                                                  const uint32_t* __restrict__ off, uint32_t NB,
                                                  const uint32_t* __restrict__ bases,
                                                  uint32_t* __restrict__ partials) {
  const uint32_t M = off[NB], K = off[NB + 1];
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t e0_64 = (uint64_t)k * K;
  if (e0_64 >= M) return;
  const uint32_t e0 = (uint32_t)e0_64;
  const uint32_t e1 = (uint32_t)std::min<uint64_t>(e0_64 + K, M);
  // bucket holding entry e0: the last b with off[b] <= e0
  uint32_t lo = 0, hi = NB;   // off[lo] <= e0 < off[hi] (off[NB] = M)
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= e0) lo = mid; else hi = mid;
  }
  uint32_t cb = lo;
  uint32_t next = off[cb + 1];
  // the boundary after next, loaded ahead: a lane that closes a bucket reads it from a
  // register instead of waiting for a dependent load (off[NB] = M ends every search)
  uint32_t next2 = off[std::min(cb + 2, NB)];
  g1x28 acc = g1x28_inf();
  uint32_t r_next = refs[e0];
  RawPoint<REC> raw;
  issue_point<REC>(bases, r_next, raw);
  for (uint32_t e = e0; e < e1; e++) {
    const uint32_t r = r_next;
    f28 x, y;
    const bool live = decode_point<REC>(raw, x, y);
    if (e + 1 < e1) {   // the next gather, in flight during this addition
      r_next = refs[e + 1];
      issue_point<REC>(bases, r_next, raw);
    }
    if (e >= next) {
      store_x28_raw(partials + MZK_X28_WORDS * (size_t)(k + cb), acc);
      acc = g1x28_inf();
      cb++;
      next = next2;
      uint32_t steps = 0;
      while (next <= e && ++steps < 16) {   // a few empty buckets (small sets)
        cb++;
        next = off[cb + 1];
      }
      if (__builtin_expect(next <= e, 0)) {   // a long empty run (skewed scalars): bisection
        uint32_t lo = cb, hi = NB;           // off[lo] <= e < off[hi] (off[NB] = M > e)
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (off[mid] <= e) lo = mid; else hi = mid;
        }
        cb = lo;
        next = off[cb + 1];
      }
      next2 = off[std::min(cb + 2, NB)];
    }
    if (live) acc = g1x28_madd<REC>(acc, x, y, (r & NEG_FLAG) != 0);
  }
  store_x28_raw(partials + MZK_X28_WORDS * (size_t)(k + cb), acc);
}

// A bucket's chunk partials are the slots k + b for its chunks k0..k1 (a chunk's thread
// writes the partial of each bucket it touches there).  Skewed scalars (a boolean or
// small-valued column, repeated values) put most digits in a few buckets, whose partials
// a reduction kernel would merge as one dependent chain of additions (a boolean column
// of 2^18 terms: ~4000 partials, ~60 ms).  A bucket spanning more than LONG_CHUNKS chunks
// is therefore merged by k_long_merge first, by a whole wave, into its first slot, and the
// reductions read only that slot (bucket_chunks).
static constexpr uint32_t LONG_CHUNKS = 64;

__device__ __forceinline__ bool bucket_chunks(const uint32_t* __restrict__ off, uint32_t b, uint32_t K, uint32_t& k0,
                                              uint32_t& k1) {
  const uint32_t lo = off[b], hi = off[b + 1];
  if (hi <= lo) return false;
  k0 = lo / K;
  k1 = (hi - 1) / K;
  if (k1 - k0 >= LONG_CHUNKS) k1 = k0;   // merged by k_long_merge
  return true;
}

// A butterfly step of a wave-wide point sum: lane l gets lane (l ^ d)'s point.
__device__ __forceinline__ g1x28 shfl_xor_x28(const g1x28& a, int d) {
  g1x28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) {
    r.X.l[i] = (uint32_t)__shfl_xor((int)a.X.l[i], d);
    r.Y.l[i] = (uint32_t)__shfl_xor((int)a.Y.l[i], d);
    r.ZZ.l[i] = (uint32_t)__shfl_xor((int)a.ZZ.l[i], d);
    r.ZZZ.l[i] = (uint32_t)__shfl_xor((int)a.ZZZ.l[i], d);
  }
  return r;
}

// A workgroup of four waves per 64 buckets: every wave reads the 64 buckets' spans (lane
// i: bucket 64 blockIdx + i).  Four or more long buckets: each wave takes every fourth
// one, its lanes sum the bucket's partials strided by 64 (each lane's first slot is its
// own) and a butterfly over the lanes by shuffles (no barrier: each wave runs alone) leaves
// the sum in the bucket's first slot.  Fewer (one hot bucket is the common skewed case):
// the four waves share each bucket, lanes strided by 256, and add their four sums through
// LDS.  With no long bucket a workgroup reads 65 offsets and leaves.
__global__ void __launch_bounds__(256) k_long_merge(const uint32_t* __restrict__ off, uint32_t NB,
                                                    uint32_t* __restrict__ partials) {
  __shared__ uint32_t lds[4 * MZK_X28_WORDS];
  const uint32_t K = off[NB + 1];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint32_t b_lane = blockIdx.x * 64 + lane;
  uint32_t k0 = 0, n = 0;
  if (b_lane < NB) {
    const uint32_t lo = off[b_lane], hi = off[b_lane + 1];
    if (hi > lo) {
      k0 = lo / K;
      n = (hi - 1) / K - k0 + 1;
    }
  }
  uint64_t todo = __ballot(n > LONG_CHUNKS);   // the same in every wave of the workgroup
  const bool share = __popcll(todo) < 4;
  for (uint32_t r = 0; todo; r++) {
    const uint32_t i = (uint32_t)__ffsll((long long)todo) - 1;
    todo &= todo - 1;
    if (!share && (r & 3u) != wv) continue;
    const uint32_t first = (uint32_t)__shfl((int)k0, (int)i) + blockIdx.x * 64 + i;   // slot k0 + b
    const uint32_t cnt = (uint32_t)__shfl((int)n, (int)i);                             // > 64 slots
    const uint32_t me = share ? threadIdx.x : lane, stride = share ? 256u : 64u;
    g1x28 x = g1x28_inf();
    for (uint32_t k = me; k < cnt; k += stride) x = g1x28_add(x, load_x28(partials + MZK_X28_WORDS * (size_t)(first + k)));
    for (int d = 32; d >= 1; d >>= 1) x = g1x28_add(x, shfl_xor_x28(x, d));
    if (share) {   // uniform: every wave of the workgroup is here
      if (lane == 0) store_x28_raw(lds + MZK_X28_WORDS * wv, x);
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) x = g1x28_add(x, load_x28(lds + MZK_X28_WORDS * w));
        store_x28_raw(partials + MZK_X28_WORDS * (size_t)first, x);
      }
      __syncthreads();   // the LDS is reused by the next bucket
    } else if (lane == 0) {
      store_x28_raw(partials + MZK_X28_WORDS * (size_t)first, x);
    }
  }
}

// Segment t of L buckets of one set (b = t L + i): S_b = sum of the chunk partials
// of bucket b, then running sums from the top: R_t = sum_i S_b and
// A_t = sum_i (i + 1) S_b.  The set's weighted sum is then
//   sum_b (b + 1) S_b = sum_t A_t + L sum_t t R_t.
__global__ void __launch_bounds__(64) k_bucket_seg(const uint32_t* __restrict__ off, uint32_t NB,
                                                   uint32_t L, const uint32_t* __restrict__ partials,
                                                   uint32_t* __restrict__ A, uint32_t* __restrict__ R) {
  const uint32_t K = off[NB + 1];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if ((size_t)t * L >= NB) return;
  g1x28 run = g1x28_inf(), acc = g1x28_inf();
  for (uint32_t i = L; i-- > 0;) {
    const uint32_t b = t * L + i;
    uint32_t k0, k1;
    if (bucket_chunks(off, b, K, k0, k1))
      for (uint32_t k = k0; k <= k1; k++) run = g1x28_add(run, load_x28(partials + MZK_X28_WORDS * (size_t)(k + b)));
    acc = g1x28_add(acc, run);
  }
  store_x28_raw(A + MZK_X28_WORDS * (size_t)t, acc);
  store_x28_raw(R + MZK_X28_WORDS * (size_t)t, run);
}

// Small bucket sets (per-window sets of B <= 256 buckets, B a power of two): one
// workgroup reduces 256 / B whole sets.  Thread i of a set merges bucket i's chunk
// partials into S_i, a suffix scan in LDS gives U_i = sum_(i' >= i) S_i', and a tree sum
// gives sum_i U_i = sum_i (i + 1) S_i, the set's weighted bucket sum, written in the
// blst layout for the host's window combination.
__global__ void __launch_bounds__(256) k_set_reduce(const uint32_t* __restrict__ off, uint32_t NB,
                                                    uint32_t B, uint32_t nsets,
                                                    const uint32_t* __restrict__ partials,
                                                    uint32_t* __restrict__ out) {
  const uint32_t K = off[NB + 1];
  extern __shared__ uint32_t lds[];
  const uint32_t t = threadIdx.x;
  const uint32_t i = t % B;
  const uint32_t set = blockIdx.x * (blockDim.x / B) + t / B;
  const bool live = set < nsets;
  g1x28 x = g1x28_inf();
  if (live) {
    const uint32_t b = set * B + i;
    uint32_t k0, k1;
    if (bucket_chunks(off, b, K, k0, k1))
      for (uint32_t k = k0; k <= k1; k++) x = g1x28_add(x, load_x28(partials + MZK_X28_WORDS * (size_t)(k + b)));
  }
  uint32_t* mine = lds + MZK_X28_WORDS * t;
  store_x28_raw(mine, x);
  __syncthreads();
  for (uint32_t d = 1; d < B; d <<= 1) {   // suffix sums within the set's B threads
    g1x28 nb = g1x28_inf();
    if (i + d < B) nb = load_x28(lds + MZK_X28_WORDS * (t + d));
    __syncthreads();
    if (i + d < B) {
      x = g1x28_add(x, nb);
      store_x28_raw(mine, x);
    }
    __syncthreads();
  }
  for (uint32_t d = B / 2; d >= 1; d >>= 1) {   // tree sum of the suffix sums
    if (i < d) {
      x = g1x28_add(x, load_x28(lds + MZK_X28_WORDS * (t + d)));
      store_x28_raw(mine, x);
    }
    __syncthreads();
  }
  if (live && i == 0) {
    uint32_t o[48];
    g1x28_to_std(x, o);
    uint4* qv = reinterpret_cast<uint4*>(out + 48 * (size_t)set);
#pragma unroll
    for (int w = 0; w < 12; w++) qv[w] = make_uint4(o[4 * w], o[4 * w + 1], o[4 * w + 2], o[4 * w + 3]);
  }
}

// k_set_reduce with a quad of lanes per bucket (g1x28_add_quad): for batches whose
// sets x buckets are too few to put a wave on every SIMD, where each level's addition
// latency is the kernel's time.  128 quads per workgroup, 128 / B whole sets (B <= 128).
static constexpr uint32_t QUAD_SETS_MAX_BUCKETS = 16384;   // sets x B at most: 1 wave per SIMD as quads

__global__ void __launch_bounds__(512) k_set_reduce_quad(const uint32_t* __restrict__ off, uint32_t NB,
                                                         uint32_t B, uint32_t nsets,
                                                         const uint32_t* __restrict__ partials,
                                                         uint32_t* __restrict__ out) {
  const uint32_t K = off[NB + 1];
  extern __shared__ uint32_t lds[];   // two buffers of 128 points: the scan reads one, writes the other
  const uint32_t t = threadIdx.x, qd = t >> 2, ql = t & 3u;
  const uint32_t i = qd % B;
  const uint32_t set = blockIdx.x * (128u / B) + qd / B;
  const bool live = set < nsets;
  g1x28 x = g1x28_inf();
  if (live) {
    const uint32_t b = set * B + i;
    uint32_t k0, k1;
    if (bucket_chunks(off, b, K, k0, k1)) {   // the first partial loaded, the rest added (an addition to the identity costs a full one here)
      x = load_x28(partials + MZK_X28_WORDS * (size_t)(k0 + b));
      for (uint32_t k = k0 + 1; k <= k1; k++) x = g1x28_add_quad(x, partials + MZK_X28_WORDS * (size_t)(k + b));
    }
  }
  uint32_t* cur = lds;
  uint32_t* nxt = lds + 128 * MZK_X28_WORDS;
  if (ql == 0) store_x28_raw(cur + MZK_X28_WORDS * qd, x);
  __syncthreads();
  for (uint32_t d = 1; d < B; d <<= 1) {   // suffix sums within the set's B quads
    if (i + d < B) x = g1x28_add_quad(x, cur + MZK_X28_WORDS * (qd + d));
    if (ql == 0) store_x28_raw(nxt + MZK_X28_WORDS * qd, x);
    __syncthreads();
    uint32_t* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  for (uint32_t d = B / 2; d >= 1; d >>= 1) {   // tree sum of the suffix sums: reads slots >= d, writes < d
    if (i < d) {
      x = g1x28_add_quad(x, cur + MZK_X28_WORDS * (qd + d));
      if (ql == 0) store_x28_raw(cur + MZK_X28_WORDS * qd, x);
    }
    __syncthreads();
  }
  if (live && i == 0 && ql == 0) {
    uint32_t o[48];
    g1x28_to_std(x, o);
    uint4* qv = reinterpret_cast<uint4*>(out + 48 * (size_t)set);
#pragma unroll
    for (int w = 0; w < 12; w++) qv[w] = make_uint4(o[4 * w], o[4 * w + 1], o[4 * w + 2], o[4 * w + 3]);
  }
}

// k_bucket_seg for segments of one bucket (a single set): a quad per bucket merges its
// chunk partials with quad-lane additions; A_t = R_t = S_b.  2^15 quads put two waves on
// each SIMD where one lane per bucket left half of them idle.
__global__ void __launch_bounds__(256) k_bucket_merge_quad(const uint32_t* __restrict__ off, uint32_t NB,
                                                           const uint32_t* __restrict__ partials,
                                                           uint32_t* __restrict__ A, uint32_t* __restrict__ R) {
  const uint32_t K = off[NB + 1];
  const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
  if (b >= NB) return;   // whole quads leave together (NB quads, 4 lanes each)
  g1x28 x = g1x28_inf();
  uint32_t k0, k1;
  if (bucket_chunks(off, b, K, k0, k1)) {
    x = load_x28(partials + MZK_X28_WORDS * (size_t)(k0 + b));
    for (uint32_t k = k0 + 1; k <= k1; k++) x = g1x28_add_quad(x, partials + MZK_X28_WORDS * (size_t)(k + b));
  }
  if ((threadIdx.x & 3u) == 0) {
    store_x28_raw(A + MZK_X28_WORDS * (size_t)b, x);
    store_x28_raw(R + MZK_X28_WORDS * (size_t)b, x);
  }
}

// Block-wide sum of one point per thread (LDS tree); thread 0 returns it.
__device__ __forceinline__ g1x28 block_sum(g1x28 x, uint32_t* lds) {
  const uint32_t t = threadIdx.x;
  store_x28_raw(lds + MZK_X28_WORDS * t, x);
  __syncthreads();
  for (uint32_t d = blockDim.x / 2; d >= 1; d >>= 1) {
    if (t < d) {
      x = g1x28_add(x, load_x28(lds + MZK_X28_WORDS * (t + d)));
      store_x28_raw(lds + MZK_X28_WORDS * t, x);
    }
    __syncthreads();
  }
  return x;
}

// 6'+7'. A single set of T = 2^t buckets (segments of one bucket, A_t = R_t = S_t) by
// blocks of CT_LEAVES = 2^CT_LEVELS buckets.  The channels ask for sum_t S_t and, for
// each bit j, U_j = sum of the S_t whose index t has bit j set.  Inside a block, a binary
// tree over its buckets gives every node its sum P and, for each bit i below the node's
// level, A_i = the sum of its buckets with bit i set: a parent's A_i is the sum of its
// children's A_i, and its A_j (j its children's level) is its right child's P.  The root
// carries the block's P and A_0..A_6 with (CT_LEAVES - 1) + (CT_LEAVES - 1 - CT_LEVELS)
// additions, where the channel loop spends CT_LEAVES (1 + CT_LEVELS / 2); every level is
// one addition step, the (CT_LEAVES / 2 >> j) (j + 1) additions of level j on separate
// quads.  k_chan_top then sums U_j = sum over blocks of A_j for j < 7, and for the
// higher bits the block sums P_k of the blocks k with bit j - 7 set.
static constexpr uint32_t CT_LEVELS = 7;
static constexpr uint32_t CT_LEAVES = 1u << CT_LEVELS;   // buckets per block, a quad each
static constexpr uint32_t CT_REC = CT_LEVELS + 1;         // points per block record: P, A_0..A_6
static constexpr uint32_t CT_TOP_QUADS = 128;
static constexpr size_t TREE_SETS = 4;   // most sets reduced by the block tree

// Slot of A_i of node m at the current level: A_i exists for 64 >> i nodes when created.
__device__ __forceinline__ uint32_t ct_acc_slot(uint32_t i, uint32_t m) {
  return CT_LEAVES + (CT_LEAVES - (CT_LEAVES >> i)) + m;
}

// Block k: buckets [k CT_LEAVES, (k + 1) CT_LEAVES), each the merge of its chunk partials
// (as k_bucket_merge_quad), then the tree; writes out[CT_REC k + 0] = P, out[CT_REC k + 1 + i]
// = A_i (radix-2^28 XYZZ, MZK_X28_WORDS words each).  LDS: P of the current level (in place,
// CT_LEAVES slots) and the A_i (CT_LEAVES - 1 slots), (2 CT_LEAVES - 1) points.
__global__ void __launch_bounds__(4 * CT_LEAVES) k_chan_tree(const uint32_t* __restrict__ off, uint32_t NB,
                                                            const uint32_t* __restrict__ partials,
                                                            uint32_t* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  const uint32_t t = threadIdx.x, qd = t >> 2, ql = t & 3u;
  const uint32_t K = off[NB + 1];
  {
    const uint32_t b = blockIdx.x * CT_LEAVES + qd;
    g1x28 x = g1x28_inf();
    uint32_t k0, k1;
    if (bucket_chunks(off, b, K, k0, k1)) {
      x = load_x28(partials + MZK_X28_WORDS * (size_t)(k0 + b));
      for (uint32_t k = k0 + 1; k <= k1; k++) x = g1x28_add_quad(x, partials + MZK_X28_WORDS * (size_t)(k + b));
    }
    if (ql == 0) store_x28_raw(lds + MZK_X28_WORDS * qd, x);
  }
  __syncthreads();
  for (uint32_t j = 0; j < CT_LEVELS; j++) {
    const uint32_t parents = (CT_LEAVES / 2) >> j, per = j + 1;
    // A_j of parent m = P of its right child 2m + 1 (the A_j slots are not read this level)
    for (uint32_t w = t; w < parents * MZK_X28_WORDS; w += blockDim.x) {
      const uint32_t m = w / MZK_X28_WORDS, e = w % MZK_X28_WORDS;
      lds[MZK_X28_WORDS * ct_acc_slot(j, m) + e] = lds[MZK_X28_WORDS * (2 * m + 1) + e];
    }
    // task (m, u): u = 0 adds the children's P, u = i + 1 their A_i (i < j)
    const bool act = qd < parents * per;
    const uint32_t m = qd / per, u = qd % per;
    const uint32_t s0 = u == 0 ? 2 * m : ct_acc_slot(u - 1, 2 * m);
    g1x28 x = g1x28_inf();
    if (act) x = g1x28_add_quad(load_x28(lds + MZK_X28_WORDS * s0), lds + MZK_X28_WORDS * (s0 + 1));
    __syncthreads();   // every read of this level before any write
    if (act && ql == 0) store_x28_raw(lds + MZK_X28_WORDS * (u == 0 ? m : ct_acc_slot(u - 1, m)), x);
    __syncthreads();
  }
  // the root: P in slot 0, A_i in ct_acc_slot(i, 0)
  for (uint32_t w = t; w < CT_REC * MZK_X28_WORDS; w += blockDim.x) {
    const uint32_t r = w / MZK_X28_WORDS, e = w % MZK_X28_WORDS;
    const uint32_t s = r == 0 ? 0 : ct_acc_slot(r - 1, 0);
    out[MZK_X28_WORDS * ((size_t)blockIdx.x * CT_REC + r) + e] = lds[MZK_X28_WORDS * s + e];
  }
}

// Channel h (blockIdx.x) of set s (blockIdx.y), from the set's nblk block records
// (nblk = 2^kbits, set s's at in[CT_REC nblk s ...]): h = 0 sums every P_k; h = 1 + i
// (i < CT_LEVELS) sums every A_i; h = 1 + CT_LEVELS + j sums the P_k whose k has bit j
// set.  A quad per CT_TOP_QUADS-th record (the first of its terms loaded, not added to the
// identity: each addition step is a serial step of the kernel), then a tree over the quads
// in LDS; writes the blst-layout XYZZ (canonical 12 x 32-bit limbs) of the channel to
// out[48 (s gridDim.x + h)].
__global__ void __launch_bounds__(4 * CT_TOP_QUADS) k_chan_top(const uint32_t* __restrict__ in, uint32_t nblk,
                                                              uint32_t* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  const uint32_t qd = threadIdx.x >> 2, ql = threadIdx.x & 3u;
  const uint32_t h = blockIdx.x;
  in += (size_t)MZK_X28_WORDS * CT_REC * nblk * blockIdx.y;
  out += (size_t)48 * gridDim.x * blockIdx.y;
  const uint32_t slot = (h >= 1 && h <= CT_LEVELS) ? h : 0;
  const uint32_t bit = h > CT_LEVELS ? h - 1 - CT_LEVELS : 0;
  const bool all = h <= CT_LEVELS;
  uint32_t k = qd;
  while (k < nblk && !all && !((k >> bit) & 1u)) k += CT_TOP_QUADS;
  g1x28 x = g1x28_inf();
  if (k < nblk) {
    x = load_x28(in + MZK_X28_WORDS * ((size_t)k * CT_REC + slot));
    for (k += CT_TOP_QUADS; k < nblk; k += CT_TOP_QUADS)
      if (all || ((k >> bit) & 1u)) x = g1x28_add_quad(x, in + MZK_X28_WORDS * ((size_t)k * CT_REC + slot));
  }
  if (ql == 0) store_x28_raw(lds + MZK_X28_WORDS * qd, x);
  __syncthreads();
  for (uint32_t d = CT_TOP_QUADS / 2; d >= 1; d >>= 1) {   // reads slots >= d, writes < d
    if (qd < d) {
      x = g1x28_add_quad(x, lds + MZK_X28_WORDS * (qd + d));
      if (ql == 0) store_x28_raw(lds + MZK_X28_WORDS * qd, x);
    }
    __syncthreads();
  }
  if (qd == 0 && ql == 0) {
    uint32_t o[48];
    g1x28_to_std(x, o);
    uint4* qv = reinterpret_cast<uint4*>(out + 48 * (size_t)h);
#pragma unroll
    for (int w = 0; w < 12; w++) qv[w] = make_uint4(o[4 * w], o[4 * w + 1], o[4 * w + 2], o[4 * w + 3]);
  }
}

// The same tree for sets of segments (L > 1, T = 2^t segments per set, t >= 7): leaves are
// the segments' (R_t, A_t), ST_LEAVES per block; a node also carries Asum, the sum of its
// A_t (channel 0).  Level j has (ST_LEAVES / 2 >> j)(j + 2) additions, a quad each.
// The records (P, Asum, A_0..A_5) of a set's T / ST_LEAVES blocks go to k_seg_top.
static constexpr uint32_t ST_LEVELS = 6;
static constexpr uint32_t ST_LEAVES = 1u << ST_LEVELS;
static constexpr uint32_t ST_REC = ST_LEVELS + 2;
static constexpr uint32_t ST_TOP_QUADS = 128;   // most quads of a k_seg_top block

__device__ __forceinline__ uint32_t st_acc_slot(uint32_t i, uint32_t m) {
  return 2 * ST_LEAVES + (ST_LEAVES - (ST_LEAVES >> i)) + m;
}

// LDS: P (ST_LEAVES slots, in place), Asum (ST_LEAVES), the A_i (ST_LEAVES - 1).
__global__ void __launch_bounds__(4 * ST_LEAVES) k_seg_tree(const uint32_t* __restrict__ A,
                                                           const uint32_t* __restrict__ R,
                                                           uint32_t* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  const uint32_t t = threadIdx.x, qd = t >> 2, ql = t & 3u;
  const size_t leaf0 = (size_t)blockIdx.x * ST_LEAVES;
  constexpr uint32_t V = MZK_X28_WORDS / 4;   // uint4 per point
  for (uint32_t w = t; w < ST_LEAVES * V; w += blockDim.x) {
    const uint32_t leaf = w / V, e = 4 * (w % V);
    *reinterpret_cast<uint4*>(lds + MZK_X28_WORDS * leaf + e) =
        *reinterpret_cast<const uint4*>(R + MZK_X28_WORDS * (leaf0 + leaf) + e);
    *reinterpret_cast<uint4*>(lds + MZK_X28_WORDS * (ST_LEAVES + leaf) + e) =
        *reinterpret_cast<const uint4*>(A + MZK_X28_WORDS * (leaf0 + leaf) + e);
  }
  __syncthreads();
  for (uint32_t j = 0; j < ST_LEVELS; j++) {
    const uint32_t parents = (ST_LEAVES / 2) >> j, per = j + 2;
    for (uint32_t w = t; w < parents * MZK_X28_WORDS; w += blockDim.x) {   // A_j of parent m = P of child 2m + 1
      const uint32_t m = w / MZK_X28_WORDS, e = w % MZK_X28_WORDS;
      lds[MZK_X28_WORDS * st_acc_slot(j, m) + e] = lds[MZK_X28_WORDS * (2 * m + 1) + e];
    }
    // task (m, u) of quad qd: u = 0 adds the children's P, u = 1 their Asum, u = i + 2
    // their A_i (i < j); the A_j slots written above are not read this level
    const bool act = qd < parents * per;
    const uint32_t m = qd / per, u = qd % per;
    const uint32_t s0 = u == 0 ? 2 * m : u == 1 ? ST_LEAVES + 2 * m : st_acc_slot(u - 2, 2 * m);
    g1x28 x = g1x28_inf();
    if (act) x = g1x28_add_quad(load_x28(lds + MZK_X28_WORDS * s0), lds + MZK_X28_WORDS * (s0 + 1));
    __syncthreads();
    if (act && ql == 0)
      store_x28_raw(lds + MZK_X28_WORDS * (u == 0 ? m : u == 1 ? ST_LEAVES + m : st_acc_slot(u - 2, m)), x);
    __syncthreads();
  }
  for (uint32_t w = t; w < ST_REC * MZK_X28_WORDS; w += blockDim.x) {
    const uint32_t r = w / MZK_X28_WORDS, e = w % MZK_X28_WORDS;
    const uint32_t s = r == 0 ? 0 : r == 1 ? ST_LEAVES : st_acc_slot(r - 2, 0);
    out[MZK_X28_WORDS * ((size_t)blockIdx.x * ST_REC + r) + e] = lds[MZK_X28_WORDS * s + e];
  }
}

// Channel h of set s from its nrec block records (block (s nch + h)): h = 0 sums the Asum,
// h = 1 + i (i < ST_LEVELS) the A_i, h = 1 + ST_LEVELS + j the P of the records whose
// index has bit j set.  nq = blockDim.x / 4 quads (a power of two): quad q takes records
// q, q + nq, ... (the first of them loaded, the rest added), then a tree over the quads in
// LDS; writes the blst-layout XYZZ of the channel at out[48 (s nch + h)].
__global__ void __launch_bounds__(4 * ST_TOP_QUADS) k_seg_top(const uint32_t* __restrict__ in, uint32_t nrec,
                                                             uint32_t nch, uint32_t* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  const uint32_t qd = threadIdx.x >> 2, ql = threadIdx.x & 3u, nq = blockDim.x >> 2;
  const uint32_t s = blockIdx.x / nch, h = blockIdx.x % nch;
  const uint32_t slot = h == 0 ? 1 : h <= ST_LEVELS ? h + 1 : 0;
  const bool all = h <= ST_LEVELS;
  const uint32_t bit = all ? 0 : h - 1 - ST_LEVELS;
  const uint32_t* rec = in + MZK_X28_WORDS * ((size_t)s * nrec * ST_REC + slot);
  uint32_t k = qd;
  while (k < nrec && !all && !((k >> bit) & 1u)) k += nq;
  g1x28 x = g1x28_inf();
  if (k < nrec) {
    x = load_x28(rec + MZK_X28_WORDS * (size_t)k * ST_REC);
    for (k += nq; k < nrec; k += nq)
      if (all || ((k >> bit) & 1u)) x = g1x28_add_quad(x, rec + MZK_X28_WORDS * (size_t)k * ST_REC);
  }
  if (ql == 0) store_x28_raw(lds + MZK_X28_WORDS * qd, x);
  __syncthreads();
  for (uint32_t d = nq / 2; d >= 1; d >>= 1) {   // reads slots >= d, writes < d
    if (qd < d) {
      x = g1x28_add_quad(x, lds + MZK_X28_WORDS * (qd + d));
      if (ql == 0) store_x28_raw(lds + MZK_X28_WORDS * qd, x);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint32_t o[48];
    g1x28_to_std(x, o);
    uint4* qv = reinterpret_cast<uint4*>(out + 48 * (size_t)blockIdx.x);
#pragma unroll
    for (int w = 0; w < 12; w++) qv[w] = make_uint4(o[4 * w], o[4 * w + 1], o[4 * w + 2], o[4 * w + 3]);
  }
}

// 7. channels over the T segments of each set.  Channel 0 sums all A_t; channel
// 1 + j sums the R_t whose index t has bit j set, so sum_t t R_t = sum_j 2^j U_j.
// Block (set s, channel h, part q) writes out[(s * nch + h) * parts + q].
__global__ void __launch_bounds__(CHAN_BLOCK) k_chan_sum(const uint32_t* __restrict__ A,
                                                         const uint32_t* __restrict__ R, uint32_t T, uint32_t nch,
                                                         uint32_t parts, uint32_t* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  const uint32_t q = blockIdx.x % parts;
  const uint32_t sh = blockIdx.x / parts;
  const uint32_t h = sh % nch, s = sh / nch;
  const uint32_t per = T / parts;
  const uint32_t* src = h == 0 ? A : R;
  g1x28 acc = g1x28_inf();
  for (uint32_t t = q * per + threadIdx.x; t < (q + 1) * per; t += CHAN_BLOCK)
    if (h == 0 || ((t >> (h - 1)) & 1u)) acc = g1x28_add(acc, load_x28(src + MZK_X28_WORDS * ((size_t)s * T + t)));
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) store_x28_raw(out + MZK_X28_WORDS * (size_t)blockIdx.x, acc);
}

// Sums `len` consecutive points per segment; one block per segment.  Writes the
// blst-layout XYZZ (canonical 12 x 32-bit limbs) the host tail reads.
__global__ void __launch_bounds__(CHAN_BLOCK) k_seg_sum(const uint32_t* __restrict__ in, uint32_t len,
                                                        uint32_t* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  g1x28 acc = g1x28_inf();
  for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
    acc = g1x28_add(acc, load_x28(in + MZK_X28_WORDS * ((size_t)blockIdx.x * len + i)));
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) {
    uint32_t o[48];
    g1x28_to_std(acc, o);
    uint4* qv = reinterpret_cast<uint4*>(out + 48 * (size_t)blockIdx.x);
#pragma unroll
    for (int i = 0; i < 12; i++) qv[i] = make_uint4(o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]);
  }
}

// --------------------------------------------------------------------- host side
static host::Xyzz to_host(const uint32_t* p) {
  host::Xyzz r;
  memcpy(r.X.l, p, 48);
  memcpy(r.Y.l, p + 12, 48);
  memcpy(r.ZZ.l, p + 24, 48);
  memcpy(r.ZZZ.l, p + 36, 48);
  return r;
}

static int ceil_log2(uint64_t x) {
  int l = 0;
  while ((1ull << l) < x) l++;
  return l;
}

// Window bits minimising the modelled cost: accumulation W n mixed additions
// (10 products each) plus reduction ~3.5 B additions per set (14 products each).
// Only c in {4, 8, 16}: for these 255 = c (W - 1) + (c - 1), so the top window is a
// full window and uniform scalars spread over all its buckets (other widths leave
// a top window of a few bits whose handful of buckets receive n digits each).
static int best_c(size_t n, int per_window) {
  int best = 4;
  double best_cost = 1e300;
  for (int c = 4; c <= 16; c *= 2) {
    const double W = 255 / c + 1;
    const double B = (double)(1u << (c - 1));
    const double sets = per_window ? W : 1.0;
    const double cost = 10.0 * W * (double)n + 14.0 * 3.5 * sets * B;
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

MsmPlan msm_plan(size_t n, int per_window, int num_cus, int c_override) {
  MsmPlan p;
  const MsmTune t = tune_msm_get(n);
  int c = c_override > 0 ? c_override : (per_window && t.c > 0 ? t.c : best_c(n, per_window));
  p.c = c;
  p.W = 255 / c + 1;
  p.B = 1u << (c - 1);
  p.per_window = per_window;
  p.sets = per_window ? p.W : 1;
  p.NB = (uint32_t)p.sets * p.B;
  size_t M = (size_t)p.W * n;
  const uint32_t tpc = per_window ? (t.tpc_window ? t.tpc_window : 512) : (t.tpc_shared ? t.tpc_shared : 512);
  size_t target_threads = (size_t)std::max(num_cus, 1) * tpc;
  // Chunks of K digits in whole rounds of target_threads: the fewest rounds that keep
  // K <= 256, then K rounded up.  Rounded down, K left a few more chunks than threads
  // and a second, nearly empty round (2^18 - 1 terms: K = 31, 135 300 chunks for
  // 131 072 threads, accumulation 0.67 -> 0.87 ms).
  const size_t rounds = std::max<size_t>(1, (M + 256 * target_threads - 1) / (256 * target_threads));
  const size_t K = (M + rounds * target_threads - 1) / (rounds * target_threads);
  p.K = (uint32_t)std::max<size_t>(2, K);
  return p;
}

// Raises k_sort_scatter's dynamic-LDS limit to what SCAT_MAX_D1 bins need.
static int scatter_lds_attr() {
  static std::once_flag once;
  static int rc = MZK_OK;
  std::call_once(once, [] {
    const int bytes = (int)(scatter_lds_words(SCAT_MAX_D1, false) * 4);
    if (hipFuncSetAttribute((const void*)k_sort_bins, hipFuncAttributeMaxDynamicSharedMemorySize, 128 << 10) !=
            hipSuccess ||
        hipFuncSetAttribute((const void*)k_sort_scatter<false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
            hipSuccess ||
        hipFuncSetAttribute((const void*)k_sort_scatter<true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
            hipSuccess) {
      set_error("msm: cannot raise the scatter's LDS limit to %d bytes", bytes);
      rc = MZK_ERR_HIP;
    }
  });
  return rc;
}

// Host workers for the per-job tails: up to 15 threads started once and kept for the
// process (a batch's tails took ~0.1 ms more when each call started its own threads).
// One batch runs at a time; a caller that finds the pool busy runs its tails itself.
//
// A sleeping worker takes ~50-100 us to wake, longer than a five-job tail takes on the
// calling thread alone, so a caller about to wait for the device calls prepare(n): up to
// SPIN_WORKERS workers wake then and poll for the batch (at most SPIN_NS) while the device
// finishes.  run() waits only for the batch's jobs and for the workers that took it, never
// for workers still asleep: the batch is withdrawn once the caller has drained it, and a
// worker that wakes later finds nothing and sleeps again.
class TailPool {
 public:
  static TailPool& get() {
    static TailPool* p = new TailPool();   // never destroyed: its workers wait until exit
    return *p;
  }
  void prepare(size_t n) {
    if (n < 2 || workers_.empty() || forked_child()) return;
    spin_until.store(now_ns() + SPIN_NS, std::memory_order_relaxed);
    spin_slots.store((int)std::min<size_t>(n - 1, SPIN_WORKERS), std::memory_order_relaxed);
    {
      std::lock_guard<std::mutex> g(mu);
      wake++;
    }
    cv.notify_all();
  }
  // fn(j) for j < n on the workers and the calling thread
  void run(size_t n, const std::function<void(size_t)>& fn) {
    std::unique_lock<std::mutex> busy(run_mu, std::try_to_lock);
    if (!busy.owns_lock() || workers_.empty() || n < 2 || forked_child()) {
      for (size_t j = 0; j < n; j++) fn(j);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu);
      task = &fn;
      total = n;
      next.store(0, std::memory_order_relaxed);
      completed.store(0, std::memory_order_relaxed);
      holders = 0;
      gen.fetch_add(1, std::memory_order_release);
    }
    cv.notify_all();
    for (size_t j; (j = next.fetch_add(1)) < n;) {
      fn(j);
      completed.fetch_add(1, std::memory_order_acq_rel);
    }
    std::unique_lock<std::mutex> g(mu);
    task = nullptr;   // withdrawn: workers that wake from now on leave it alone
    done.wait(g, [&] { return holders == 0 && completed.load(std::memory_order_acquire) == n; });
  }

 private:
  static constexpr size_t SPIN_WORKERS = 4;
  static constexpr uint64_t SPIN_NS = 5000000;   // 5 ms
  static uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
  }
  // a child of fork() has none of the parent's worker threads: it runs its tails itself
  static std::atomic<bool>& forked_flag() {
    static std::atomic<bool> f{false};
    return f;
  }
  static bool forked_child() { return forked_flag().load(std::memory_order_relaxed); }
  TailPool() {
    pthread_atfork(nullptr, nullptr, [] { forked_flag().store(true); });
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(15, hw > 1 ? hw - 1 : 0);
    for (size_t t = 0; t < nt; t++) workers_.emplace_back([this] { loop(); });
    for (auto& w : workers_) w.detach();
  }
  void loop() {
    uint64_t seen = 0, seen_wake = 0;
    for (;;) {
      bool spin = false;
      {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&] { return gen.load(std::memory_order_relaxed) != seen || wake != seen_wake; });
        spin = gen.load(std::memory_order_relaxed) == seen;   // a prepare() and no batch yet
        seen_wake = wake;
      }
      if (spin) {
        if (spin_slots.fetch_sub(1, std::memory_order_relaxed) <= 0) continue;   // enough pollers
        const uint64_t until = spin_until.load(std::memory_order_relaxed);
        while (gen.load(std::memory_order_acquire) == seen && now_ns() < until) __builtin_ia32_pause();
        if (gen.load(std::memory_order_acquire) == seen) continue;   // no batch came: sleep again
      }
      const std::function<void(size_t)>* fn = nullptr;
      size_t n = 0;
      {
        std::lock_guard<std::mutex> g(mu);
        seen = gen.load(std::memory_order_relaxed);
        if (task == nullptr) continue;   // already drained and withdrawn
        fn = task;
        n = total;
        holders++;
      }
      for (size_t j; (j = next.fetch_add(1)) < n;) {
        (*fn)(j);
        completed.fetch_add(1, std::memory_order_acq_rel);
      }
      std::lock_guard<std::mutex> g(mu);
      if (--holders == 0) done.notify_all();
    }
  }
  std::mutex run_mu, mu;
  std::condition_variable cv, done;
  const std::function<void(size_t)>* task = nullptr;   // the current batch, under mu
  size_t total = 0;
  int holders = 0;                  // workers that took the current batch, under mu
  uint64_t wake = 0;                // prepare() requests, under mu
  std::atomic<uint64_t> gen{0};     // batches started (written under mu)
  std::atomic<size_t> next{0}, completed{0};
  std::atomic<uint64_t> spin_until{0};
  std::atomic<int> spin_slots{0};
  std::vector<std::thread> workers_;
};

static void parallel_jobs(size_t n, const std::function<void(size_t)>& fn) { TailPool::get().run(n, fn); }
static void prepare_jobs(size_t n) { TailPool::get().prepare(n); }

// Host Horner over one job's channel positions.  Set s (of sets_per_job) weighs
// 2^(c s); inside it, sum_t A_t has weight 1 and U_j (bit j of t) 2^(log2 L + j).
static void job_tail(const uint32_t* h, int sets_per_job, int c, uint32_t nch, uint32_t lbits, uint32_t tbits,
                     uint64_t out[18]) {
  const int top = (sets_per_job - 1) * c + (int)(lbits + tbits);
  std::vector<int> at(top + 1, -1), at2(top + 1, -1);   // channel per position (two may share)
  for (int s = 0; s < sets_per_job; s++) {
    for (uint32_t h2 = 0; h2 < nch; h2++) {
      const int pos = s * c + (h2 == 0 ? 0 : (int)(lbits + h2 - 1));
      (at[pos] < 0 ? at[pos] : at2[pos]) = (int)(s * nch + h2);
    }
  }
  host::Xyzz acc = host::Xyzz::inf();
  for (int pos = top; pos >= 0; pos--) {
    acc = host::xyzz_dbl(acc);
    if (at[pos] >= 0) acc = host::xyzz_add(acc, to_host(h + 48 * (size_t)at[pos]));
    if (at2[pos] >= 0) acc = host::xyzz_add(acc, to_host(h + 48 * (size_t)at2[pos]));
  }
  host::xyzz_to_jacobian(acc, out);
}

// Host Horner over one job's per-window set sums: sum_w 2^(c w) T_w.
static void job_tail_sets(const uint32_t* h, int windows, int c, uint64_t out[18]) {
  host::Xyzz acc = host::Xyzz::inf();
  for (int w = windows - 1; w >= 0; w--) {
    for (int d = 0; d < c; d++) acc = host::xyzz_dbl(acc);
    acc = host::xyzz_add(acc, to_host(h + 48 * (size_t)w));
  }
  host::xyzz_to_jacobian(acc, out);
}

int msm_run_jobs(mzk_ctx* c, const MsmPlan& p, const std::vector<MsmJob>& jobs, const uint32_t* d_bases,
                 uint32_t ref_stride, uint64_t* outs, bool records, const std::function<int()>* on_enqueued) {
  const size_t nj = jobs.size();
  size_t ntot = 0;
  for (auto& j : jobs) ntot += j.n;
  if (ntot == 0) {
    memset(outs, 0, nj * 18 * 8);
    return MZK_OK;
  }
  const size_t M_max = (size_t)p.W * ntot;
  MZK_REQUIRE(M_max < 0x7fffffffull, "msm: %zu window digits exceed the 2^31 index range", M_max);
  const int spj = p.per_window ? p.W : 1;   // bucket sets per job
  const size_t sets = nj * (size_t)spj;
  MZK_REQUIRE(sets * p.B < (1ull << 31), "msm: %zu bucket sets too many", sets);
  const uint32_t NB = (uint32_t)(sets * p.B);

  // tiles of at most ~2048 scalars, never crossing a job
  std::vector<JobDesc> jd(nj);
  std::vector<TileDesc> td;
  // sort groups: consecutive jobs whose bucket sets hold at most SORT_GROUP_BUCKETS
  // buckets together, each group sorted on its own key range (one histogram block, one
  // set of bins per group) so that bins stay short with few second-level counters
  const size_t job_buckets = (size_t)spj * p.B;
  const size_t GJ = NB <= SORT_GROUP_BUCKETS ? nj : std::max<size_t>(1, SORT_GROUP_BUCKETS / job_buckets);
  const size_t ngroups = (nj + GJ - 1) / GJ;
  const uint32_t NBg = (uint32_t)std::min<size_t>(NB, GJ * job_buckets);
  // top-level bins: 512 when the second level's 2^l2 counters allow it (longer runs per
  // bin in the scatter), more for large key spaces (the second level keeps <= 2^11)
  const int nb_bits = ceil_log2(NBg);
  const uint32_t l2_top = (uint32_t)std::min(11, std::max(0, nb_bits - 9));
  const size_t D1_top = ((size_t)NBg + (1u << l2_top) - 1) >> l2_top;
  // ~1024-1536 tiles plus at most one partial tile per job, fewer when the bins are many:
  // the (bin, tile) histogram stays within the two-level scan (2048 x 2048 entries)
  const size_t max_tiles = (size_t)SCAN_TILE * SCAN_TILE / std::max<size_t>(D1_top, 1);
  MZK_REQUIRE(max_tiles > nj + 1, "msm: %zu bucket sets too many for the sort", sets);
  const size_t tile_target = std::min<size_t>(1536, max_tiles - nj - 1);
  const size_t tile = std::max<size_t>(256, (ntot + tile_target - 1) / tile_target);
  MZK_REQUIRE(nj < 512, "msm: at most 511 jobs per batch (%zu given)", nj);
  size_t canon_off = 0;
  std::vector<GroupDesc> gd(ngroups);
  for (size_t j = 0; j < nj; j++) {
    jd[j].coeffs = (const uint32_t*)jobs[j].d_coeffs;
    jd[j].n = (uint32_t)jobs[j].n;
    jd[j].ref_offset = jobs[j].ref_offset;
    jd[j].canon_off = (uint32_t)canon_off;
    jd[j].set_base = (uint32_t)(j * spj);
    const size_t g = j / GJ;
    if (j % GJ == 0) gd[g].tile0 = (uint32_t)td.size();
    for (size_t lo = 0; lo < jobs[j].n; lo += tile)
      td.push_back(TileDesc{(uint32_t)j, (uint32_t)lo, (uint32_t)std::min(lo + tile, jobs[j].n), 0, 0,
                            (uint32_t)(g * NBg)});
    canon_off += jobs[j].n;
  }
  for (size_t g = 0; g < ngroups; g++)
    gd[g].G = (uint32_t)((g + 1 < ngroups ? gd[g + 1].tile0 : td.size()) - gd[g].tile0);
  SortArgs sa;
  memset(&sa, 0, sizeof sa);   // padding too: the struct is part of the graph key
  sa.G = (uint32_t)td.size();
  sa.c = p.c;
  sa.W = p.W;
  sa.per_window = p.per_window;
  sa.l2 = l2_top;
  sa.D1 = (uint32_t)D1_top;
  sa.ref_stride = ref_stride;
  {   // one word per sorted digit when the low key bits, the reference and its sign fit
    uint64_t ref_end = 0;   // references are below this
    for (auto& j : jobs) ref_end = std::max<uint64_t>(ref_end, (uint64_t)j.ref_offset + j.n);
    if (!p.per_window) ref_end += (uint64_t)(p.W - 1) * ref_stride;
    const uint32_t ref_bits = (uint32_t)ceil_log2(std::max<uint64_t>(ref_end, 2));
    // fewer low bits (more top-level bins) when that makes the word fit: one word per
    // digit through the scatter and the bins, and more, shorter bins for k_sort_bins
    // (5 x 2^18 shared sets against a 2^20 SRS: 320 bins of 2^9 -> 1280 of 2^7,
    // k_sort_bins 0.39 -> 0.13 ms)
    if (sa.l2 + ref_bits + 1 > 32 && ref_bits + 1 <= 32) {
      const uint32_t l2 = 31 - ref_bits;
      const size_t D1 = ((size_t)NBg + (1u << l2) - 1) >> l2;
      const size_t hn_tiles = (D1 * sa.G + SCAN_TILE - 1) / SCAN_TILE;
      if (D1 <= SCAT_MAX_D1 && hn_tiles <= (size_t)SCAN_TILE) {
        sa.l2 = l2;
        sa.D1 = (uint32_t)D1;
      }
    }
    sa.pack = sa.l2 + ref_bits + 1 <= 32 ? ref_bits + 1 : 0;
  }
  sa.NBg = NBg;
  for (size_t g = 0; g < ngroups; g++)   // each tile's histogram column within its group's block
    for (uint32_t t = gd[g].tile0; t < gd[g].tile0 + gd[g].G; t++) {
      td[t].hbase = sa.D1 * gd[g].tile0 + (t - gd[g].tile0);
      td[t].hstride = gd[g].G;
    }
  const size_t HN = (size_t)sa.D1 * sa.G;   // histogram entries
  const size_t tiles = (HN + SCAN_TILE - 1) / SCAN_TILE;
  MZK_REQUIRE(tiles <= (size_t)SCAN_TILE, "msm: sort histogram too large (%zu)", HN);
  // Per-window sets: a bucket holds ~n / B digits of its job, merged from its chunk
  // partials by one dependent chain in the reduction; chunks of at least n / (4 B)
  // digits keep that chain within ~5 additions (small batches would otherwise cut
  // their few digits into chunks of 2 or 3).
  uint32_t K = p.K;
  if (p.per_window) {
    size_t nmax = 0;
    for (auto& j : jobs) nmax = std::max(nmax, j.n);
    K = (uint32_t)std::min<size_t>(256, std::max<size_t>(K, (nmax + 4 * (size_t)p.B - 1) / (4 * (size_t)p.B)));
  }
  const size_t nchunks = (M_max + K - 1) / K;
  // The chunk size the kernels use comes from the digits the sort found (k_chunk_size):
  // for shared sets ceil(M / nchunks) refs per thread, at least 2 and at least sqrt(D) / 4
  // for the largest bucket's D digits, so sparse scalars (zeros give no digits: a
  // delta-encoded column, a mostly-zero advice column) still spread their digits over
  // every thread while skewed ones (a boolean column) keep their long buckets' partials
  // few; per-window sets keep K (their chains' bound).
  sa.kchunks = (uint32_t)nchunks;
  sa.kmax = K;
  sa.kmin = p.per_window ? K : std::min<uint32_t>(2, K);
  // reduction geometry: segments of L buckets, T per set; channels A + bits of t.
  // One MSM alone: a thread per bucket keeps chains short.  Many sets: segments of
  // 4 cut the channel work, which is then parallel across sets; with segments enough for
  // ~3 waves per SIMD at 8 buckets (48 shared sets and more), the reduction is bound by
  // throughput rather than by its dependent chains and segments of 8 halve the segment
  // tree's work (96 uneven commits -1.3 %; five sets were 4 % slower that way).
  // Up to TREE_SETS sets of 2^t buckets (t >= 8) take the block tree of a single set, a
  // quad of lanes per bucket: two to four sets in segments leave the SIMDs mostly idle
  // through the segments' dependent additions (2^15-bucket sets, reduction kernels: two
  // sets 0.51 -> 0.30 ms, four 0.49 -> 0.46 ms; from five sets on the tree's extra
  // additions cost more than the segments' chains: five 0.43 -> 0.47 ms, eight 0.58 -> 0.64).
  const bool small_sets = p.per_window && p.B <= 256;
  const bool many_sets = (size_t)sets * (p.B / 8) >= (size_t)3 * 64 * 1024;
  const bool tree = sets <= TREE_SETS && !small_sets && (p.B & (p.B - 1)) == 0 && p.B >= 2 * CT_LEAVES;
  const uint32_t L = std::min<uint32_t>(tree || sets == 1 ? 1 : (many_sets ? 8 : 4), p.B);
  const uint32_t T = p.B / L;
  const uint32_t tbits = (uint32_t)ceil_log2(T);
  const uint32_t nch = 1 + tbits;
  const uint32_t parts = std::max<uint32_t>(1, T / 2048);
  const size_t nchan = sets * nch;
  const uint32_t lbits = (uint32_t)ceil_log2(L);
  // segments of L > 1 buckets, 2^7 <= T <= 2^14 per set: channels by the segment tree
  // (k_seg_tree, k_seg_top), its records written over the chunk partials once merged
  const bool segtree = L > 1 && !(p.per_window && p.B <= 256) && (T & (T - 1)) == 0 && T >= 2 * ST_LEAVES &&
                       T <= 256 * ST_LEAVES;

  // scratch A: descriptors | canonical scalars | hist | hoff (HN+1) | tile sums | off (NB+1) | K |
  // the bins' largest buckets
  const size_t desc_words =
      (nj * sizeof(JobDesc) + td.size() * sizeof(TileDesc) + ngroups * sizeof(GroupDesc) + 31) / 4 + 8;
  // off: NB + 1 offsets, K, then the bins' largest buckets (8-word aligned)
  size_t a_words = desc_words + 8 * ntot + HN + (HN + 1) + tiles + ((size_t)NB + 8) + ngroups * sa.D1 + 64;
  MZK_TRY(c->scratch[SCR_MSM_A].reserve(a_words * 4));
  uint32_t* base = (uint32_t*)c->scratch[SCR_MSM_A].ptr;
  JobDesc* d_jobs = (JobDesc*)base;
  TileDesc* d_tiles = (TileDesc*)((char*)base + ((nj * sizeof(JobDesc) + 15) & ~(size_t)15));
  uint32_t* canon = base + ((desc_words + 3) & ~(size_t)3);
  uint32_t* hist = canon + 8 * ntot;
  uint32_t* hoff = hist + HN;
  uint32_t* tsums = hoff + HN + 1;
  uint32_t* off = tsums + tiles;
  uint32_t* bmax = off + (((size_t)NB + 2 + 7) & ~(size_t)7);
  // descriptors through pinned memory: [jobs | pad | tiles] as laid out in scratch A, one
  // asynchronous copy; the channel sums come back through the same buffer after it
  const size_t tiles_at = (size_t)((char*)d_tiles - (char*)d_jobs);
  const size_t groups_at = (tiles_at + td.size() * sizeof(TileDesc) + 15) & ~(size_t)15;
  GroupDesc* d_groups = (GroupDesc*)((char*)d_jobs + groups_at);
  const size_t desc_bytes = groups_at + ngroups * sizeof(GroupDesc);
  const size_t res_at = (desc_bytes + 255) & ~(size_t)255;
  const size_t res_words = std::max(nchan, sets) * 48 + 4;
  MZK_TRY(c->pinned.reserve(res_at + res_words * 4));
  char* pin = (char*)c->pinned.ptr;
  memcpy(pin, jd.data(), nj * sizeof(JobDesc));
  memcpy(pin + tiles_at, td.data(), td.size() * sizeof(TileDesc));
  memcpy(pin + groups_at, gd.data(), ngroups * sizeof(GroupDesc));
  MZK_HIP(hipMemcpyAsync(d_jobs, pin, desc_bytes, hipMemcpyHostToDevice, c->stream));
  uint32_t* res = (uint32_t*)(pin + res_at);
  sa.jobs = d_jobs;
  sa.tiles = d_tiles;
  sa.groups = d_groups;
  // scratch B: tmp keys | tmp refs | sorted refs
  MZK_TRY(c->scratch[SCR_MSM_B].reserve(3 * M_max * 4 + 64));
  uint32_t* tmp_key = (uint32_t*)c->scratch[SCR_MSM_B].ptr;
  uint32_t* tmp_ref = tmp_key + M_max;
  uint32_t* refs = tmp_ref + M_max;
  // scratch C: chunk partials (nchunks + NB slots) then segment sums A, R (sets * T each),
  // all in the radix-2^28 point form (MZK_X28_WORDS words each)
  const size_t nseg = sets * T;
  size_t c_words = (nchunks + NB) * MZK_X28_WORDS + 2 * nseg * MZK_X28_WORDS;
  MZK_TRY(c->scratch[SCR_MSM_C].reserve(c_words * 4));
  uint32_t* partials = (uint32_t*)c->scratch[SCR_MSM_C].ptr;
  uint32_t* segA = partials + (nchunks + NB) * MZK_X28_WORDS;
  uint32_t* segR = segA + nseg * MZK_X28_WORDS;
  // scratch D: channel partials and channel sums (small sets: one sum per set)
  size_t d_words = nchan * parts * MZK_X28_WORDS + std::max(nchan, sets) * 48 + 64;
  MZK_TRY(c->scratch[SCR_MSM_D].reserve(d_words * 4));
  uint32_t* chan_parts = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  uint32_t* chan = chan_parts + nchan * parts * MZK_X28_WORDS;

  // the scatter's LDS (scatter_lds_words): above the 64 KiB a kernel gets without asking
  // once D1 passes ~1300 bins (unpacked) or ~2700 (packed)
  MZK_REQUIRE(sa.D1 <= SCAT_MAX_D1, "msm: %u top-level sort bins exceed %u (too many bucket sets)", sa.D1,
              SCAT_MAX_D1);
  const size_t scat_lds = scatter_lds_words(sa.D1, sa.pack != 0) * 4;
  // k_sort_bins places each reference with a 4-byte store at its bucket's cursor: a
  // workgroup keeps 2^l2 partly written lines open, and once an XCD's resident bins hold
  // more than its L2 can merge, the partial lines go back to HBM (fifteen 2^18 commits,
  // 1024 bins of 512 buckets, 4 bins per CU: 0.675 ms; 1 per CU: 0.42 ms).  Extra LDS caps
  // the resident bins at ~512 / 2^l2 per CU (one XCD's 32 CUs then hold ~2 MB of open lines).
  const uint32_t bins_per_cu = std::max<uint32_t>(1, 512u >> sa.l2);
  const size_t bins_lds = bins_per_cu < 8 && ngroups * sa.D1 > (size_t)c->num_cus * bins_per_cu
                              ? std::max<size_t>((1u << sa.l2) * 4, BINS_LDS_CU / (2 * bins_per_cu + 1) * 2)
                              : (1u << sa.l2) * 4;
  MZK_TRY(scatter_lds_attr());
  // the device sequence (sort, accumulation, reduction) as one cached graph: every value
  // its launches take is in the key; the descriptors above are copied before it runs
  GraphKey gk;
  gk.add('M').add(sa).add(canon).add(hist).add(hoff).add(HN).add(tsums).add(tiles).add(tmp_key).add(tmp_ref);
  gk.add(refs).add(off).add(NB).add(K).add(nchunks).add(records).add(d_bases).add(partials).add(small_sets);
  gk.add(p.B).add(sets).add(chan).add(segA).add(segR).add(L).add(nseg).add(T).add(nch).add(parts).add(chan_parts);
  gk.add(nchan).add(tree).add(segtree).add(ngroups);
  MZK_TRY(run_graph(c, gk.s, [&]() -> int {
    MZK_LAUNCH(c, "k_sort_hist", k_sort_hist, dim3(sa.G), dim3(SORT_BLOCK), sa.D1 * 4, sa, canon, hist);
    MZK_LAUNCH(c, "k_scan_tiles", k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, hist, hoff, HN, tsums);
    MZK_LAUNCH(c, "k_scan_sums", k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, tsums, (int)tiles, hoff, HN);
    MZK_LAUNCH(c, "k_scan_add", k_scan_add, dim3(grid_for(HN, 256)), dim3(256), 0, hoff, HN, tsums);
    if (sa.pack)
      MZK_LAUNCH(c, "k_sort_scatter", k_sort_scatter<true>, dim3(sa.G), dim3(SORT_BLOCK), scat_lds, canon, sa, hoff,
                 tmp_key, tmp_ref);
    else
      MZK_LAUNCH(c, "k_sort_scatter", k_sort_scatter<false>, dim3(sa.G), dim3(SORT_BLOCK), scat_lds, canon, sa, hoff,
                 tmp_key, tmp_ref);
    MZK_LAUNCH(c, "k_sort_bins", k_sort_bins, dim3((unsigned)(ngroups * sa.D1)), dim3(SORT_BLOCK), bins_lds,
               hoff, sa, NB,
               tmp_key, tmp_ref, refs, off, bmax);
    MZK_LAUNCH(c, "k_chunk_size", k_chunk_size, dim3(1), dim3(256), 0, off, NB, bmax, (uint32_t)(ngroups * sa.D1), sa);
    if (records)
      MZK_LAUNCH(c, "k_msm_accum", k_msm_accum<true>, dim3(grid_for(nchunks, 64)), dim3(64), 0, refs, off, NB,
                 d_bases, partials);
    else
      MZK_LAUNCH(c, "k_msm_accum", k_msm_accum<false>, dim3(grid_for(nchunks, 64)), dim3(64), 0, refs, off, NB,
                 d_bases, partials);
    MZK_LAUNCH(c, "k_long_merge", k_long_merge, dim3(grid_for(NB, 64)), dim3(256), 0, off, NB, partials);
    if (small_sets) {
      // per-set weighted sums in LDS (one per set, into chan); a quad of lanes per bucket
      // when the sets are too few to fill the SIMDs one lane per bucket
      if (p.B <= 128 && sets * p.B <= QUAD_SETS_MAX_BUCKETS) {
        const uint32_t spb = 128 / p.B;
        MZK_LAUNCH(c, "k_set_reduce", k_set_reduce_quad, dim3((unsigned)((sets + spb - 1) / spb)), dim3(512),
                   2 * 128 * MZK_X28_WORDS * 4, off, NB, p.B, (uint32_t)sets, partials, chan);
      } else {
        const uint32_t spb = 256 / p.B;
        MZK_LAUNCH(c, "k_set_reduce", k_set_reduce, dim3((unsigned)((sets + spb - 1) / spb)), dim3(256),
                   256 * MZK_X28_WORDS * 4, off, NB, p.B, (uint32_t)sets, partials, chan);
      }
    } else if (tree) {
      // sets of 2^t buckets: the merge and the tree by blocks, then the channels of each set
      MZK_LAUNCH(c, "k_chan_tree", k_chan_tree, dim3(NB / CT_LEAVES), dim3(4 * CT_LEAVES),
                 (2 * CT_LEAVES - 1) * MZK_X28_WORDS * 4, off, NB, partials, segA);
      MZK_LAUNCH(c, "k_chan_top", k_chan_top, dim3((unsigned)nch, (unsigned)sets), dim3(4 * CT_TOP_QUADS),
                 CT_TOP_QUADS * MZK_X28_WORDS * 4, segA, T / CT_LEAVES, chan);
    } else {
      if (L == 1)   // a single set: a quad per bucket
        MZK_LAUNCH(c, "k_bucket_seg", k_bucket_merge_quad, dim3(grid_for((size_t)NB * 4, 256)), dim3(256), 0, off, NB,
                   partials, segA, segR);
      else
        MZK_LAUNCH(c, "k_bucket_seg", k_bucket_seg, dim3(grid_for(nseg, 64)), dim3(64), 0, off, NB, L, partials,
                   segA, segR);
      if (segtree) {   // the segments' channels by the block tree: records into the chunk partials' space
        // nrec = T / ST_LEAVES is a power of two >= 2; k_seg_top's quads: min(nrec, 128), at least 16
        const uint32_t nrec = T / ST_LEAVES;
        const uint32_t top_quads = std::min<uint32_t>(ST_TOP_QUADS, std::max<uint32_t>(16, nrec));
        MZK_LAUNCH(c, "k_seg_tree", k_seg_tree, dim3((unsigned)(nseg / ST_LEAVES)), dim3(4 * ST_LEAVES),
                   (3 * ST_LEAVES - 1) * MZK_X28_WORDS * 4, segA, segR, partials);
        MZK_LAUNCH(c, "k_seg_top", k_seg_top, dim3((unsigned)nchan), dim3(4 * top_quads),
                   top_quads * MZK_X28_WORDS * 4, partials, nrec, nch, chan);
      } else {
        MZK_LAUNCH(c, "k_chan_sum", k_chan_sum, dim3((unsigned)(nchan * parts)), dim3(CHAN_BLOCK),
                   CHAN_BLOCK * MZK_X28_WORDS * 4, segA, segR, T, nch, parts, chan_parts);
        MZK_LAUNCH(c, "k_seg_sum", k_seg_sum, dim3((unsigned)nchan), dim3(64), 64 * MZK_X28_WORDS * 4, chan_parts,
                   parts, chan);
      }
    }
    MZK_HIP(hipGetLastError());
    return MZK_OK;
  }));
  if (on_enqueued) MZK_TRY((*on_enqueued)());
  if (small_sets) {
    // the window combination on the host
    uint32_t* tot = chan;
    const uint32_t* h = res;
    double h0 = host_ms();
    prepare_jobs(nj);
    MZK_HIP(hipMemcpyAsync(res, tot, sets * 48 * 4, hipMemcpyDeviceToHost, c->stream));
    MZK_HIP(hipMemcpyAsync(res + sets * 48, off + NB, 8, hipMemcpyDeviceToHost, c->stream));
    MZK_HIP(hipStreamSynchronize(c->stream));
    const uint32_t M = res[sets * 48], Keff = res[sets * 48 + 1];
    double h1 = host_ms();
    c->prof.host("host_wait", h1 - h0);
    c->last_msm[0] = (uint64_t)p.c;
    c->last_msm[1] = (uint64_t)p.W;
    c->last_msm[2] = (uint64_t)sets;
    c->last_msm[3] = Keff;
    c->last_msm[4] = M;
    c->last_msm[5] = ntot;
    parallel_jobs(nj, [&](size_t j) {
      if (jobs[j].n == 0) memset(outs + 18 * j, 0, 18 * 8);
      else job_tail_sets(h + j * (size_t)spj * 48, spj, p.c, outs + 18 * j);
    });
    c->prof.host("host_horner", host_ms() - h1);
    return MZK_OK;
  }
  const uint32_t* h = res;
  double h0 = host_ms();
  prepare_jobs(nj);
  MZK_HIP(hipMemcpyAsync(res, chan, nchan * 48 * 4, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipMemcpyAsync(res + nchan * 48, off + NB, 8, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  const uint32_t M = res[nchan * 48], Keff = res[nchan * 48 + 1];
  double h1 = host_ms();
  c->prof.host("host_wait", h1 - h0);
  c->last_msm[0] = (uint64_t)p.c;
  c->last_msm[1] = (uint64_t)p.W;
  c->last_msm[2] = (uint64_t)sets;
  c->last_msm[3] = Keff;
  c->last_msm[4] = M;
  c->last_msm[5] = ntot;
  const size_t per_job = (size_t)spj * nch * 48;
  parallel_jobs(nj, [&](size_t j) {
    if (jobs[j].n == 0) memset(outs + 18 * j, 0, 18 * 8);
    else job_tail(h + j * per_job, spj, p.c, nch, lbits, tbits, outs + 18 * j);
  });
  c->prof.host("host_horner", host_ms() - h1);
  return MZK_OK;
}

int msm_run(mzk_ctx* c, const MsmPlan& p, const uint32_t* d_scalars, size_t n, const uint32_t* d_bases,
            uint32_t ref_stride, uint32_t ref_offset, bool records, uint64_t out[18]) {
  std::vector<MsmJob> jobs(1);
  jobs[0].d_coeffs = d_scalars;
  jobs[0].n = n;
  jobs[0].ref_offset = ref_offset;
  return msm_run_jobs(c, p, jobs, d_bases, ref_stride, out, records);
}

}  // namespace mzk

using namespace mzk;

extern "C" {

int mzk_msm_dev(mzk_ctx* c, const void* d_coeffs, const void* d_bases, size_t n, uint64_t out[18]) {
  MZK_REQUIRE(c != nullptr && out != nullptr, "mzk_msm_dev: null argument");
  if (n == 0) {
    memset(out, 0, 18 * 8);
    return MZK_OK;
  }
  MZK_REQUIRE(d_coeffs && d_bases, "mzk_msm_dev: null input");
  MZK_REQUIRE(n < (1ull << 31), "mzk_msm_dev: n too large");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  MsmPlan p = msm_plan(n, 1, c->num_cus, 0);
  return msm_run(c, p, (const uint32_t*)d_coeffs, n, (const uint32_t*)d_bases, 0, 0, false, out);
}

static int msm_host(mzk_ctx* c, const uint64_t* coeffs, const uint64_t* bases, size_t n, uint64_t out[18]) {
  MZK_REQUIRE(c != nullptr && out != nullptr, "msm: null argument");
  if (n == 0) {
    memset(out, 0, 18 * 8);
    return MZK_OK;
  }
  MZK_REQUIRE(coeffs && bases, "msm: null input");
  MZK_REQUIRE(n < (1ull << 31), "msm: n too large");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  MZK_TRY(c->scratch[SCR_IN0].reserve(n * 32));
  MZK_TRY(c->scratch[SCR_IN1].reserve(n * 96));
  MZK_HIP(hipMemcpyAsync(c->scratch[SCR_IN0].ptr, coeffs, n * 32, hipMemcpyHostToDevice, c->stream));
  MZK_HIP(hipMemcpyAsync(c->scratch[SCR_IN1].ptr, bases, n * 96, hipMemcpyHostToDevice, c->stream));
  MsmPlan p = msm_plan(n, 1, c->num_cus, 0);
  return msm_run(c, p, (const uint32_t*)c->scratch[SCR_IN0].ptr, n, (const uint32_t*)c->scratch[SCR_IN1].ptr, 0,
                 0, false, out);
}

int mzk_msm_best(mzk_ctx* c, const uint64_t* coeffs, const uint64_t* bases, size_t n, uint64_t out[18]) {
  return msm_host(c, coeffs, bases, n, out);
}

// Zero scalars produce no window digits, so the zero filter of msm_specific
// (kzg/msm.rs:176-181) is implicit in the bucket histogram.
int mzk_msm_specific(mzk_ctx* c, const uint64_t* coeffs, const uint64_t* bases, size_t n, uint64_t out[18]) {
  return msm_host(c, coeffs, bases, n, out);
}

// msm_parallel (msm.rs:430-455): the reference splits the terms over threads; the sum
// is the same unique point msm_best computes.
int mzk_msm_parallel(mzk_ctx* c, const uint64_t* coeffs, const uint64_t* bases, size_t n, uint64_t out[18]) {
  return msm_host(c, coeffs, bases, n, out);
}

// Jacobian (X, Y, Z) as XYZZ (X, Y, Z^2, Z^3); Z = 0 is the identity.
static host::Xyzz jacobian_to_xyzz(const uint64_t jac[18]) {
  host::Xyzz p;
  memcpy(p.X.l, jac, 48);
  memcpy(p.Y.l, jac + 6, 48);
  host::Fp Z;
  memcpy(Z.l, jac + 12, 48);
  if (host::fp_is_zero(Z)) return host::Xyzz::inf();
  p.ZZ = host::fp_sqr(Z);
  p.ZZZ = host::fp_mul(p.ZZ, Z);
  return p;
}

// The doublings msm_serial (msm.rs:338-423) applies to its incoming accumulator: c per
// window (:364-367), W = 8 max_byte / c + 1 windows (:359), c by the rule of :340-346 and
// max_byte the highest non-zero byte over the canonical scalars; 0 when every scalar is
// zero (the reference then returns before touching acc, :360-362).
static uint64_t serial_doublings(const uint64_t* coeffs, size_t n) {
  const size_t c = n < 4 ? 1 : (n < 32 ? 3 : (size_t)std::ceil(std::log((double)(uint32_t)n)));
  host::Fr raw_one = host::fr_zero();
  raw_one.l[0] = 1;
  int max_byte = 0;
  for (size_t i = 0; i < n && max_byte < 32; i++) {
    const host::Fr s = host::fr_mul(host::fr_const(coeffs + 4 * i), raw_one);   // canonical
    for (int b = 31; b >= max_byte; b--)
      if ((s.l[b / 8] >> (8 * (b % 8))) & 0xff) {
        max_byte = b + 1;
        break;
      }
  }
  return max_byte == 0 ? 0 : (uint64_t)c * ((size_t)max_byte * 8 / c + 1);
}

// msm_serial (msm.rs:338-423), acc Jacobian in and out: acc <- [2^d] acc + sum_i coeffs[i]
// bases[i] with d = serial_doublings (d = 0 and acc unchanged for all-zero scalars).
// From the identity -- the only way the reference itself calls it (msm.rs:430-455) --
// acc becomes the MSM.
int mzk_msm_serial(mzk_ctx* c, const uint64_t* coeffs, const uint64_t* bases, size_t n, uint64_t acc[18]) {
  MZK_REQUIRE(acc != nullptr, "mzk_msm_serial: null accumulator");
  uint64_t m[18];
  MZK_TRY(msm_host(c, coeffs, bases, n, m));
  host::Xyzz a = jacobian_to_xyzz(acc);
  if (!a.is_inf()) {
    const uint64_t d = serial_doublings(coeffs, n);
    for (uint64_t i = 0; i < d; i++) a = host::xyzz_dbl(a);
  }
  host::xyzz_to_jacobian(host::xyzz_add(a, jacobian_to_xyzz(m)), acc);
  return MZK_OK;
}

int mzk_g1_jacobian_to_affine(const uint64_t jac[18], uint64_t aff[12]) {
  MZK_REQUIRE(jac != nullptr && aff != nullptr, "mzk_g1_jacobian_to_affine: null argument");
  host::Fp X, Y, Z;
  memcpy(X.l, jac, 48);
  memcpy(Y.l, jac + 6, 48);
  memcpy(Z.l, jac + 12, 48);
  if (host::fp_is_zero(Z)) {
    memset(aff, 0, 96);
    return MZK_OK;
  }
  host::Fp zi = host::fp_inv(Z), zi2 = host::fp_mul(zi, zi);
  host::Fp x = host::fp_mul(X, zi2), y = host::fp_mul(Y, host::fp_mul(zi2, zi));
  memcpy(aff, x.l, 48);
  memcpy(aff + 6, y.l, 48);
  return MZK_OK;
}

int mzk_msm_last_plan(mzk_ctx* c, uint64_t out[6]) {
  MZK_REQUIRE(c != nullptr && out != nullptr, "mzk_msm_last_plan: null argument");
  std::lock_guard<std::mutex> g(c->mu);
  memcpy(out, c->last_msm, sizeof c->last_msm);
  return MZK_OK;
}

}  // extern "C"
