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
//                     partial bucket sums (mixed additions, 8M + 2S): equal work per
//                     thread whatever the bucket sizes
//  6. k_bucket_seg    per segment of L = 4 buckets: merge each bucket's chunk
//                     partials, running sums -> A_t = sum_i (i+1) S, R_t = sum_i S
//  7. k_chan_sum      channels: sum_t A_t, and U_j = sum of the R_t whose index t has
//                     bit j set, so sum_b (b+1) S_b = sum_t A_t + L sum_j 2^j U_j --
//                     plain sums in any order, short dependency chains
//  8. host tail       Horner over the channel bit positions (one host core runs a
//                     dependent chain of point operations ~20x faster than one wave).
//
// Two bucket layouts share the kernels:
//  * per-window sets (any bases): key = w * B + |d| - 1, set weight 2^(c w);
//  * one shared set (resident SRS with window-shifted copies 2^(cw) g_i built once):
//    key = |d| - 1, reference = w * N + offset + i, a single set of weight 1.
#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "g1.hpp"
#include "host_ec.hpp"
#include "msm.hpp"

namespace mzk {

static constexpr uint32_t NEG_FLAG = 0x80000000u;
static constexpr int SORT_BLOCK = 256;
static constexpr int CHAN_BLOCK = 256;

// Signed window digit w of a canonical scalar (8 x u32 LE); `carry` threads the
// recoding from window to window: digit = bits + carry, digits above 2^(c-1)
// become digit - 2^c with a carry into the next window.
__device__ __forceinline__ int window_digit(const uint32_t s[8], int w, int c, uint32_t& carry) {
  const int bit = w * c;
  uint32_t bits = 0;
  if (bit < 256) {
    const int limb = bit >> 5, sh = bit & 31;
    uint64_t two = (uint64_t)s[limb];
    if (limb + 1 < 8) two |= (uint64_t)s[limb + 1] << 32;
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

struct SortArgs {
  uint32_t n, tile, G;   // scalars, scalars per tile, tiles
  int c, W, per_window;
  uint32_t l2;           // low key bits sorted in k_sort_bins
  uint32_t D1;           // bins of the top key bits
  uint32_t ref_stride, ref_offset;
};

// 1. digits + per-tile histogram of key >> l2 (LDS atomics)
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_hist(const uint32_t* __restrict__ scalars, SortArgs a,
                                                          uint32_t* __restrict__ canon, uint32_t* __restrict__ hist) {
  extern __shared__ uint32_t cnt[];   // D1 counters
  for (uint32_t i = threadIdx.x; i < a.D1; i += SORT_BLOCK) cnt[i] = 0;
  __syncthreads();
  const uint32_t B = 1u << (a.c - 1);
  const uint32_t lo = blockIdx.x * a.tile;
  const uint32_t hi = min(lo + a.tile, a.n);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += SORT_BLOCK) {
    fr_t s = from_mont(load<FrP>(scalars + 8 * (size_t)i));
    store<FrP>(canon + 8 * (size_t)i, s);
    uint32_t carry = 0;
    for (int w = 0; w < a.W; w++) {
      int d = window_digit(s.v, w, a.c, carry);
      if (d != 0) {
        uint32_t key = (uint32_t)(d > 0 ? d : -d) - 1u;
        if (a.per_window) key += (uint32_t)w * B;
        atomicAdd(&cnt[key >> a.l2], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < a.D1; b += SORT_BLOCK) hist[(size_t)b * a.G + blockIdx.x] = cnt[b];
}

// 3. scatter into bins: tmp_key[pos] = key & (2^l2 - 1), tmp_ref[pos] = reference | sign
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_scatter(const uint32_t* __restrict__ canon, SortArgs a,
                                                             const uint32_t* __restrict__ hoff,
                                                             uint32_t* __restrict__ tmp_key,
                                                             uint32_t* __restrict__ tmp_ref) {
  extern __shared__ uint32_t cur[];   // D1 cursors
  for (uint32_t b = threadIdx.x; b < a.D1; b += SORT_BLOCK) cur[b] = hoff[(size_t)b * a.G + blockIdx.x];
  __syncthreads();
  const uint32_t B = 1u << (a.c - 1);
  const uint32_t mask2 = (1u << a.l2) - 1u;
  const uint32_t lo = blockIdx.x * a.tile;
  const uint32_t hi = min(lo + a.tile, a.n);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += SORT_BLOCK) {
    fr_t s = load<FrP>(canon + 8 * (size_t)i);
    uint32_t carry = 0;
    for (int w = 0; w < a.W; w++) {
      int d = window_digit(s.v, w, a.c, carry);
      if (d != 0) {
        uint32_t key = (uint32_t)(d > 0 ? d : -d) - 1u;
        uint32_t ref = a.ref_offset + i;
        if (a.per_window) key += (uint32_t)w * B;
        else ref += (uint32_t)w * a.ref_stride;
        uint32_t pos = atomicAdd(&cur[key >> a.l2], 1u);
        tmp_key[pos] = key & mask2;
        tmp_ref[pos] = ref | (d < 0 ? NEG_FLAG : 0u);
      }
    }
  }
}

// 4. per top-bit bin: counting sort on the low bits; writes bucket offsets off[]
//    (off[NB] = M) and the references in bucket order.
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_bins(const uint32_t* __restrict__ hoff, SortArgs a, uint32_t NB,
                                                          const uint32_t* __restrict__ tmp_key,
                                                          const uint32_t* __restrict__ tmp_ref,
                                                          uint32_t* __restrict__ refs, uint32_t* __restrict__ off) {
  extern __shared__ uint32_t c2[];   // 2^l2 counters, then scan scratch
  const uint32_t bin = blockIdx.x;
  const uint32_t R = 1u << a.l2;
  const uint32_t lo = hoff[(size_t)bin * a.G];
  const uint32_t hi = hoff[(size_t)(bin + 1) * a.G];   // hoff has D1 * G + 1 entries
  for (uint32_t d = threadIdx.x; d < R; d += SORT_BLOCK) c2[d] = 0;
  __syncthreads();
  for (uint32_t e = lo + threadIdx.x; e < hi; e += SORT_BLOCK) atomicAdd(&c2[tmp_key[e]], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {   // R <= 8192 counters: a serial exclusive scan is short
    uint32_t run = lo;
    for (uint32_t d = 0; d < R; d++) {
      uint32_t v = c2[d];
      c2[d] = run;
      run += v;
    }
    if (bin == 0) off[NB] = hoff[(size_t)a.D1 * a.G];
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < R; d += SORT_BLOCK) {
    const uint32_t b = bin * R + d;
    if (b < NB) off[b] = c2[d];
  }
  __syncthreads();
  for (uint32_t e = lo + threadIdx.x; e < hi; e += SORT_BLOCK) {
    uint32_t pos = atomicAdd(&c2[tmp_key[e]], 1u);
    refs[pos] = tmp_ref[e];
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
// stores the sum for bucket b at partial slot k + b.  Slot k + b is written by
// exactly the chunks k that hold entries of bucket b, so no initialisation is needed.
// The grid covers the largest possible M; M itself is read from off[NB].
__global__ void __launch_bounds__(64) k_msm_accum(const uint32_t* __restrict__ refs,
                                                  const uint32_t* __restrict__ off, uint32_t NB, uint32_t K,
                                                  const uint32_t* __restrict__ bases,
                                                  uint32_t* __restrict__ partials) {
  const uint32_t M = off[NB];
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
  g1_xyzz_t acc = g1_xyzz_t::inf();
  for (uint32_t e = e0; e < e1; e++) {
    if (e >= next) {
      store_xyzz(partials + 48 * (size_t)(k + cb), acc);
      acc = g1_xyzz_t::inf();
      do {
        cb++;
        next = off[cb + 1];
      } while (next <= e);
    }
    const uint32_t r = refs[e];
    g1_aff_t q = load_affine(bases + 24 * (size_t)(r & ~NEG_FLAG));
    if (r & NEG_FLAG) q = g1_neg(q);
    acc = g1_add_affine(acc, q);
  }
  store_xyzz(partials + 48 * (size_t)(k + cb), acc);
}

// Segment t of L buckets of one set (b = t L + i): S_b = sum of the chunk partials
// of bucket b, then running sums from the top: R_t = sum_i S_b and
// A_t = sum_i (i + 1) S_b.  The set's weighted sum is then
//   sum_b (b + 1) S_b = sum_t A_t + L sum_t t R_t.
__global__ void __launch_bounds__(64) k_bucket_seg(const uint32_t* __restrict__ off, uint32_t NB, uint32_t K,
                                                   uint32_t L, const uint32_t* __restrict__ partials,
                                                   uint32_t* __restrict__ A, uint32_t* __restrict__ R) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if ((size_t)t * L >= NB) return;
  g1_xyzz_t run = g1_xyzz_t::inf(), acc = g1_xyzz_t::inf();
  for (uint32_t i = L; i-- > 0;) {
    const uint32_t b = t * L + i;
    const uint32_t lo = off[b], hi = off[b + 1];
    if (hi > lo) {
      const uint32_t k0 = lo / K, k1 = (hi - 1) / K;
      for (uint32_t k = k0; k <= k1; k++) run = g1_add(run, load_xyzz(partials + 48 * (size_t)(k + b)));
    }
    acc = g1_add(acc, run);
  }
  store_xyzz(A + 48 * (size_t)t, acc);
  store_xyzz(R + 48 * (size_t)t, run);
}

// Block-wide sum of one XYZZ point per thread (LDS tree); thread 0 returns it.
__device__ __forceinline__ g1_xyzz_t block_sum(g1_xyzz_t x, uint32_t* lds) {
  const uint32_t t = threadIdx.x;
  store_xyzz(lds + 48 * t, x);
  __syncthreads();
  for (uint32_t d = blockDim.x / 2; d >= 1; d >>= 1) {
    if (t < d) {
      x = g1_add(x, load_xyzz(lds + 48 * (t + d)));
      store_xyzz(lds + 48 * t, x);
    }
    __syncthreads();
  }
  return x;
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
  g1_xyzz_t acc = g1_xyzz_t::inf();
  for (uint32_t t = q * per + threadIdx.x; t < (q + 1) * per; t += CHAN_BLOCK)
    if (h == 0 || ((t >> (h - 1)) & 1u)) acc = g1_add(acc, load_xyzz(src + 48 * ((size_t)s * T + t)));
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) store_xyzz(out + 48 * (size_t)blockIdx.x, acc);
}

// Sums `len` consecutive points per segment; one block per segment.
__global__ void __launch_bounds__(CHAN_BLOCK) k_seg_sum(const uint32_t* __restrict__ in, uint32_t len,
                                                        uint32_t* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  g1_xyzz_t acc = g1_xyzz_t::inf();
  for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
    acc = g1_add(acc, load_xyzz(in + 48 * ((size_t)blockIdx.x * len + i)));
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) store_xyzz(out + 48 * (size_t)blockIdx.x, acc);
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
  int c = c_override > 0 ? c_override : best_c(n, per_window);
  p.c = c;
  p.W = 255 / c + 1;
  p.B = 1u << (c - 1);
  p.per_window = per_window;
  p.sets = per_window ? p.W : 1;
  p.NB = (uint32_t)p.sets * p.B;
  size_t M = (size_t)p.W * n;
  size_t target_threads = (size_t)std::max(num_cus, 1) * 512;
  size_t K = M / target_threads;
  p.K = (uint32_t)std::max<size_t>(2, std::min<size_t>(256, K));
  return p;
}

int msm_run(mzk_ctx* c, const MsmPlan& p, const uint32_t* d_scalars, size_t n, const uint32_t* d_bases,
            uint32_t ref_stride, uint32_t ref_offset, uint64_t out[18]) {
  if (n == 0) {
    memset(out, 0, 18 * 8);
    return MZK_OK;
  }
  const size_t M_max = (size_t)p.W * n;
  MZK_REQUIRE(M_max < 0x7fffffffull, "msm: %zu window digits exceed the 2^31 index range", M_max);
  const uint32_t NB = p.NB;

  // sort geometry: D1 bins of the top key bits (LDS histogram), G tiles of scalars
  SortArgs sa;
  sa.n = (uint32_t)n;
  sa.G = (uint32_t)std::min<size_t>(1024, std::max<size_t>(1, (n + 2047) / 2048));
  sa.tile = (uint32_t)((n + sa.G - 1) / sa.G);
  sa.c = p.c;
  sa.W = p.W;
  sa.per_window = p.per_window;
  const int nb_bits = ceil_log2(NB);
  const int l1 = std::min(nb_bits, 11);
  sa.l2 = (uint32_t)(nb_bits - l1);
  sa.D1 = (uint32_t)((NB + (1u << sa.l2) - 1) >> sa.l2);
  sa.ref_stride = ref_stride;
  sa.ref_offset = ref_offset;
  const size_t HN = (size_t)sa.D1 * sa.G;   // histogram entries
  const size_t tiles = (HN + SCAN_TILE - 1) / SCAN_TILE;
  MZK_REQUIRE(tiles <= (size_t)SCAN_TILE, "msm: sort histogram too large (%zu)", HN);
  const size_t nchunks = (M_max + p.K - 1) / p.K;
  // reduction geometry: segments of L buckets, T per set; channels A + bits of t.
  // One shared set has few, large buckets: a thread per bucket keeps chains short.
  // Many per-window sets have many small buckets: segments of 4 cut the channel work.
  const uint32_t L = std::min<uint32_t>(p.per_window ? 4 : 1, p.B);
  const uint32_t T = p.B / L;
  const uint32_t tbits = (uint32_t)ceil_log2(T);
  const uint32_t nch = 1 + tbits;
  const uint32_t parts = std::max<uint32_t>(1, T / 2048);
  const size_t nchan = (size_t)p.sets * nch;
  const uint32_t lbits = (uint32_t)ceil_log2(L);

  // scratch A: canonical scalars | hist | hoff (HN+1) | tile sums | off (NB+1)
  size_t a_words = 8 * n + HN + (HN + 1) + tiles + (NB + 1) + 64;
  MZK_TRY(c->scratch[SCR_MSM_A].reserve(a_words * 4));
  uint32_t* canon = (uint32_t*)c->scratch[SCR_MSM_A].ptr;
  uint32_t* hist = canon + 8 * n;
  uint32_t* hoff = hist + HN;
  uint32_t* tsums = hoff + HN + 1;
  uint32_t* off = tsums + tiles;
  // scratch B: tmp keys | tmp refs | sorted refs
  MZK_TRY(c->scratch[SCR_MSM_B].reserve(3 * M_max * 4 + 64));
  uint32_t* tmp_key = (uint32_t*)c->scratch[SCR_MSM_B].ptr;
  uint32_t* tmp_ref = tmp_key + M_max;
  uint32_t* refs = tmp_ref + M_max;
  // scratch C: chunk partials (nchunks + NB slots) then segment sums A, R (sets * T each)
  const size_t nseg = (size_t)p.sets * T;
  size_t c_words = (nchunks + NB) * 48 + 2 * nseg * 48;
  MZK_TRY(c->scratch[SCR_MSM_C].reserve(c_words * 4));
  uint32_t* partials = (uint32_t*)c->scratch[SCR_MSM_C].ptr;
  uint32_t* segA = partials + (nchunks + NB) * 48;
  uint32_t* segR = segA + nseg * 48;
  // scratch D: channel partials and channel sums
  size_t d_words = (nchan * parts + nchan) * 48 + 64;
  MZK_TRY(c->scratch[SCR_MSM_D].reserve(d_words * 4));
  uint32_t* chan_parts = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  uint32_t* chan = chan_parts + nchan * parts * 48;

  MZK_LAUNCH(c, "k_sort_hist", k_sort_hist, dim3(sa.G), dim3(SORT_BLOCK), sa.D1 * 4, d_scalars, sa, canon, hist);
  MZK_LAUNCH(c, "k_scan_tiles", k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, hist, hoff, HN, tsums);
  MZK_LAUNCH(c, "k_scan_sums", k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, tsums, (int)tiles, hoff, HN);
  MZK_LAUNCH(c, "k_scan_add", k_scan_add, dim3(grid_for(HN, 256)), dim3(256), 0, hoff, HN, tsums);
  MZK_LAUNCH(c, "k_sort_scatter", k_sort_scatter, dim3(sa.G), dim3(SORT_BLOCK), sa.D1 * 4, canon, sa, hoff, tmp_key,
             tmp_ref);
  MZK_LAUNCH(c, "k_sort_bins", k_sort_bins, dim3(sa.D1), dim3(SORT_BLOCK), (1u << sa.l2) * 4, hoff, sa, NB, tmp_key,
             tmp_ref, refs, off);
  MZK_LAUNCH(c, "k_msm_accum", k_msm_accum, dim3(grid_for(nchunks, 64)), dim3(64), 0, refs, off, NB, p.K, d_bases,
             partials);
  MZK_LAUNCH(c, "k_bucket_seg", k_bucket_seg, dim3(grid_for(nseg, 64)), dim3(64), 0, off, NB, p.K, L, partials, segA,
             segR);
  MZK_LAUNCH(c, "k_chan_sum", k_chan_sum, dim3((unsigned)(nchan * parts)), dim3(CHAN_BLOCK), CHAN_BLOCK * 48 * 4, segA,
             segR, T, nch, parts, chan_parts);
  MZK_LAUNCH(c, "k_seg_sum", k_seg_sum, dim3((unsigned)nchan), dim3(64), 64 * 48 * 4, chan_parts, parts, chan);
  MZK_HIP(hipGetLastError());

  std::vector<uint32_t> h(nchan * 48);
  uint32_t M = 0;
  double h0 = host_ms();
  MZK_HIP(hipMemcpyAsync(h.data(), chan, nchan * 48 * 4, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipMemcpyAsync(&M, off + NB, 4, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  double h1 = host_ms();
  c->prof.host("host_wait", h1 - h0);
  c->last_msm[0] = (uint64_t)p.c;
  c->last_msm[1] = (uint64_t)p.W;
  c->last_msm[2] = (uint64_t)p.sets;
  c->last_msm[3] = p.K;
  c->last_msm[4] = M;
  c->last_msm[5] = n;
  // Set s weighs 2^(c s); inside it, sum_t A_t has weight 1 and U_j (bit j of t)
  // weight 2^(log2 L + j).  Horner over bit positions from the top.
  const int top = (p.sets - 1) * p.c + (int)(lbits + tbits);
  std::vector<int> at(top + 1, -1), at2(top + 1, -1);   // channel index per position (two may share)
  for (int s = 0; s < p.sets; s++) {
    for (uint32_t h2 = 0; h2 < nch; h2++) {
      const int pos = s * p.c + (h2 == 0 ? 0 : (int)(lbits + h2 - 1));
      (at[pos] < 0 ? at[pos] : at2[pos]) = (int)(s * nch + h2);
    }
  }
  host::Xyzz acc = host::Xyzz::inf();
  for (int pos = top; pos >= 0; pos--) {
    acc = host::xyzz_dbl(acc);
    if (at[pos] >= 0) acc = host::xyzz_add(acc, to_host(h.data() + 48 * (size_t)at[pos]));
    if (at2[pos] >= 0) acc = host::xyzz_add(acc, to_host(h.data() + 48 * (size_t)at2[pos]));
  }
  host::xyzz_to_jacobian(acc, out);
  c->prof.host("host_horner", host_ms() - h1);
  return MZK_OK;
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
  return msm_run(c, p, (const uint32_t*)d_coeffs, n, (const uint32_t*)d_bases, 0, 0, out);
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
                 0, out);
}

int mzk_msm_best(mzk_ctx* c, const uint64_t* coeffs, const uint64_t* bases, size_t n, uint64_t out[18]) {
  return msm_host(c, coeffs, bases, n, out);
}

// Zero scalars produce no window digits, so the zero filter of msm_specific
// (kzg/msm.rs:176-181) is implicit in the bucket histogram.
int mzk_msm_specific(mzk_ctx* c, const uint64_t* coeffs, const uint64_t* bases, size_t n, uint64_t out[18]) {
  return msm_host(c, coeffs, bases, n, out);
}

int mzk_msm_last_plan(mzk_ctx* c, uint64_t out[6]) {
  MZK_REQUIRE(c != nullptr && out != nullptr, "mzk_msm_last_plan: null argument");
  std::lock_guard<std::mutex> g(c->mu);
  memcpy(out, c->last_msm, sizeof c->last_msm);
  return MZK_OK;
}

}  // extern "C"
