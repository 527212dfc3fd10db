// Pippenger multi-scalar multiplication on G1 for gfx950.
//
// Replaces curves/src/msm.rs:460-532 (msm_best) and proofs/src/poly/kzg/msm.rs:174-201
// (msm_specific).  The result is the unique point sum_i s_i P_i; the reference's
// Booth digits and batch-affine buckets are a CPU schedule, re-planned here for the
// GPU:
//
//  1. k_msm_prep      scalars -> canonical form, signed c-bit window digits, bucket
//                     histogram (one thread per scalar)
//  2. scan            exclusive prefix sum of the histogram = bucket offsets
//  3. k_msm_scatter   counting sort: (point, sign) references grouped by bucket
//  4. k_msm_accum     each thread adds a fixed chunk of K sorted references into
//                     XYZZ partial bucket sums (mixed additions, 8M + 2S); equal load
//                     per thread whatever the bucket sizes
//  5. k_bucket_sum    per bucket, the sum of its (few) chunk partials
//  6. k_wsum          per block of 256 buckets, sum_j (j+1) S_j and sum_j S_j via a
//                     suffix scan in LDS; applied recursively to the block outputs
//  7. host tail       a few dozen point operations (serial; one host core runs them
//                     ~50x faster than one GPU wave) and the window Horner step.
//
// Two bucket layouts share the kernels:
//  * per-window sets (ad-hoc bases): bucket id = w * B + |d| - 1, one weighted
//    sum per window, combined by sum_w 2^(cw) S_w on the host;
//  * one shared set (resident SRS with window-shifted copies 2^(cw) g_i, built
//    once at registration): bucket id = |d| - 1, reference = w * N + i, and no
//    window combination at all.
#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "g1.hpp"
#include "host_ec.hpp"
#include "msm.hpp"

namespace mzk {

static constexpr uint32_t NEG_FLAG = 0x80000000u;
static constexpr int WS_BLOCK = 256;

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

__global__ void k_msm_prep(const uint32_t* __restrict__ scalars, size_t n, int c, int W, int per_window,
                           uint32_t* __restrict__ canon, uint32_t* __restrict__ counts) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t s = from_mont(load<FrP>(scalars + 8 * i));
  store<FrP>(canon + 8 * i, s);
  const uint32_t B = 1u << (c - 1);
  uint32_t carry = 0;
  for (int w = 0; w < W; w++) {
    int d = window_digit(s.v, w, c, carry);
    if (d != 0) {
      uint32_t b = (uint32_t)(d > 0 ? d : -d) - 1u;
      if (per_window) b += (uint32_t)w * B;
      atomicAdd(&counts[b], 1u);
    }
  }
}

__global__ void k_msm_scatter(const uint32_t* __restrict__ canon, size_t n, int c, int W, int per_window,
                              uint32_t ref_stride, uint32_t ref_offset, uint32_t* __restrict__ cursor,
                              uint32_t* __restrict__ refs) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t s = load<FrP>(canon + 8 * i);
  const uint32_t B = 1u << (c - 1);
  uint32_t carry = 0;
  for (int w = 0; w < W; w++) {
    int d = window_digit(s.v, w, c, carry);
    if (d != 0) {
      uint32_t b = (uint32_t)(d > 0 ? d : -d) - 1u;
      uint32_t ref = ref_offset + (uint32_t)i;
      if (per_window) b += (uint32_t)w * B;
      else ref += (uint32_t)w * ref_stride;
      uint32_t pos = atomicAdd(&cursor[b], 1u);
      refs[pos] = ref | (d < 0 ? NEG_FLAG : 0u);
    }
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

// Phase 3: add tile prefixes; also initialise the scatter cursors.
__global__ void k_scan_add(uint32_t* __restrict__ out, size_t n, const uint32_t* __restrict__ tile_sums,
                           uint32_t* __restrict__ cursor) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t v = out[i] + tile_sums[i / SCAN_TILE];
  out[i] = v;
  cursor[i] = v;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// --------------------------------------------------------------------- accumulation
// Thread k adds refs[kK, (k+1)K) into running XYZZ sums, one per bucket run, and
// stores the sum for bucket b at partial slot k + b.  Slot k + b is written by
// exactly the chunks k that hold entries of bucket b, so no initialisation is needed.
__global__ void __launch_bounds__(64) k_msm_accum(const uint32_t* __restrict__ refs, uint32_t M,
                                                  const uint32_t* __restrict__ off, uint32_t NB, uint32_t K,
                                                  const uint32_t* __restrict__ bases,
                                                  uint32_t* __restrict__ partials) {
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

// S_b = sum of the chunk partials of bucket b (identity for empty buckets).
__global__ void __launch_bounds__(64) k_bucket_sum(const uint32_t* __restrict__ off, uint32_t NB, uint32_t K,
                                                   const uint32_t* __restrict__ partials,
                                                   uint32_t* __restrict__ sums) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= NB) return;
  const uint32_t lo = off[b], hi = off[b + 1];
  g1_xyzz_t acc = g1_xyzz_t::inf();
  if (hi > lo) {
    const uint32_t k0 = lo / K, k1 = (hi - 1) / K;
    acc = load_xyzz(partials + 48 * (size_t)(k0 + b));
    for (uint32_t k = k0 + 1; k <= k1; k++) acc = g1_add(acc, load_xyzz(partials + 48 * (size_t)(k + b)));
  }
  store_xyzz(sums + 48 * (size_t)b, acc);
}

// Per block of WS_BLOCK consecutive items X_j (j local) of one set:
//   P = sum_j (j+1) X_j = sum_j suffix(j),  Q = sum_j X_j = suffix(0),
// with suffix(j) = sum_{i>=j} X_i from a Hillis-Steele scan in LDS.
// Items per set: `per_set`; blocks per set: `gps`; outputs P[set*gps + g], Q[...].
__global__ void __launch_bounds__(WS_BLOCK) k_wsum(const uint32_t* __restrict__ items, uint32_t per_set,
                                                   uint32_t gps, uint32_t* __restrict__ P,
                                                   uint32_t* __restrict__ Q) {
  extern __shared__ uint32_t lds[];   // WS_BLOCK * 48 words
  const uint32_t set = blockIdx.x / gps, g = blockIdx.x % gps;
  const uint32_t j = threadIdx.x;
  const uint32_t idx = g * WS_BLOCK + j;
  g1_xyzz_t x = idx < per_set ? load_xyzz(items + 48 * ((size_t)set * per_set + idx)) : g1_xyzz_t::inf();
  store_xyzz(lds + 48 * j, x);
  __syncthreads();
  for (uint32_t d = 1; d < WS_BLOCK; d <<= 1) {
    g1_xyzz_t y = j + d < WS_BLOCK ? load_xyzz(lds + 48 * (j + d)) : g1_xyzz_t::inf();
    __syncthreads();
    x = g1_add(x, y);
    store_xyzz(lds + 48 * j, x);
    __syncthreads();
  }
  // x = suffix(j); thread 0 holds Q
  if (j == 0) store_xyzz(Q + 48 * (size_t)blockIdx.x, x);
  // tree sum of the suffixes
  for (uint32_t d = WS_BLOCK / 2; d >= 1; d >>= 1) {
    if (j < d) {
      x = g1_add(x, load_xyzz(lds + 48 * (j + d)));
      store_xyzz(lds + 48 * j, x);
    }
    __syncthreads();
  }
  if (j == 0) store_xyzz(P + 48 * (size_t)blockIdx.x, x);
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

static host::Xyzz host_neg(host::Xyzz p) {
  if (!p.is_inf()) p.Y = host::fp_sub(host::fp_zero(), p.Y);
  return p;
}

static host::Xyzz host_scale_pow2(host::Xyzz p, int k) {
  for (int i = 0; i < k; i++) p = host::xyzz_dbl(p);
  return p;
}

MsmPlan msm_plan(size_t n, int per_window, int num_cus, int c_override) {
  MsmPlan p;
  int lg = 0;
  while (((size_t)1 << (lg + 1)) <= n) lg++;
  int c = c_override > 0 ? c_override : std::max(4, std::min(16, lg - 4));
  p.c = c;
  p.W = 255 / c + 1;
  p.B = 1u << (c - 1);
  p.per_window = per_window;
  p.sets = per_window ? p.W : 1;
  p.NB = (uint32_t)p.sets * p.B;
  size_t M = (size_t)p.W * n;
  size_t target_threads = (size_t)std::max(num_cus, 1) * 512;
  size_t K = M / target_threads;
  p.K = (uint32_t)std::max<size_t>(2, std::min<size_t>(64, K));
  return p;
}

// Weighted sum per set over device items (sets x per_set XYZZ), returned to the host:
// out_ws[s] = sum_j (j+1) X_{s,j}, out_sum[s] = sum_j X_{s,j}.
static int wsum_sets(mzk_ctx* c, const uint32_t* items, uint32_t sets, uint32_t per_set, uint32_t* work,
                     size_t work_words, std::vector<host::Xyzz>& out_ws, std::vector<host::Xyzz>& out_sum) {
  const uint32_t gps = (per_set + WS_BLOCK - 1) / WS_BLOCK;
  const size_t blocks = (size_t)sets * gps;
  MZK_REQUIRE(4 * blocks * 48 <= work_words, "msm: reduction workspace too small");
  uint32_t* P = work;
  uint32_t* Q = work + blocks * 48;
  MZK_LAUNCH(c, "k_wsum", k_wsum, dim3((unsigned)blocks), dim3(WS_BLOCK), WS_BLOCK * 48 * 4, items,
                     per_set, gps, P, Q);
  MZK_HIP(hipGetLastError());
  out_ws.assign(sets, host::Xyzz::inf());
  out_sum.assign(sets, host::Xyzz::inf());
  if (gps == 1) {
    std::vector<uint32_t> h(2 * blocks * 48);
    MZK_HIP(hipMemcpyAsync(h.data(), P, 2 * blocks * 48 * 4, hipMemcpyDeviceToHost, c->stream));
    MZK_HIP(hipStreamSynchronize(c->stream));
    for (uint32_t s = 0; s < sets; s++) {
      out_ws[s] = to_host(h.data() + 48 * s);
      out_sum[s] = to_host(h.data() + 48 * (blocks + s));
    }
    return MZK_OK;
  }
  // WS(X) = sum_g P_g + m * (WS(Q) - sum_g Q_g), m = WS_BLOCK; sum_g P_g = SUM of the P array.
  std::vector<host::Xyzz> wsq, sumq, wsp, sump;
  uint32_t* next = work + 2 * blocks * 48;
  MZK_TRY(wsum_sets(c, Q, sets, gps, next, work_words - 2 * blocks * 48, wsq, sumq));
  MZK_TRY(wsum_sets(c, P, sets, gps, next, work_words - 2 * blocks * 48, wsp, sump));
  int shift = 0;
  while ((1 << shift) < WS_BLOCK) shift++;
  for (uint32_t s = 0; s < sets; s++) {
    host::Xyzz t = host::xyzz_add(wsq[s], host_neg(sumq[s]));
    t = host_scale_pow2(t, shift);
    out_ws[s] = host::xyzz_add(sump[s], t);
    out_sum[s] = sumq[s];
  }
  return MZK_OK;
}

int msm_run(mzk_ctx* c, const MsmPlan& p, const uint32_t* d_scalars, size_t n, const uint32_t* d_bases,
            uint32_t ref_stride, uint32_t ref_offset, uint64_t out[18]) {
  if (n == 0) {
    memset(out, 0, 18 * 8);
    return MZK_OK;
  }
  const size_t M_max = (size_t)p.W * n;
  MZK_REQUIRE(M_max < 0x7fffffffull, "msm: %zu window digits exceed the 2^31 index range", M_max);
  const size_t NB = p.NB;
  const size_t tiles = (NB + SCAN_TILE - 1) / SCAN_TILE;
  MZK_REQUIRE(tiles <= (size_t)SCAN_TILE, "msm: too many buckets (%zu)", NB);
  const size_t nchunks = (M_max + p.K - 1) / p.K;

  // scratch A: canonical scalars | counts | off | cursor | tile sums
  size_t a_words = 8 * n + NB + (NB + 1) + NB + tiles + 64;
  MZK_TRY(c->scratch[SCR_MSM_A].reserve(a_words * 4));
  uint32_t* canon = (uint32_t*)c->scratch[SCR_MSM_A].ptr;
  uint32_t* counts = canon + 8 * n;
  uint32_t* off = counts + NB;
  uint32_t* cursor = off + NB + 1;
  uint32_t* tsums = cursor + NB;
  // scratch B: sorted references
  MZK_TRY(c->scratch[SCR_MSM_B].reserve(M_max * 4 + 64));
  uint32_t* refs = (uint32_t*)c->scratch[SCR_MSM_B].ptr;
  // scratch C: chunk partials (nchunks + NB slots) then bucket sums (NB)
  size_t c_words = (nchunks + NB) * 48 + NB * 48;
  MZK_TRY(c->scratch[SCR_MSM_C].reserve(c_words * 4));
  uint32_t* partials = (uint32_t*)c->scratch[SCR_MSM_C].ptr;
  uint32_t* sums = partials + (nchunks + NB) * 48;
  // scratch D: weighted-sum pyramid
  size_t blocks1 = (size_t)p.sets * ((p.B + WS_BLOCK - 1) / WS_BLOCK);
  size_t d_words = 8 * (blocks1 + 64) * 48;
  MZK_TRY(c->scratch[SCR_MSM_D].reserve(d_words * 4));
  uint32_t* work = (uint32_t*)c->scratch[SCR_MSM_D].ptr;

  hipStream_t st = c->stream;
  MZK_LAUNCH(c, "k_fill_u32", k_fill_u32, dim3(grid_for(NB, 256)), dim3(256), 0, counts, NB, 0u);
  MZK_LAUNCH(c, "k_msm_prep", k_msm_prep, dim3(grid_for(n, 128)), dim3(128), 0, d_scalars, n, p.c, p.W, p.per_window,
                     canon, counts);
  MZK_LAUNCH(c, "k_scan_tiles", k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, counts, off, NB, tsums);
  MZK_LAUNCH(c, "k_scan_sums", k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, tsums, (int)tiles, off, NB);
  MZK_LAUNCH(c, "k_scan_add", k_scan_add, dim3(grid_for(NB, 256)), dim3(256), 0, off, NB, tsums, cursor);
  MZK_LAUNCH(c, "k_msm_scatter", k_msm_scatter, dim3(grid_for(n, 128)), dim3(128), 0, canon, n, p.c, p.W, p.per_window,
                     ref_stride, ref_offset, cursor, refs);
  MZK_HIP(hipGetLastError());
  uint32_t M = 0;
  MZK_HIP(hipMemcpyAsync(&M, off + NB, 4, hipMemcpyDeviceToHost, st));
  MZK_HIP(hipStreamSynchronize(st));
  if (M == 0) {
    memset(out, 0, 18 * 8);
    return MZK_OK;
  }
  const uint32_t chunks = (M + p.K - 1) / p.K;
  c->last_msm[0] = (uint64_t)p.c;
  c->last_msm[1] = (uint64_t)p.W;
  c->last_msm[2] = (uint64_t)p.sets;
  c->last_msm[3] = p.K;
  c->last_msm[4] = M;
  c->last_msm[5] = n;
  MZK_LAUNCH(c, "k_msm_accum", k_msm_accum, dim3(grid_for(chunks, 64)), dim3(64), 0, refs, M, off, (uint32_t)NB, p.K,
                     d_bases, partials);
  MZK_LAUNCH(c, "k_bucket_sum", k_bucket_sum, dim3(grid_for(NB, 64)), dim3(64), 0, off, (uint32_t)NB, p.K, partials,
                     sums);
  MZK_HIP(hipGetLastError());
  std::vector<host::Xyzz> ws, tot;
  MZK_TRY(wsum_sets(c, sums, (uint32_t)p.sets, p.B, work, d_words, ws, tot));
  host::Xyzz acc = host::Xyzz::inf();
  for (int s = p.sets - 1; s >= 0; s--) {
    if (s != p.sets - 1) acc = host_scale_pow2(acc, p.c);
    acc = host::xyzz_add(acc, ws[s]);
  }
  host::xyzz_to_jacobian(acc, out);
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

}  // extern "C"

extern "C" int mzk_msm_last_plan(mzk_ctx* c, uint64_t out[6]) {
  MZK_REQUIRE(c != nullptr && out != nullptr, "mzk_msm_last_plan: null argument");
  std::lock_guard<std::mutex> g(c->mu);
  memcpy(out, c->last_msm, sizeof c->last_msm);
  return MZK_OK;
}
