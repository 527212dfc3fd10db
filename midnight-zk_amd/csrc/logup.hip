// The LogUp lookup argument's vectors on the device (proofs/src/plonk/logup/prover.rs):
//
//   mzk_logup_multiplicities_dev  compute_multiplicities (:411-477): m_i = (how often the
//       table value t_i is looked up by a selected input row) / (how often t_i appears in
//       the table), over the usable rows; the blinding rows take the caller's values.
//       Counting goes through an open-addressing hash table in HBM keyed by the table row
//       that first claimed a slot (its value stays in the table, so a slot needs no key
//       copy): insert every usable table row (claim by compare-and-swap, count), then look
//       up every selected input row (count; a miss is the reference's panic, reported as
//       the first failing column and row), then one pass writes numerators and
//       denominators, and the denominators are batch-inverted.
//   mzk_logup_logderivative_dev   compute_logderivative (:189-310): every 1/(f_j + beta)
//       and 1/(t + beta) in one batch inversion, the helper columns h_chunk = sum over the
//       chunk's inputs, the per-row log-derivative sel h - m / (t + beta), and the
//       aggregator Z = [0, running sums ...] up to the blinding rows, then the blinding
//       values.
//
// Field values are compared as their Montgomery words (reduced values: equal elements
// have equal words).
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "ctx.hpp"
#include "field.hpp"

namespace mzk {

static constexpr int LOGUP_MAX_INPUTS = 32;

struct InputPtrs {
  const uint32_t* p[LOGUP_MAX_INPUTS];
};

struct LogupWords {
  uint32_t v[8];
};

__device__ __forceinline__ void load8(const uint32_t* __restrict__ p, uint32_t w[8]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  const uint4 a = q[0], b = q[1];
  w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
  w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
}

__device__ __forceinline__ bool eq8(const uint32_t* __restrict__ p, const uint32_t w[8]) {
  uint32_t x[8];
  load8(p, x);
  uint32_t d = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d |= x[i] ^ w[i];
  return d == 0;
}

__device__ __forceinline__ uint32_t hash8(const uint32_t w[8]) {
  uint32_t h = 0x811C9DC5u;   // FNV-1a over the words, then a murmur3 finaliser
#pragma unroll
  for (int i = 0; i < 8; i++) {
    h ^= w[i];
    h *= 0x01000193u;
  }
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

static constexpr uint32_t SLOT_NONE = 0xffffffffu;

// The slot holding value w, or SLOT_NONE (inserts have completed: plain loads).
__device__ __forceinline__ uint32_t find_slot(const uint32_t w[8], const uint32_t* __restrict__ table,
                                              const uint32_t* __restrict__ owner, uint32_t mask) {
  uint32_t p = hash8(w) & mask;
  for (uint32_t probe = 0; probe <= mask; probe++) {
    const uint32_t o = owner[p];
    if (o == 0) return SLOT_NONE;
    if (eq8(table + 8 * (size_t)(o - 1), w)) return p;
    p = (p + 1) & mask;
  }
  return SLOT_NONE;
}

__global__ void k_logup_clear(uint32_t* __restrict__ words, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) words[i] = 0;
}

// Every usable table row: find or claim the slot of its value, count it.
__global__ void k_logup_insert(const uint32_t* __restrict__ table, size_t usable, uint32_t* owner,
                               uint32_t* tcount, uint32_t mask) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= usable) return;
  uint32_t w[8];
  load8(table + 8 * i, w);
  uint32_t p = hash8(w) & mask;
  for (uint32_t probe = 0; probe <= mask; probe++) {   // the table is at most half full
    const uint32_t o = atomicCAS(&owner[p], 0u, (uint32_t)i + 1);
    if (o == 0 || eq8(table + 8 * (size_t)(o - 1), w)) {
      atomicAdd(&tcount[p], 1u);
      return;
    }
    p = (p + 1) & mask;
  }
}

// Every selected input row of every column: count its value's slot; a value missing from
// the table records (column << 32 | row), the smallest one wins.
__global__ void k_logup_count(InputPtrs in, uint32_t ninputs, const uint32_t* __restrict__ sel,
                              const uint32_t* __restrict__ table, size_t usable, const uint32_t* __restrict__ owner,
                              uint32_t* icount, uint32_t mask, unsigned long long* miss) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= usable * ninputs) return;
  const uint32_t c = (uint32_t)(g / usable);
  const size_t i = g % usable;
  uint32_t s[8];
  load8(sel + 8 * i, s);
  if ((s[0] | s[1] | s[2] | s[3] | s[4] | s[5] | s[6] | s[7]) == 0) return;
  uint32_t w[8];
  load8(in.p[c] + 8 * i, w);
  const uint32_t p = find_slot(w, table, owner, mask);
  if (p == SLOT_NONE) {
    atomicMin(miss, ((unsigned long long)c << 32) | (unsigned long long)i);
    return;
  }
  atomicAdd(&icount[p], 1u);
}

// k in Montgomery form: k R2 R^-1 = k R
__device__ __forceinline__ fr_t fr_small(uint32_t k) {
  constexpr uint32_t R2[8] = {0xf3f29c6du, 0xc999e990u, 0x87925c23u, 0x2b6cedcbu,
                              0x7254398fu, 0x05d31496u, 0x9f59ff11u, 0x0748d9d9u};
  fr_t x = fr_t::zero(), r2;
  x.v[0] = k;
#pragma unroll
  for (int j = 0; j < 8; j++) r2.v[j] = R2[j];
  return mul(x, r2);
}

// Numerators (input counts) and denominators (table counts) of the usable rows; the
// blinding rows get the caller's values over 1.
__global__ void k_logup_finalize(const uint32_t* __restrict__ table, size_t usable, size_t n,
                                 const uint32_t* __restrict__ owner, const uint32_t* __restrict__ tcount,
                                 const uint32_t* __restrict__ icount, uint32_t mask,
                                 const uint32_t* __restrict__ blind, uint32_t* __restrict__ num,
                                 uint32_t* __restrict__ den) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i < usable) {
    uint32_t w[8];
    load8(table + 8 * i, w);
    const uint32_t p = find_slot(w, table, owner, mask);   // present: the row itself inserted it
    store<FrP>(num + 8 * i, fr_small(icount[p]));
    store<FrP>(den + 8 * i, fr_small(tcount[p]));
  } else {
    store<FrP>(num + 8 * i, load<FrP>(blind + 8 * (i - usable)));
    store<FrP>(den + 8 * i, fr_t::one());
  }
}

// den[c n + i] = (c < ninputs ? input c : table)[i] + beta
__global__ void k_logup_denoms(InputPtrs in, uint32_t ninputs, const uint32_t* __restrict__ table, size_t n,
                               LogupWords beta, uint32_t* __restrict__ den) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (size_t)(ninputs + 1) * n) return;
  const uint32_t c = (uint32_t)(g / n);
  const size_t i = g % n;
  const uint32_t* src = c < ninputs ? in.p[c] : table;
  fr_t b;
#pragma unroll
  for (int j = 0; j < 8; j++) b.v[j] = beta.v[j];
  store<FrP>(den + 8 * g, add(load<FrP>(src + 8 * i), b));
}

struct HelperOut {
  uint32_t* h[LOGUP_MAX_INPUTS];
  uint32_t chunk_end[LOGUP_MAX_INPUTS];   // inputs [chunk_end[k-1], chunk_end[k]) form chunk k
};

// helpers[k][i] = sum over chunk k of 1/(f_c[i] + beta); logd[i] = sel[i] sum_k helpers[k][i]
// - m[i] / (t[i] + beta)  (logup/prover.rs:225-267)
__global__ void k_logup_rows(HelperOut H, uint32_t nchunks, uint32_t ninputs, const uint32_t* __restrict__ inv,
                             const uint32_t* __restrict__ sel, const uint32_t* __restrict__ m, size_t n,
                             uint32_t* __restrict__ logd) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t total = fr_t::zero();
  uint32_t c = 0;
  for (uint32_t k = 0; k < nchunks; k++) {
    fr_t h = fr_t::zero();
    for (; c < H.chunk_end[k]; c++) h = add(h, load<FrP>(inv + 8 * ((size_t)c * n + i)));
    store<FrP>(H.h[k] + 8 * i, h);
    total = add(total, h);
  }
  const fr_t t_inv = load<FrP>(inv + 8 * ((size_t)ninputs * n + i));
  store<FrP>(logd + 8 * i, sub(mul(load<FrP>(sel + 8 * i), total), mul(load<FrP>(m + 8 * i), t_inv)));
}

// Z[0] = 0 and the blinding rows Z[n - bf + t] = blind[t]
__global__ void k_logup_z_ends(uint32_t* __restrict__ z, size_t n, uint32_t bf, const uint32_t* __restrict__ blind) {
  const uint32_t t = threadIdx.x;
  if (t == 0) store<FrP>(z, fr_t::zero());
  for (uint32_t j = t; j < bf; j += blockDim.x) store<FrP>(z + 8 * (n - bf + j), load<FrP>(blind + 8 * j));
}

static bool overlaps(const void* p, size_t pb, const void* q, size_t qb) {
  const uintptr_t a = (uintptr_t)p, b = (uintptr_t)q;
  return p && q && a < b + qb && b < a + pb;
}

static uint32_t pow2_at_least(size_t x) {
  uint32_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

}  // namespace mzk

using namespace mzk;

extern "C" {

int mzk_logup_multiplicities_dev(mzk_ctx* c, const void* selector, const void* const* inputs, size_t ninputs,
                                 const void* table, size_t n, size_t usable_rows, const uint64_t* blinding,
                                 void* d_m) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_REQUIRE(selector && table && d_m && (inputs || ninputs == 0), "mzk_logup_multiplicities_dev: null argument");
  MZK_REQUIRE(ninputs <= (size_t)LOGUP_MAX_INPUTS, "mzk_logup_multiplicities_dev: %zu inputs above %d", ninputs,
              LOGUP_MAX_INPUTS);
  MZK_REQUIRE(n >= 1 && n < (1ull << 30), "mzk_logup_multiplicities_dev: n = %zu out of range", n);
  MZK_REQUIRE(usable_rows <= n, "mzk_logup_multiplicities_dev: usable rows above n");
  MZK_REQUIRE(usable_rows == n || blinding, "mzk_logup_multiplicities_dev: blinding values missing");
  InputPtrs in;
  memset(&in, 0, sizeof in);
  for (size_t j = 0; j < ninputs; j++) {
    MZK_REQUIRE(inputs[j] != nullptr, "mzk_logup_multiplicities_dev: input %zu missing", j);
    MZK_REQUIRE(!overlaps(d_m, n * 32, inputs[j], n * 32), "mzk_logup_multiplicities_dev: output overlaps input %zu",
                j);
    in.p[j] = (const uint32_t*)inputs[j];
  }
  MZK_REQUIRE(!overlaps(d_m, n * 32, table, n * 32) && !overlaps(d_m, n * 32, selector, n * 32),
              "mzk_logup_multiplicities_dev: output overlaps the table or the selector");
  const uint32_t cap = pow2_at_least(std::max<size_t>(2 * usable_rows, 64));
  const uint32_t mask = cap - 1;
  const size_t bl = n - usable_rows;
  uint32_t *owner, *tcount, *icount, *blind, *num, *den;
  unsigned long long* miss;
  unsigned long long miss_h = ~0ull;
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipSetDevice(c->device));
    MZK_TRY(c->scratch[SCR_MSM_A].reserve((size_t)3 * cap * 4 + bl * 32 + 64));
    MZK_TRY(c->scratch[SCR_MSM_B].reserve(n * 32 + 64));
    owner = (uint32_t*)c->scratch[SCR_MSM_A].ptr;
    tcount = owner + cap;
    icount = tcount + cap;
    miss = (unsigned long long*)(icount + cap);
    blind = (uint32_t*)(miss + 2);   // 16-byte aligned
    num = (uint32_t*)d_m;            // numerators in place, denominators in scratch
    den = (uint32_t*)c->scratch[SCR_MSM_B].ptr;
    hipStream_t st = c->stream;
    MZK_LAUNCH(c, "k_logup_clear", k_logup_clear, dim3(grid_for((size_t)3 * cap, 256)), dim3(256), 0, owner,
               (size_t)3 * cap);
    MZK_HIP(hipMemcpyAsync(miss, &miss_h, 8, hipMemcpyHostToDevice, st));
    if (bl) MZK_HIP(hipMemcpyAsync(blind, blinding, bl * 32, hipMemcpyHostToDevice, st));
    if (usable_rows) {
      MZK_LAUNCH(c, "k_logup_insert", k_logup_insert, dim3(grid_for(usable_rows, 256)), dim3(256), 0,
                 (const uint32_t*)table, usable_rows, owner, tcount, mask);
      if (ninputs)
        MZK_LAUNCH(c, "k_logup_count", k_logup_count, dim3(grid_for(usable_rows * ninputs, 256)), dim3(256), 0, in,
                   (uint32_t)ninputs, (const uint32_t*)selector, (const uint32_t*)table, usable_rows, owner, icount,
                   mask, miss);
    }
    MZK_HIP(hipGetLastError());
    MZK_HIP(hipMemcpyAsync(&miss_h, miss, 8, hipMemcpyDeviceToHost, st));
    MZK_HIP(hipStreamSynchronize(st));
    if (miss_h != ~0ull) {
      set_error("mzk_logup_multiplicities_dev: input %llu row %llu: value not found in the lookup table",
                miss_h >> 32, miss_h & 0xffffffffull);
      return MZK_ERR_INVALID;
    }
    MZK_LAUNCH(c, "k_logup_finalize", k_logup_finalize, dim3(grid_for(n, 256)), dim3(256), 0,
               (const uint32_t*)table, usable_rows, n, owner, tcount, icount, mask, blind, num, den);
    MZK_HIP(hipGetLastError());
  }
  // m = input count / table count (logup/prover.rs:451-476)
  MZK_TRY(mzk_fr_batch_invert(c, den, n));
  return mzk_fr_vec_op(c, 2, num, den, num, n);
}

int mzk_logup_logderivative_dev(mzk_ctx* c, const mzk_logup_args* a) {
  MZK_REQUIRE(c != nullptr && a != nullptr, "mzk_logup_logderivative_dev: null argument");
  const size_t n = a->n, bf = a->blinding_factors;
  MZK_REQUIRE(n >= 2 && n < (1ull << 30) && bf + 1 <= n, "mzk_logup_logderivative_dev: n = %zu, blinding %zu", n, bf);
  MZK_REQUIRE(a->selector && a->multiplicities && a->table && a->aggregator && (bf == 0 || a->blinding),
              "mzk_logup_logderivative_dev: null vector");
  MZK_REQUIRE(a->ninputs <= (size_t)LOGUP_MAX_INPUTS && a->nchunks <= (size_t)LOGUP_MAX_INPUTS &&
                  (a->ninputs == 0 || (a->inputs && a->chunk_sizes && a->helpers)),
              "mzk_logup_logderivative_dev: inputs / chunks out of range");
  InputPtrs in;
  HelperOut H;
  memset(&in, 0, sizeof in);
  memset(&H, 0, sizeof H);
  size_t end = 0;
  for (size_t k = 0; k < a->nchunks; k++) {
    MZK_REQUIRE(a->helpers[k] != nullptr, "mzk_logup_logderivative_dev: helper %zu missing", k);
    end += a->chunk_sizes[k];
    H.h[k] = (uint32_t*)a->helpers[k];
    H.chunk_end[k] = (uint32_t)end;
  }
  MZK_REQUIRE(end == a->ninputs, "mzk_logup_logderivative_dev: chunk sizes sum to %zu, not %zu inputs", end,
              a->ninputs);
  for (size_t j = 0; j < a->ninputs; j++) {
    MZK_REQUIRE(a->inputs[j] != nullptr, "mzk_logup_logderivative_dev: input %zu missing", j);
    in.p[j] = (const uint32_t*)a->inputs[j];
  }
  const size_t ni = a->ninputs;
  // outputs are written while the inputs are still read: no output may overlap an input
  std::vector<const void*> ins = {a->selector, a->multiplicities, a->table};
  for (size_t j = 0; j < ni; j++) ins.push_back(a->inputs[j]);
  std::vector<const void*> outs = {a->aggregator};
  for (size_t k = 0; k < a->nchunks; k++) outs.push_back(a->helpers[k]);
  for (size_t o = 0; o < outs.size(); o++) {
    for (const void* q : ins)
      MZK_REQUIRE(!overlaps(outs[o], n * 32, q, n * 32), "mzk_logup_logderivative_dev: an output overlaps an input");
    for (size_t o2 = o + 1; o2 < outs.size(); o2++)
      MZK_REQUIRE(!overlaps(outs[o], n * 32, outs[o2], n * 32), "mzk_logup_logderivative_dev: two outputs overlap");
  }
  uint32_t *den, *logd, *blind;
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipSetDevice(c->device));
    MZK_TRY(c->scratch[SCR_MSM_B].reserve((ni + 2) * n * 32 + 64));
    MZK_TRY(c->scratch[SCR_MSM_A].reserve(bf * 32 + 64));
    den = (uint32_t*)c->scratch[SCR_MSM_B].ptr;
    logd = den + 8 * (ni + 1) * n;
    blind = (uint32_t*)c->scratch[SCR_MSM_A].ptr;
    if (bf) MZK_HIP(hipMemcpyAsync(blind, a->blinding, bf * 32, hipMemcpyHostToDevice, c->stream));
    LogupWords beta;
    memcpy(beta.v, a->beta, 32);
    MZK_LAUNCH(c, "k_logup_denoms", k_logup_denoms, dim3(grid_for((ni + 1) * n, 256)), dim3(256), 0, in,
               (uint32_t)ni, (const uint32_t*)a->table, n, beta, den);
    MZK_HIP(hipGetLastError());
  }
  // 1/(f_j + beta) and 1/(t + beta) in one batch inversion (zeros stay zero, as BatchInvert)
  MZK_TRY(mzk_fr_batch_invert(c, den, (ni + 1) * n));
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_LAUNCH(c, "k_logup_rows", k_logup_rows, dim3(grid_for(n, 256)), dim3(256), 0, H, (uint32_t)a->nchunks,
               (uint32_t)ni, (const uint32_t*)den, (const uint32_t*)a->selector, (const uint32_t*)a->multiplicities, n,
               logd);
    MZK_LAUNCH(c, "k_logup_z_ends", k_logup_z_ends, dim3(1), dim3(64), 0, (uint32_t*)a->aggregator, n, (uint32_t)bf,
               (const uint32_t*)blind);
    MZK_HIP(hipGetLastError());
  }
  // Z[k] = logd[0] + ... + logd[k-1] for 1 <= k < n - bf (logup/prover.rs:269-278)
  return mzk_fr_prefix_sum(c, logd, (uint32_t*)a->aggregator + 8, n - bf - 1);
}

}  // extern "C"
