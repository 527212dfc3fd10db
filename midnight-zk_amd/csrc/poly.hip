// Scalar-field vector kernels the prover needs around its MSMs and NTTs:
// element-wise arithmetic (Polynomial ops, proofs/src/poly/mod.rs), batch inversion
// (ff BatchInvert, used by the permutation argument), prefix products (the grand
// product z, permutation/prover.rs:177-215), Horner evaluation (eval_polynomial,
// proofs/src/utils/arithmetic.rs:59-87), Kate division (arithmetic.rs:106-125) and
// powers of a point.  Scans and the Kate recurrence are parallel three-phase scans.
#include <vector>

#include "ctx.hpp"
#include "field.hpp"
#include "host_ec.hpp"

namespace mzk {

struct FrArg {
  uint32_t v[8];
};

__device__ __forceinline__ fr_t arg_fr(const FrArg& a) {
  fr_t r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = a.v[i];
  return r;
}

static FrArg to_arg(const uint64_t s[4]) {
  FrArg a;
  memcpy(a.v, s, 32);
  return a;
}

static FrArg to_arg(const host::Fr& f) {
  FrArg a;
  memcpy(a.v, f.l, 32);
  return a;
}

__global__ void k_vec_binop(int op, const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                            uint32_t* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t x = load<FrP>(a + 8 * i), y = load<FrP>(b + 8 * i);
  fr_t r = op == 0 ? add(x, y) : op == 1 ? sub(x, y) : mul(x, y);
  store<FrP>(out + 8 * i, r);
}

// Delta encodings of a Lagrange vector (poly/mod.rs:290-335): order 1, b_0 = a_0 and
// b_i = a_i - a_(i-1); order 2, c_0 = a_0, c_1 = a_1 - 2 a_0, c_i = a_i - 2 a_(i-1) + a_(i-2).
__global__ void k_to_delta(const uint32_t* __restrict__ a, uint32_t* __restrict__ out, size_t n, int order) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t r = load<FrP>(a + 8 * i);
  if (i >= 1) {
    const fr_t p = load<FrP>(a + 8 * (i - 1));
    r = sub(r, p);
    if (order == 2) {
      r = sub(r, p);
      if (i >= 2) r = add(r, load<FrP>(a + 8 * (i - 2)));
    }
  }
  store<FrP>(out + 8 * i, r);
}

// op 0: out = a * s; 1: out = a + s; 2: out = out + a * s
__global__ void k_vec_scalar(int op, const uint32_t* __restrict__ a, FrArg s, uint32_t* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t x = load<FrP>(a + 8 * i);
  fr_t r;
  if (op == 0) r = mul(x, arg_fr(s));
  else if (op == 1) r = add(x, arg_fr(s));
  else r = add(load<FrP>(out + 8 * i), mul(x, arg_fr(s)));
  store<FrP>(out + 8 * i, r);
}

static constexpr int SCAN_T_INV = 256;   // threads per batch-inversion workgroup

// Batch inversion (ff BatchInvert): zeros stay zero, every other value is replaced by its
// inverse.  The inverse of x_i is T^-1 times the product of all the other non-zero values
// (T the product of all of them), and any grouping of "the others" will do: a workgroup
// tile of INV_TILE values, thread t holding the tile's values t, t + 256, ... (loads
// coalesce).  The others of value (t, k) are: the other tiles (one product per tile from
// the tile totals' prefix and suffix), the tile's other threads (the threads' products'
// exclusive prefix and suffix in the workgroup) and the thread's other values (Montgomery's
// trick over its INV_E values).  Three launches and one inversion of T on the host (a
// Fermat inversion on one GPU lane is ~400 dependent products); no dependent chain is
// longer than a few tens of products.
static constexpr int INV_E = 4;   // values per thread (at 8 the two arrays below leave registers for scratch)
static constexpr uint32_t INV_TILE = SCAN_T_INV * INV_E;          // values per tile

// Workgroup-wide exclusive prefix and suffix products of one value per thread (zeros
// are the caller's to replace): pre = product over threads < t, suf = over threads > t.
__device__ void block_pre_suf(fr_t v, uint32_t* lp, uint32_t* ls, fr_t& pre, fr_t& suf, fr_t& total) {
  const int t = threadIdx.x;
  fr_t a = v, b = v;   // inclusive prefix, inclusive suffix
  store<FrP>(lp + 8 * t, a);
  store<FrP>(ls + 8 * t, b);
  __syncthreads();
  for (int d = 1; d < SCAN_T_INV; d <<= 1) {
    const fr_t ya = t >= d ? load<FrP>(lp + 8 * (t - d)) : fr_t::one();
    const fr_t yb = t + d < SCAN_T_INV ? load<FrP>(ls + 8 * (t + d)) : fr_t::one();
    __syncthreads();
    a = mul(a, ya);
    b = mul(b, yb);
    store<FrP>(lp + 8 * t, a);
    store<FrP>(ls + 8 * t, b);
    __syncthreads();
  }
  pre = t > 0 ? load<FrP>(lp + 8 * (t - 1)) : fr_t::one();
  suf = t + 1 < SCAN_T_INV ? load<FrP>(ls + 8 * (t + 1)) : fr_t::one();
  total = load<FrP>(lp + 8 * (SCAN_T_INV - 1));
  __syncthreads();
}

// Value k of thread t in tile `tile`: index tile INV_TILE + k 256 + t.
__device__ __forceinline__ size_t inv_index(uint32_t tile, int k) {
  return (size_t)tile * INV_TILE + (size_t)k * SCAN_T_INV + threadIdx.x;
}

// Pass 1: the product of each tile's non-zero values.
__global__ void __launch_bounds__(SCAN_T_INV) k_binv_tiles(const uint32_t* __restrict__ x, size_t n,
                                                            uint32_t* __restrict__ tot) {
  __shared__ uint32_t lds[8 * SCAN_T_INV];
  fr_t p = fr_t::one();
#pragma unroll
  for (int k = 0; k < INV_E; k++) {
    const size_t i = inv_index(blockIdx.x, k);
    if (i < n) {
      const fr_t v = load<FrP>(x + 8 * i);
      if (!v.is_zero()) p = mul(p, v);
    }
  }
  store<FrP>(lds + 8 * threadIdx.x, p);
  __syncthreads();
  for (int d = SCAN_T_INV / 2; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d) {
      p = mul(p, load<FrP>(lds + 8 * (threadIdx.x + d)));
      store<FrP>(lds + 8 * threadIdx.x, p);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) store<FrP>(tot + 8 * blockIdx.x, p);
}

// Pass 2, one workgroup: others[j] = product of the totals of every tile but j, and
// others[m] = T, the product of all m totals.  Thread t holds tiles t, t + 256, ...; its
// tiles' exclusive products are walked in others[] (forward products, then the backward
// trick), the other threads' by block_pre_suf.
__global__ void __launch_bounds__(SCAN_T_INV) k_binv_totals(const uint32_t* __restrict__ tot, size_t m,
                                                             uint32_t* __restrict__ others) {
  __shared__ uint32_t lp[8 * SCAN_T_INV], ls[8 * SCAN_T_INV];
  fr_t p = fr_t::one();
  for (size_t j = threadIdx.x; j < m; j += SCAN_T_INV) {
    store<FrP>(others + 8 * j, p);   // product of this thread's earlier tiles
    p = mul(p, load<FrP>(tot + 8 * j));
  }
  fr_t pre, suf, total;
  block_pre_suf(p, lp, ls, pre, suf, total);
  if (threadIdx.x < m) {
    fr_t b = mul(pre, suf);   // the other threads' tiles
    const size_t last = threadIdx.x + ((m - 1 - threadIdx.x) / SCAN_T_INV) * SCAN_T_INV;
    for (size_t j = last;; j -= SCAN_T_INV) {   // this thread's tiles, last to first
      store<FrP>(others + 8 * j, mul(load<FrP>(others + 8 * j), b));
      b = mul(b, load<FrP>(tot + 8 * j));
      if (j < SCAN_T_INV) break;
    }
  }
  if (threadIdx.x == 0) store<FrP>(others + 8 * m, total);
}

// Pass 3: inv[i] = x[i]^-1 (0 for 0) with tinv = T^-1.  inv may be x: a thread reads its
// values before it writes them, and no other thread touches them.
__global__ void __launch_bounds__(SCAN_T_INV) k_binv_apply(const uint32_t* x, size_t n,
                                                            const uint32_t* __restrict__ others, FrArg tinv,
                                                            uint32_t* inv) {
  __shared__ uint32_t lp[8 * SCAN_T_INV], ls[8 * SCAN_T_INV];
  fr_t v[INV_E], fw[INV_E];
  fr_t p = fr_t::one();
#pragma unroll
  for (int k = 0; k < INV_E; k++) {
    const size_t i = inv_index(blockIdx.x, k);
    v[k] = i < n ? load<FrP>(x + 8 * i) : fr_t::zero();
    fw[k] = p;   // product of this thread's values before k
    if (!v[k].is_zero()) p = mul(p, v[k]);
  }
  fr_t pre, suf, total;
  block_pre_suf(p, lp, ls, pre, suf, total);
  fr_t b = mul(mul(arg_fr(tinv), load<FrP>(others + 8 * blockIdx.x)), mul(pre, suf));
#pragma unroll
  for (int k = INV_E - 1; k >= 0; k--) {
    const size_t i = inv_index(blockIdx.x, k);
    const bool z = v[k].is_zero();
    if (i < n) store<FrP>(inv + 8 * i, z ? fr_t::zero() : mul(fw[k], b));
    if (!z) b = mul(b, v[k]);
  }
}

// ---------------------------------------------------------------- product scan
static constexpr int SCAN_T = 256;
static constexpr int SCAN_E = 16;   // elements per thread

// The scans run under multiplication (running products) or addition (running sums), a
// uniform choice per launch.
__device__ __forceinline__ fr_t scan_op(bool sum, const fr_t& x, const fr_t& y) { return sum ? add(x, y) : mul(x, y); }
__device__ __forceinline__ fr_t scan_id(bool sum) { return sum ? fr_t::zero() : fr_t::one(); }

// Exclusive block scan of one field element per thread.
__device__ fr_t block_excl_scan(fr_t v, uint32_t* lds, fr_t& total, bool sum) {
  const int t = threadIdx.x;
  store<FrP>(lds + 8 * t, v);
  __syncthreads();
  fr_t x = v;
  for (int d = 1; d < SCAN_T; d <<= 1) {
    fr_t y = t >= d ? load<FrP>(lds + 8 * (t - d)) : scan_id(sum);
    __syncthreads();
    x = scan_op(sum, x, y);
    store<FrP>(lds + 8 * t, x);
    __syncthreads();
  }
  total = load<FrP>(lds + 8 * (SCAN_T - 1));
  fr_t excl = t > 0 ? load<FrP>(lds + 8 * (t - 1)) : scan_id(sum);
  __syncthreads();
  return excl;
}

// Several independent scans share each phase's launch (the permutation argument's running
// products of every set and proof, permutation/prover.rs:190-197): a workgroup finds its
// scan from the first-tile table.
static constexpr int PROD_MAXJ = 48;
static constexpr uint32_t PROD_TILE = SCAN_T * SCAN_E;

struct ProdJob {
  const uint32_t* a;
  uint32_t* out;
  uint32_t* totals;   // one per tile
  uint64_t n;
};

struct ProdBatch {
  uint32_t njobs;
  uint32_t sum;   // 1: running sums, 0: running products
  uint32_t first_tile[PROD_MAXJ + 1];
  ProdJob j[PROD_MAXJ];
};
static_assert(sizeof(ProdBatch) <= 4000, "kernel argument block too large");

__device__ __forceinline__ uint32_t prod_job_of(const ProdBatch& B, uint32_t tile) {
  uint32_t lo = 0, hi = B.njobs;   // first_tile[lo] <= tile < first_tile[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (B.first_tile[mid] <= tile) lo = mid; else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// Phase 1: inclusive products within each tile of SCAN_T * SCAN_E; tile totals.
__global__ void __launch_bounds__(SCAN_T) k_prod_tiles(const ProdBatch B) {
  __shared__ uint32_t lds[8 * SCAN_T];
  const uint32_t jb = prod_job_of(B, blockIdx.x);
  const ProdJob& J = B.j[jb];
  const uint32_t tile = blockIdx.x - B.first_tile[jb];
  const size_t base = ((size_t)tile * SCAN_T + threadIdx.x) * SCAN_E;
  const bool sum = B.sum != 0;
  fr_t p = scan_id(sum);
  for (int k = 0; k < SCAN_E; k++)
    if (base + k < J.n) p = scan_op(sum, p, load<FrP>(J.a + 8 * (base + k)));
  fr_t total;
  fr_t run = block_excl_scan(p, lds, total, sum);
  for (int k = 0; k < SCAN_E; k++) {
    if (base + k >= J.n) break;
    run = scan_op(sum, run, load<FrP>(J.a + 8 * (base + k)));
    store<FrP>(J.out + 8 * (base + k), run);
  }
  if (threadIdx.x == 0) store<FrP>(J.totals + 8 * tile, total);
}

// Phase 2: one workgroup per scan turns its tile totals into exclusive prefixes
// (tiles <= SCAN_T * SCAN_E).
__global__ void __launch_bounds__(SCAN_T) k_prod_totals(const ProdBatch B) {
  __shared__ uint32_t lds[8 * SCAN_T];
  const ProdJob& J = B.j[blockIdx.x];
  const size_t tiles = B.first_tile[blockIdx.x + 1] - B.first_tile[blockIdx.x];
  const size_t base = (size_t)threadIdx.x * SCAN_E;
  const bool sum = B.sum != 0;
  fr_t p = scan_id(sum);
  for (int k = 0; k < SCAN_E; k++)
    if (base + k < tiles) p = scan_op(sum, p, load<FrP>(J.totals + 8 * (base + k)));
  fr_t total;
  fr_t run = block_excl_scan(p, lds, total, sum);
  for (int k = 0; k < SCAN_E; k++) {
    if (base + k >= tiles) break;
    fr_t v = load<FrP>(J.totals + 8 * (base + k));
    store<FrP>(J.totals + 8 * (base + k), run);
    run = scan_op(sum, run, v);
  }
}

// Phase 3: out[i] *= exclusive prefix of its tile; PROD_TILE / 256 workgroups per tile.
__global__ void __launch_bounds__(256) k_prod_apply(const ProdBatch B) {
  constexpr uint32_t per_tile = PROD_TILE / 256;
  const uint32_t gtile = blockIdx.x / per_tile;
  const uint32_t jb = prod_job_of(B, gtile);
  const ProdJob& J = B.j[jb];
  const uint32_t tile = gtile - B.first_tile[jb];
  if (tile == 0) return;
  const size_t i = (size_t)tile * PROD_TILE + (blockIdx.x % per_tile) * 256 + threadIdx.x;
  if (i >= J.n) return;
  store<FrP>(J.out + 8 * i, scan_op(B.sum != 0, load<FrP>(J.out + 8 * i), load<FrP>(J.totals + 8 * tile)));
}

// ---------------------------------------------------------------- evaluation
struct Pows {
  FrArg p[32];   // x^(2^j)
};

__device__ __forceinline__ fr_t pow_index(const Pows& P, uint64_t e) {
  fr_t r = fr_t::one();
  for (int j = 0; e; j++, e >>= 1)
    if (e & 1) r = mul(r, arg_fr(P.p[j]));
  return r;
}

static constexpr int EVAL_CHUNK = 64;

// Horner per chunk of EVAL_CHUNK coefficients, times x^start; block sums in `partials`.
__global__ void __launch_bounds__(256) k_eval(const uint32_t* __restrict__ a, size_t n, FrArg x, Pows P,
                                              uint32_t* __restrict__ partials) {
  __shared__ uint32_t lds[8 * 256];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t lo = t * EVAL_CHUNK;
  fr_t acc = fr_t::zero();
  if (lo < n) {
    const size_t hi = lo + EVAL_CHUNK < n ? lo + EVAL_CHUNK : n;
    const fr_t xx = arg_fr(x);
    for (size_t i = hi; i-- > lo;) acc = add(mul(acc, xx), load<FrP>(a + 8 * i));
    acc = mul(acc, pow_index(P, lo));
  }
  store<FrP>(lds + 8 * threadIdx.x, acc);
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d) {
      acc = add(acc, load<FrP>(lds + 8 * (threadIdx.x + d)));
      store<FrP>(lds + 8 * threadIdx.x, acc);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) store<FrP>(partials + 8 * blockIdx.x, acc);
}

// Many evaluations in one launch: block b belongs to the job whose block range holds it.
struct EvalJob {
  const uint32_t* a;
  uint64_t n;
  uint32_t first_block;
  uint32_t pad;
};

__global__ void __launch_bounds__(256) k_eval_batch(const EvalJob* __restrict__ jobs, int njobs,
                                                    const uint32_t* __restrict__ xpows,
                                                    uint32_t* __restrict__ partials) {
  __shared__ uint32_t lds[8 * 256];
  int lo_j = 0, hi_j = njobs - 1;   // last job with first_block <= blockIdx.x
  while (lo_j < hi_j) {
    const int mid = (lo_j + hi_j + 1) >> 1;
    if (jobs[mid].first_block <= blockIdx.x) lo_j = mid;
    else hi_j = mid - 1;
  }
  const EvalJob J = jobs[lo_j];
  const uint32_t* P = xpows + (size_t)lo_j * 32 * 8;   // x^(2^j), j < 32
  const size_t t = (size_t)(blockIdx.x - J.first_block) * blockDim.x + threadIdx.x;
  const size_t lo = t * EVAL_CHUNK;
  fr_t acc = fr_t::zero();
  if (lo < J.n) {
    const size_t hi = lo + EVAL_CHUNK < J.n ? lo + EVAL_CHUNK : J.n;
    const fr_t xx = load<FrP>(P);
    for (size_t i = hi; i-- > lo;) acc = add(mul(acc, xx), load<FrP>(J.a + 8 * i));
    fr_t s = fr_t::one();
    uint64_t e = lo;
    for (int j = 0; e; j++, e >>= 1)
      if (e & 1) s = mul(s, load<FrP>(P + 8 * j));
    acc = mul(acc, s);
  }
  store<FrP>(lds + 8 * threadIdx.x, acc);
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d) {
      acc = add(acc, load<FrP>(lds + 8 * (threadIdx.x + d)));
      store<FrP>(lds + 8 * threadIdx.x, acc);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) store<FrP>(partials + 8 * blockIdx.x, acc);
}

// ---------------------------------------------------------------- Kate division
// q_i = a_(i+1) + b q_(i+1), i < m = n - 1, with q_m = 0 (kate_division).  Chunk c holds
// q indices [c L, c L + L): its bottom value is f_c(x) = loc_c + g x for the value x at
// the bottom of the chunk above (g = b^L).  Four passes: loc_c per chunk; per workgroup
// of 256 chunks the suffix compositions f_c o ... o f_(last of the group) (uniform
// multiplier g, so a Hillis-Steele scan of the loc values with g^(2^k)); one workgroup
// per division carries the groups' totals down; each chunk recomputes its q values from
// its carry-in.  Up to KATE_MAXJ independent divisions share each pass's launch (the
// multi-open divides every point set's polynomial at once, kzg/mod.rs:226-236): a
// workgroup finds its division from the per-division first-workgroup tables.
static constexpr uint32_t KATE_L = 8;
static constexpr uint32_t KATE_GROUP = 256;
static constexpr int KATE_MAXJ = 9;   // the kernel argument block stays below 4 KB
static constexpr uint32_t KATE_CHUNK_BLOCK = 64;

struct KatePows {
  FrArg g2k[8];   // g^(2^k), g = b^L
};

struct KateJob {
  const uint32_t* a;
  uint32_t* q;
  uint32_t *loc, *suf, *tot, *gcarry;   // scratch: C, G whole groups, G, G elements
  uint64_t m, C, G;
  FrArg b, g256;
  KatePows P;
};

struct KateBatch {
  uint32_t njobs;
  uint32_t first_chunk_block[KATE_MAXJ + 1];   // k_kate_local / k_kate_final workgroups
  uint32_t first_group[KATE_MAXJ + 1];         // k_kate_group workgroups
  KateJob j[KATE_MAXJ];
};
static_assert(sizeof(KateBatch) <= 4000, "kernel argument block too large");

// the division a workgroup belongs to (uniform across the workgroup)
__device__ __forceinline__ uint32_t kate_job_of(const uint32_t* first, uint32_t njobs, uint32_t blk) {
  uint32_t jb = 0;
  while (jb + 1 < njobs && blk >= first[jb + 1]) jb++;
  return __builtin_amdgcn_readfirstlane(jb);
}

// Phase 1: loc_c (the chunk's bottom value with a zero carry-in).
__global__ void __launch_bounds__(KATE_CHUNK_BLOCK) k_kate_local(const KateBatch B) {
  const uint32_t jb = kate_job_of(B.first_chunk_block, B.njobs, blockIdx.x);
  const KateJob& J = B.j[jb];
  const size_t c = (size_t)(blockIdx.x - B.first_chunk_block[jb]) * KATE_CHUNK_BLOCK + threadIdx.x;
  const size_t lo = c * KATE_L;
  if (lo >= J.m) return;
  const size_t hi = lo + KATE_L < J.m ? lo + KATE_L : J.m;
  const fr_t bb = arg_fr(J.b);
  fr_t q = fr_t::zero();
  for (size_t i = hi; i-- > lo;) q = add(load<FrP>(J.a + 8 * (i + 1)), mul(bb, q));
  store<FrP>(J.loc + 8 * c, q);
}

// Phase 3 (one workgroup per division): carry-in of every group, i.e. the value at the
// bottom of the group above.  The group maps F_g(x) = tot_g + g256 x compose from the top;
// thread t owns groups [t S, t S + S) (S = ceil(G / 256)); a suffix scan of the composed
// per-thread maps gives each thread its incoming value.
__global__ void __launch_bounds__(256) k_kate_carries(const KateBatch B) {
  __shared__ uint32_t la[8 * 256], lm[8 * 256];
  const KateJob& J = B.j[blockIdx.x];
  const int t = threadIdx.x;
  const size_t C = J.G;
  const size_t S = (C + 255) / 256;
  const size_t c0 = t * S;
  const size_t c1 = c0 + S < C ? c0 + S : C;
  const fr_t g = arg_fr(J.g256);
  // thread map F_t = f_c0 o f_(c0+1) o ... o f_(c1-1): F(x) = A + M x
  fr_t A = fr_t::zero(), M = fr_t::one();
  for (size_t c = c1; c-- > c0;) {   // apply the lower group last: F <- f_c o F
    A = add(load<FrP>(J.tot + 8 * c), mul(g, A));
    M = mul(g, M);
  }
  if (c0 >= C) {
    A = fr_t::zero();
    M = fr_t::one();
  }
  // suffix scan over threads: incoming(t) = F_(t+1) o F_(t+2) o ... (0)
  store<FrP>(la + 8 * t, A);
  store<FrP>(lm + 8 * t, M);
  __syncthreads();
  fr_t SA = A, SM = M;   // composition of F_t .. F_(t + 2^k - 1)
  for (int d = 1; d < 256; d <<= 1) {
    fr_t nA = fr_t::zero(), nM = fr_t::one();
    if (t + d < 256) {
      nA = load<FrP>(la + 8 * (t + d));
      nM = load<FrP>(lm + 8 * (t + d));
    }
    __syncthreads();
    // (S o N)(x) = SA + SM (nA + nM x)
    SA = add(SA, mul(SM, nA));
    SM = mul(SM, nM);
    store<FrP>(la + 8 * t, SA);
    store<FrP>(lm + 8 * t, SM);
    __syncthreads();
  }
  // value entering thread t from above = (F_(t+1) o ...)(0) = suffix A of t + 1
  fr_t in = t + 1 < 256 ? load<FrP>(la + 8 * (t + 1)) : fr_t::zero();
  for (size_t c = c1; c-- > c0;) {
    store<FrP>(J.gcarry + 8 * c, in);
    in = add(load<FrP>(J.tot + 8 * c), mul(g, in));
  }
}

// Phase 2, group suffix compositions: suf[c] = sum_(j >= c, j in c's group) g^(j - c) loc_j
// (chunks past C count as loc = 0); tot[group] = suf of the group's first chunk.
__global__ void __launch_bounds__(KATE_GROUP) k_kate_group(const KateBatch B) {
  __shared__ uint32_t la[8 * KATE_GROUP];
  const uint32_t jb = kate_job_of(B.first_group, B.njobs, blockIdx.x);
  const KateJob& J = B.j[jb];
  const uint32_t grp = blockIdx.x - B.first_group[jb];
  const uint32_t t = threadIdx.x;
  const size_t c = (size_t)grp * KATE_GROUP + t;
  fr_t A = c < J.C ? load<FrP>(J.loc + 8 * c) : fr_t::zero();
  store<FrP>(la + 8 * t, A);
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < 8; k++) {   // A_t <- A_t + g^(2^k) A_(t + 2^k)
    const uint32_t d = 1u << k;
    const fr_t nb = t + d < KATE_GROUP ? load<FrP>(la + 8 * (t + d)) : fr_t::zero();
    __syncthreads();
    A = add(A, mul(arg_fr(J.P.g2k[k]), nb));
    store<FrP>(la + 8 * t, A);
    __syncthreads();
  }
  store<FrP>(J.suf + 8 * c, A);   // whole groups: chunks past C hold 0, as the final pass reads
  if (t == 0) store<FrP>(J.tot + 8 * (size_t)grp, A);
}

// Phase 4: each chunk with its true carry-in (suf of the next chunk carried by the group's
// carry-in gcarry, or gcarry itself for a group's last chunk); writes q.
__global__ void __launch_bounds__(KATE_CHUNK_BLOCK) k_kate_final(const KateBatch B) {
  const uint32_t jb = kate_job_of(B.first_chunk_block, B.njobs, blockIdx.x);
  const KateJob& J = B.j[jb];
  const size_t c = (size_t)(blockIdx.x - B.first_chunk_block[jb]) * KATE_CHUNK_BLOCK + threadIdx.x;
  const size_t lo = c * KATE_L;
  if (lo >= J.m) return;
  const size_t hi = lo + KATE_L < J.m ? lo + KATE_L : J.m;
  const uint32_t i = (uint32_t)(c % KATE_GROUP);
  const fr_t gc = load<FrP>(J.gcarry + 8 * (c / KATE_GROUP));
  fr_t v = gc;
  if (i + 1 < KATE_GROUP) {   // suf[c + 1] + g^(255 - i) gcarry (suf exists: the group is allocated whole)
    const uint32_t e = KATE_GROUP - 1 - i;
    fr_t ge = fr_t::one();
#pragma unroll 1
    for (int k = 0; k < 8; k++)
      if ((e >> k) & 1u) ge = mul(ge, arg_fr(J.P.g2k[k]));
    v = add(load<FrP>(J.suf + 8 * (c + 1)), mul(ge, gc));
  }
  const fr_t bb = arg_fr(J.b);
  for (size_t j = hi; j-- > lo;) {
    v = add(load<FrP>(J.a + 8 * (j + 1)), mul(bb, v));
    store<FrP>(J.q + 8 * j, v);
  }
}

// out[i] = x^i
__global__ void k_powers(uint32_t* __restrict__ out, size_t n, Pows P) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  store<FrP>(out + 8 * i, pow_index(P, i));
}

static Pows make_pows(const host::Fr& x) {
  Pows P;
  host::Fr t = x;
  for (int j = 0; j < 32; j++) {
    P.p[j] = to_arg(t);
    t = host::fr_mul(t, t);
  }
  return P;
}

// ---------------------------------------------------------------- permutation products
struct PermArgs {
  const uint32_t* val[MZK_PERM_MAX_CHUNK];
  const uint32_t* sig[MZK_PERM_MAX_CHUNK];
  FrArg delta_beta[MZK_PERM_MAX_CHUNK];
  FrArg beta, gamma;
  const uint32_t* omega;
  uint32_t* num;
  uint32_t* den;
  uint64_t n;
  uint32_t ncols;
};

// One element per thread: both products of the set at row i (permutation/prover.rs:129-160).
__global__ void __launch_bounds__(256) k_perm_products(PermArgs A) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const fr_t beta = arg_fr(A.beta), gamma = arg_fr(A.gamma);
  const fr_t w = load<FrP>(A.omega + 8 * i);
  fr_t nu = fr_t::one(), de = fr_t::one();
  for (uint32_t j = 0; j < A.ncols; j++) {
    const fr_t v = add(load<FrP>(A.val[j] + 8 * i), gamma);
    de = mul(de, add(v, mul(beta, load<FrP>(A.sig[j] + 8 * i))));
    nu = mul(nu, add(v, mul(arg_fr(A.delta_beta[j]), w)));
  }
  store<FrP>(A.den + 8 * i, de);
  store<FrP>(A.num + 8 * i, nu);
}

// ---------------------------------------------------------------- quotient numerator
struct NumArgs {
  uint64_t N;
  uint32_t rot_next, rot_last, ncols, nsets, chunk, tmask;
  const uint32_t *a, *b, *c, *qm, *ql, *qr, *qo, *qc;
  const uint32_t* col[8];
  const uint32_t* sigma[8];
  const uint32_t* z[4];
  const uint32_t *l0, *l_last, *l_active, *coset_x;
  FrArg y, beta, gamma;
  FrArg delta_beta[8];
  FrArg t_evals[16];
};

__global__ void __launch_bounds__(256) k_plonk_numerator(NumArgs A, uint32_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N) return;
  const size_t N = A.N;
  const size_t in = (i + A.rot_next) & (N - 1);
  const size_t il = (i + N - A.rot_last) & (N - 1);
  const fr_t y = arg_fr(A.y), beta = arg_fr(A.beta), gamma = arg_fr(A.gamma);
  const fr_t a = load<FrP>(A.a + 8 * i), b = load<FrP>(A.b + 8 * i), c = load<FrP>(A.c + 8 * i);
  // gate: a b q_m + a q_l + b q_r - c q_o + q_c
  fr_t v = add(add(add(mul(mul(a, b), load<FrP>(A.qm + 8 * i)), mul(a, load<FrP>(A.ql + 8 * i))),
                   mul(b, load<FrP>(A.qr + 8 * i))),
               sub(load<FrP>(A.qc + 8 * i), mul(c, load<FrP>(A.qo + 8 * i))));
  const fr_t l0 = load<FrP>(A.l0 + 8 * i);
  // l_0 (1 - z_0)
  v = add(mul(v, y), mul(sub(fr_t::one(), load<FrP>(A.z[0] + 8 * i)), l0));
  // l_last (z_l^2 - z_l)
  const fr_t zl = load<FrP>(A.z[A.nsets - 1] + 8 * i);
  v = add(mul(v, y), mul(sub(sqr(zl), zl), load<FrP>(A.l_last + 8 * i)));
  // l_0 (z_s - z_(s-1)(omega^last X))
  for (uint32_t s = 1; s < A.nsets; s++)
    v = add(mul(v, y), mul(sub(load<FrP>(A.z[s] + 8 * i), load<FrP>(A.z[s - 1] + 8 * il)), l0));
  const fr_t x = load<FrP>(A.coset_x + 8 * i);
  const fr_t lact = load<FrP>(A.l_active + 8 * i);
  for (uint32_t s = 0; s < A.nsets; s++) {
    fr_t left = load<FrP>(A.z[s] + 8 * in);
    fr_t right = load<FrP>(A.z[s] + 8 * i);
    const uint32_t j0 = s * A.chunk, j1 = j0 + A.chunk < A.ncols ? j0 + A.chunk : A.ncols;
    for (uint32_t j = j0; j < j1; j++) {
      const fr_t col = load<FrP>(A.col[j] + 8 * i);
      left = mul(left, add(add(col, mul(load<FrP>(A.sigma[j] + 8 * i), beta)), gamma));
      right = mul(right, add(add(col, mul(x, arg_fr(A.delta_beta[j]))), gamma));
    }
    v = add(mul(v, y), mul(sub(left, right), lact));
  }
  store<FrP>(out + 8 * i, mul(v, arg_fr(A.t_evals[i & A.tmask])));
}

// sum_j s_j p_j over up to LC_MAX polynomials (zero-extended)
static constexpr int LC_MAX = 64;
struct LcArgs {
  const uint32_t* p[LC_MAX];
  uint64_t len[LC_MAX];
  FrArg s[LC_MAX];
  uint32_t np;
};

// (the argument block, 3 KiB, travels by value in the kernel arguments)
__global__ void __launch_bounds__(256) k_lincomb(LcArgs A, uint32_t* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t acc = fr_t::zero();
  for (uint32_t j = 0; j < A.np; j++)
    if (i < A.len[j]) acc = add(acc, mul(load<FrP>(A.p[j] + 8 * i), arg_fr(A.s[j])));
  store<FrP>(out + 8 * i, acc);
}

}  // namespace mzk

using namespace mzk;

static int vec_ready(mzk_ctx* c) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_HIP(hipSetDevice(c->device));
  return MZK_OK;
}

extern "C" {

int mzk_fr_vec_op(mzk_ctx* c, int op, const void* a, const void* b, void* out, size_t n) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_REQUIRE(op >= 0 && op <= 2, "mzk_fr_vec_op: op out of range");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a && b && out, "mzk_fr_vec_op: null data");
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  MZK_LAUNCH(c, "k_vec_binop", k_vec_binop, dim3(grid_for(n, 256)), dim3(256), 0, op, (const uint32_t*)a,
             (const uint32_t*)b, (uint32_t*)out, n);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

}  // extern "C"
static bool spans_overlap(const void* p, size_t pb, const void* q, size_t qb);
extern "C" {

int mzk_fr_to_delta_dev(mzk_ctx* c, const void* a, void* out, size_t n, int order) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_REQUIRE(order == 1 || order == 2, "mzk_fr_to_delta_dev: order %d is not 1 or 2", order);
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a && out, "mzk_fr_to_delta_dev: null data");
  MZK_REQUIRE(!spans_overlap(a, n * 32, out, n * 32), "mzk_fr_to_delta_dev: output overlaps input");
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  MZK_LAUNCH(c, "k_to_delta", k_to_delta, dim3(grid_for(n, 256)), dim3(256), 0, (const uint32_t*)a, (uint32_t*)out,
             n, order);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

int mzk_fr_vec_scalar(mzk_ctx* c, int op, const void* a, const uint64_t s[4], void* out, size_t n) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_REQUIRE(op >= 0 && op <= 2, "mzk_fr_vec_scalar: op out of range");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a && s && out, "mzk_fr_vec_scalar: null data");
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  MZK_LAUNCH(c, "k_vec_scalar", k_vec_scalar, dim3(grid_for(n, 256)), dim3(256), 0, op, (const uint32_t*)a,
             to_arg(s), (uint32_t*)out, n);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

int mzk_fr_batch_invert(mzk_ctx* c, void* a, size_t n) {
  MZK_REQUIRE(c != nullptr, "null context");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a != nullptr, "mzk_fr_batch_invert: null data");
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  const size_t m = (n + INV_TILE - 1) / INV_TILE;   // tiles
  MZK_REQUIRE(m < (1ull << 31), "mzk_fr_batch_invert: n too large");
  // scratch: tile totals (m), the other tiles' products (m) and T
  MZK_TRY(c->scratch[SCR_MSM_D].reserve((2 * m + 1) * 32 + 256));
  uint32_t* tot = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  uint32_t* others = tot + 8 * m;
  MZK_LAUNCH(c, "k_binv_tiles", k_binv_tiles, dim3((unsigned)m), dim3(SCAN_T_INV), 0, (const uint32_t*)a, n, tot);
  MZK_LAUNCH(c, "k_binv_totals", k_binv_totals, dim3(1), dim3(SCAN_T_INV), 0, tot, m, others);
  host::Fr root;
  MZK_HIP(hipMemcpyAsync(root.l, others + 8 * m, 32, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  const bool zero = (root.l[0] | root.l[1] | root.l[2] | root.l[3]) == 0;   // not reached: zeros are skipped
  const host::Fr rinv = zero ? host::fr_zero() : host::fr_inv(root);
  MZK_LAUNCH(c, "k_binv_apply", k_binv_apply, dim3((unsigned)m), dim3(SCAN_T_INV), 0, (const uint32_t*)a, n,
             others, to_arg(rinv), (uint32_t*)a);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

static bool spans_overlap(const void* p, size_t pb, const void* q, size_t qb) {
  const uintptr_t a = (uintptr_t)p, b = (uintptr_t)q;
  return a < b + qb && b < a + pb;
}

static int prefix_scan_batch(mzk_ctx* c, const mzk_prefix_job* jobs, size_t njobs, bool sum);

int mzk_fr_prefix_product_batch(mzk_ctx* c, const mzk_prefix_job* jobs, size_t njobs) {
  return prefix_scan_batch(c, jobs, njobs, false);
}

int mzk_fr_prefix_sum(mzk_ctx* c, const void* a, void* out, size_t n) {
  MZK_REQUIRE(c != nullptr, "null context");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a && out, "mzk_fr_prefix_sum: null data");
  mzk_prefix_job j;
  j.d_a = a;
  j.d_out = out;
  j.n = n;
  return prefix_scan_batch(c, &j, 1, true);
}

}  // extern "C"

// Inclusive running products (sum = false) or sums (sum = true) of every job, each phase
// one launch for up to PROD_MAXJ jobs.
static int prefix_scan_batch(mzk_ctx* c, const mzk_prefix_job* jobs, size_t njobs, bool sum) {
  const char* who = sum ? "mzk_fr_prefix_sum" : "mzk_fr_prefix_product_batch";
  MZK_REQUIRE(c != nullptr && (jobs != nullptr || njobs == 0), "%s: null argument", who);
  std::vector<mzk_prefix_job> live;
  for (size_t i = 0; i < njobs; i++) {
    const mzk_prefix_job& j = jobs[i];
    if (j.n == 0) continue;
    MZK_REQUIRE(j.d_a && j.d_out, "%s: job %zu: null data", who, i);
    MZK_REQUIRE(j.n <= (size_t)PROD_TILE * PROD_TILE, "%s: job %zu: n above %zu", who, i,
                (size_t)PROD_TILE * PROD_TILE);
    live.push_back(j);
  }
  // every scan runs at once: no output may overlap any input or another output
  for (size_t i = 0; i < live.size(); i++)
    for (size_t k = 0; k < live.size(); k++) {
      MZK_REQUIRE(!spans_overlap(live[i].d_out, live[i].n * 32, live[k].d_a, live[k].n * 32),
                  "%s: an output overlaps an input", who);
      MZK_REQUIRE(i == k || !spans_overlap(live[i].d_out, live[i].n * 32, live[k].d_out, live[k].n * 32),
                  "%s: two outputs overlap", who);
    }
  if (live.empty()) return MZK_OK;
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  size_t all_tiles = 0;
  for (auto& j : live) all_tiles += (j.n + PROD_TILE - 1) / PROD_TILE;
  MZK_TRY(c->scratch[SCR_MSM_D].reserve(all_tiles * 32 + 64));
  uint32_t* totals = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  for (size_t i0 = 0; i0 < live.size(); i0 += PROD_MAXJ) {
    const size_t nj = std::min<size_t>(PROD_MAXJ, live.size() - i0);
    ProdBatch B;
    memset(&B, 0, sizeof B);
    B.njobs = (uint32_t)nj;
    B.sum = sum ? 1 : 0;
    uint32_t t = 0;
    for (size_t i = 0; i < nj; i++) {
      const mzk_prefix_job& j = live[i0 + i];
      B.j[i].a = (const uint32_t*)j.d_a;
      B.j[i].out = (uint32_t*)j.d_out;
      B.j[i].n = j.n;
      B.j[i].totals = totals;
      B.first_tile[i] = t;
      const uint32_t tiles = (uint32_t)((j.n + PROD_TILE - 1) / PROD_TILE);
      t += tiles;
      totals += 8 * (size_t)tiles;
    }
    B.first_tile[nj] = t;
    MZK_LAUNCH(c, "k_prod_tiles", k_prod_tiles, dim3(t), dim3(SCAN_T), 0, B);
    MZK_LAUNCH(c, "k_prod_totals", k_prod_totals, dim3((unsigned)nj), dim3(SCAN_T), 0, B);
    MZK_LAUNCH(c, "k_prod_apply", k_prod_apply, dim3(t * (PROD_TILE / 256)), dim3(256), 0, B);
  }
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

extern "C" {

int mzk_fr_prefix_product(mzk_ctx* c, const void* a, void* out, size_t n) {
  MZK_REQUIRE(c != nullptr, "null context");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a && out, "mzk_fr_prefix_product: null data");
  mzk_prefix_job j;
  j.d_a = a;
  j.d_out = out;
  j.n = n;
  return mzk_fr_prefix_product_batch(c, &j, 1);
}

int mzk_fr_eval(mzk_ctx* c, const void* a, size_t n, const uint64_t x[4], uint64_t out[4]) {
  MZK_REQUIRE(c != nullptr && x && out, "mzk_fr_eval: null argument");
  if (n == 0) {
    memset(out, 0, 32);
    return MZK_OK;
  }
  MZK_REQUIRE(a != nullptr, "mzk_fr_eval: null data");
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  const size_t threads = (n + EVAL_CHUNK - 1) / EVAL_CHUNK;
  const unsigned blocks = grid_for(threads, 256);
  MZK_TRY(c->scratch[SCR_MSM_D].reserve((size_t)blocks * 32 + 64));
  uint32_t* partials = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  host::Fr hx;
  memcpy(hx.l, x, 32);
  MZK_LAUNCH(c, "k_eval", k_eval, dim3(blocks), dim3(256), 0, (const uint32_t*)a, n, to_arg(x), make_pows(hx),
             partials);
  MZK_HIP(hipGetLastError());
  std::vector<uint64_t> h(4 * (size_t)blocks);
  MZK_HIP(hipMemcpyAsync(h.data(), partials, 32 * (size_t)blocks, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  host::Fr acc = host::fr_zero();
  for (unsigned b = 0; b < blocks; b++) {
    host::Fr v;
    memcpy(v.l, &h[4 * b], 32);
    acc = host::fr_add(acc, v);
  }
  memcpy(out, acc.l, 32);
  return MZK_OK;
}

int mzk_fr_eval_batch(mzk_ctx* c, const void* const* polys, const size_t* ns, const uint64_t* xs, size_t njobs,
                      uint64_t* outs) {
  MZK_REQUIRE(c != nullptr, "null context");
  if (njobs == 0) return MZK_OK;
  MZK_REQUIRE(polys && ns && xs && outs, "mzk_fr_eval_batch: null argument");
  MZK_REQUIRE(njobs <= (1u << 20), "mzk_fr_eval_batch: too many jobs");
  std::vector<EvalJob> jobs(njobs);
  std::vector<uint64_t> pows(njobs * 32 * 4);
  size_t blocks = 0;
  for (size_t j = 0; j < njobs; j++) {
    MZK_REQUIRE(ns[j] == 0 || polys[j] != nullptr, "mzk_fr_eval_batch: job %zu has null data", j);
    const size_t threads = (ns[j] + EVAL_CHUNK - 1) / EVAL_CHUNK;
    const size_t nb = threads == 0 ? 1 : (threads + 255) / 256;
    jobs[j].a = (const uint32_t*)polys[j];
    jobs[j].n = ns[j];
    jobs[j].first_block = (uint32_t)blocks;
    jobs[j].pad = 0;
    blocks += nb;
    MZK_REQUIRE(blocks < (1ull << 31), "mzk_fr_eval_batch: too much work in one call");
    host::Fr t;
    memcpy(t.l, xs + 4 * j, 32);
    for (int k = 0; k < 32; k++) {
      memcpy(&pows[(j * 32 + k) * 4], t.l, 32);
      t = host::fr_mul(t, t);
    }
  }
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  const size_t job_bytes = njobs * sizeof(EvalJob), pow_bytes = pows.size() * 8;
  MZK_TRY(c->scratch[SCR_IN1].reserve(job_bytes + pow_bytes + 256));
  MZK_TRY(c->scratch[SCR_MSM_D].reserve(blocks * 32 + 64));
  char* tab = (char*)c->scratch[SCR_IN1].ptr;
  const size_t pow_off = (job_bytes + 255) & ~(size_t)255;
  uint32_t* partials = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  MZK_HIP(hipMemcpyAsync(tab, jobs.data(), job_bytes, hipMemcpyHostToDevice, c->stream));
  MZK_HIP(hipMemcpyAsync(tab + pow_off, pows.data(), pow_bytes, hipMemcpyHostToDevice, c->stream));
  MZK_LAUNCH(c, "k_eval_batch", k_eval_batch, dim3((unsigned)blocks), dim3(256), 0, (const EvalJob*)tab, (int)njobs,
             (const uint32_t*)(tab + pow_off), partials);
  MZK_HIP(hipGetLastError());
  std::vector<uint64_t> h(4 * blocks);
  MZK_HIP(hipMemcpyAsync(h.data(), partials, 32 * blocks, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  for (size_t j = 0; j < njobs; j++) {
    const size_t b0 = jobs[j].first_block, b1 = j + 1 < njobs ? jobs[j + 1].first_block : blocks;
    host::Fr acc = host::fr_zero();
    for (size_t b = b0; b < b1; b++) {
      host::Fr v;
      memcpy(v.l, &h[4 * b], 32);
      acc = host::fr_add(acc, v);
    }
    memcpy(outs + 4 * j, acc.l, 32);
  }
  return MZK_OK;
}

// One launch of each pass for up to KATE_MAXJ divisions (every job has n >= 2).
static int kate_launch(mzk_ctx* c, const mzk_kate_job* jobs, size_t nj) {
  KateBatch B;
  memset(&B, 0, sizeof B);
  B.njobs = (uint32_t)nj;
  size_t words = 0;
  for (size_t i = 0; i < nj; i++) {
    const size_t m = jobs[i].n - 1, C = (m + KATE_L - 1) / KATE_L, G = (C + KATE_GROUP - 1) / KATE_GROUP;
    words += 8 * (C + G * KATE_GROUP + 2 * G);
  }
  MZK_TRY(c->scratch[SCR_MSM_D].reserve(words * 4 + 64));
  uint32_t* w = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  uint32_t cb = 0, gb = 0;
  for (size_t i = 0; i < nj; i++) {
    KateJob& J = B.j[i];
    J.a = (const uint32_t*)jobs[i].d_a;
    J.q = (uint32_t*)jobs[i].d_q;
    J.m = jobs[i].n - 1;
    J.C = (J.m + KATE_L - 1) / KATE_L;
    J.G = (J.C + KATE_GROUP - 1) / KATE_GROUP;
    J.loc = w;
    J.suf = J.loc + 8 * J.C;
    J.tot = J.suf + 8 * J.G * KATE_GROUP;
    J.gcarry = J.tot + 8 * J.G;
    w = J.gcarry + 8 * J.G;
    host::Fr hb;
    memcpy(hb.l, jobs[i].b, 32);
    J.b = to_arg(hb);
    host::Fr gk = host::fr_pow(hb, KATE_L);
    for (int k = 0; k < 8; k++) {
      J.P.g2k[k] = to_arg(gk);
      gk = host::fr_mul(gk, gk);
    }
    J.g256 = to_arg(gk);   // g^(2^8): one group's multiplier
    B.first_chunk_block[i] = cb;
    B.first_group[i] = gb;
    cb += (uint32_t)((J.C + KATE_CHUNK_BLOCK - 1) / KATE_CHUNK_BLOCK);
    gb += (uint32_t)J.G;
  }
  B.first_chunk_block[nj] = cb;
  B.first_group[nj] = gb;
  MZK_LAUNCH(c, "k_kate_local", k_kate_local, dim3(cb), dim3(KATE_CHUNK_BLOCK), 0, B);
  MZK_LAUNCH(c, "k_kate_group", k_kate_group, dim3(gb), dim3(KATE_GROUP), 0, B);
  MZK_LAUNCH(c, "k_kate_carries", k_kate_carries, dim3((unsigned)nj), dim3(256), 0, B);
  MZK_LAUNCH(c, "k_kate_final", k_kate_final, dim3(cb), dim3(KATE_CHUNK_BLOCK), 0, B);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}


int mzk_fr_kate_div_batch(mzk_ctx* c, const mzk_kate_job* jobs, size_t njobs) {
  MZK_REQUIRE(c != nullptr && (jobs != nullptr || njobs == 0), "mzk_fr_kate_div_batch: null argument");
  std::vector<mzk_kate_job> live;
  for (size_t i = 0; i < njobs; i++) {
    const mzk_kate_job& j = jobs[i];
    MZK_REQUIRE(j.n >= 1, "mzk_fr_kate_div_batch: job %zu: empty polynomial", i);
    MZK_REQUIRE(j.n <= ((size_t)1 << 28), "mzk_fr_kate_div_batch: job %zu: n = %zu above 2^28", i, j.n);
    if (j.n == 1) continue;   // the quotient has no coefficients
    MZK_REQUIRE(j.d_a && j.d_q, "mzk_fr_kate_div_batch: job %zu: null data", i);
    live.push_back(j);
  }
  // every job runs at once: no output may overlap any job's input or another output
  for (size_t i = 0; i < live.size(); i++)
    for (size_t k = 0; k < live.size(); k++) {
      MZK_REQUIRE(!spans_overlap(live[i].d_q, (live[i].n - 1) * 32, live[k].d_a, live[k].n * 32),
                  "mzk_fr_kate_div_batch: an output overlaps an input");
      MZK_REQUIRE(i == k || !spans_overlap(live[i].d_q, (live[i].n - 1) * 32, live[k].d_q, (live[k].n - 1) * 32),
                  "mzk_fr_kate_div_batch: two outputs overlap");
    }
  if (live.empty()) return MZK_OK;
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  for (size_t i = 0; i < live.size(); i += KATE_MAXJ)
    MZK_TRY(kate_launch(c, live.data() + i, std::min<size_t>(KATE_MAXJ, live.size() - i)));
  return MZK_OK;
}

int mzk_fr_kate_div(mzk_ctx* c, const void* a, size_t n, const uint64_t b[4], void* q) {
  MZK_REQUIRE(c != nullptr && b != nullptr, "mzk_fr_kate_div: null argument");
  MZK_REQUIRE(n >= 1, "mzk_fr_kate_div: empty polynomial");
  if (n == 1) return MZK_OK;   // the quotient has no coefficients
  MZK_REQUIRE(a && q && a != q, "mzk_fr_kate_div: null or aliased data");
  mzk_kate_job j;
  j.d_a = a;
  j.n = n;
  memcpy(j.b, b, 32);
  j.d_q = q;
  return mzk_fr_kate_div_batch(c, &j, 1);
}

int mzk_plonk_perm_products_dev(mzk_ctx* c, const void* omega_powers, size_t n, const mzk_perm_set* sets,
                                size_t nsets) {
  MZK_REQUIRE(c != nullptr && (sets != nullptr || nsets == 0), "mzk_plonk_perm_products_dev: null argument");
  if (n == 0 || nsets == 0) return MZK_OK;
  MZK_REQUIRE(omega_powers != nullptr, "mzk_plonk_perm_products_dev: null omega powers");
  for (size_t s = 0; s < nsets; s++) {
    const mzk_perm_set& e = sets[s];
    MZK_REQUIRE(e.ncols >= 1 && e.ncols <= MZK_PERM_MAX_CHUNK && e.num && e.den,
                "mzk_plonk_perm_products_dev: set %zu: %u columns / null output", s, e.ncols);
    for (uint32_t j = 0; j < e.ncols; j++)
      MZK_REQUIRE(e.values[j] && e.sigma[j], "mzk_plonk_perm_products_dev: set %zu column %u missing", s, j);
  }
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  for (size_t s = 0; s < nsets; s++) {
    const mzk_perm_set& e = sets[s];
    PermArgs A;
    memset(&A, 0, sizeof A);
    for (uint32_t j = 0; j < e.ncols; j++) {
      A.val[j] = (const uint32_t*)e.values[j];
      A.sig[j] = (const uint32_t*)e.sigma[j];
      A.delta_beta[j] = to_arg(e.delta_beta[j]);
    }
    A.beta = to_arg(e.beta);
    A.gamma = to_arg(e.gamma);
    A.omega = (const uint32_t*)omega_powers;
    A.num = (uint32_t*)e.num;
    A.den = (uint32_t*)e.den;
    A.n = n;
    A.ncols = e.ncols;
    MZK_LAUNCH(c, "k_perm_products", k_perm_products, dim3(grid_for(n, 256)), dim3(256), 0, A);
  }
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

int mzk_plonk_numerator_dev(mzk_ctx* c, const mzk_plonk_numerator* a, void* out) {
  MZK_REQUIRE(c != nullptr && a != nullptr && out != nullptr, "mzk_plonk_numerator_dev: null argument");
  const uint64_t N = a->n_ext;
  MZK_REQUIRE(N >= 2 && (N & (N - 1)) == 0, "mzk_plonk_numerator_dev: n_ext must be a power of two");
  MZK_REQUIRE(a->nsets >= 1 && a->nsets <= 4 && a->ncols >= 1 && a->ncols <= 8 && a->chunk >= 1 &&
                  (size_t)a->nsets * a->chunk >= a->ncols && (size_t)(a->nsets - 1) * a->chunk < a->ncols,
              "mzk_plonk_numerator_dev: permutation shape (%u columns, %u sets of %u) unsupported", a->ncols,
              a->nsets, a->chunk);
  MZK_REQUIRE(a->t_count >= 1 && a->t_count <= 16 && (a->t_count & (a->t_count - 1)) == 0,
              "mzk_plonk_numerator_dev: t_count must be a power of two <= 16");
  MZK_REQUIRE(a->rot_next < N && a->rot_last < N, "mzk_plonk_numerator_dev: rotation out of range");
  const void* need[] = {a->a, a->b, a->c, a->qm, a->ql, a->qr, a->qo, a->qc, a->l0, a->l_last, a->l_active,
                        a->coset_x};
  for (const void* p : need) MZK_REQUIRE(p != nullptr, "mzk_plonk_numerator_dev: null vector");
  for (uint32_t j = 0; j < a->ncols; j++)
    MZK_REQUIRE(a->col[j] && a->sigma[j], "mzk_plonk_numerator_dev: null permutation column %u", j);
  for (uint32_t s = 0; s < a->nsets; s++) MZK_REQUIRE(a->z[s], "mzk_plonk_numerator_dev: null z %u", s);
  NumArgs A;
  memset(&A, 0, sizeof A);
  A.N = N;
  A.rot_next = a->rot_next;
  A.rot_last = a->rot_last;
  A.ncols = a->ncols;
  A.nsets = a->nsets;
  A.chunk = a->chunk;
  A.tmask = a->t_count - 1;
  A.a = (const uint32_t*)a->a;
  A.b = (const uint32_t*)a->b;
  A.c = (const uint32_t*)a->c;
  A.qm = (const uint32_t*)a->qm;
  A.ql = (const uint32_t*)a->ql;
  A.qr = (const uint32_t*)a->qr;
  A.qo = (const uint32_t*)a->qo;
  A.qc = (const uint32_t*)a->qc;
  for (uint32_t j = 0; j < a->ncols; j++) {
    A.col[j] = (const uint32_t*)a->col[j];
    A.sigma[j] = (const uint32_t*)a->sigma[j];
    A.delta_beta[j] = to_arg(a->delta_beta[j]);
  }
  for (uint32_t s = 0; s < a->nsets; s++) A.z[s] = (const uint32_t*)a->z[s];
  A.l0 = (const uint32_t*)a->l0;
  A.l_last = (const uint32_t*)a->l_last;
  A.l_active = (const uint32_t*)a->l_active;
  A.coset_x = (const uint32_t*)a->coset_x;
  A.y = to_arg(a->y);
  A.beta = to_arg(a->beta);
  A.gamma = to_arg(a->gamma);
  for (uint32_t t = 0; t < a->t_count; t++) A.t_evals[t] = to_arg(a->t_evals[t]);
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  MZK_LAUNCH(c, "k_plonk_numerator", k_plonk_numerator, dim3(grid_for(N, 256)), dim3(256), 0, A, (uint32_t*)out);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

int mzk_fr_lincomb_dev(mzk_ctx* c, const void* const* polys, const size_t* lens, const uint64_t* scalars,
                       size_t np, void* out, size_t out_len) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_REQUIRE(np <= (size_t)LC_MAX, "mzk_fr_lincomb_dev: at most %d polynomials", LC_MAX);
  if (out_len == 0) return MZK_OK;
  MZK_REQUIRE(out != nullptr && (np == 0 || (polys && lens && scalars)), "mzk_fr_lincomb_dev: null argument");
  LcArgs A;
  memset(&A, 0, sizeof A);
  A.np = (uint32_t)np;
  for (size_t j = 0; j < np; j++) {
    MZK_REQUIRE(lens[j] == 0 || polys[j] != nullptr, "mzk_fr_lincomb_dev: polynomial %zu has no data", j);
    A.p[j] = (const uint32_t*)polys[j];
    A.len[j] = lens[j] < out_len ? lens[j] : out_len;
    A.s[j] = to_arg(scalars + 4 * j);
  }
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  MZK_LAUNCH(c, "k_lincomb", k_lincomb, dim3(grid_for(out_len, 256)), dim3(256), 0, A, (uint32_t*)out, out_len);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

int mzk_fr_powers(mzk_ctx* c, const uint64_t x[4], size_t n, void* out) {
  MZK_REQUIRE(c != nullptr && x != nullptr, "mzk_fr_powers: null argument");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(out != nullptr, "mzk_fr_powers: null data");
  MZK_TRY(vec_ready(c));
  std::lock_guard<std::mutex> g(c->mu);
  host::Fr hx;
  memcpy(hx.l, x, 32);
  MZK_LAUNCH(c, "k_powers", k_powers, dim3(grid_for(n, 256)), dim3(256), 0, (uint32_t*)out, n, make_pows(hx));
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

}  // extern "C"
