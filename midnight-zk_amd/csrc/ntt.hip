// Radix-2 NTTs over the BLS12-381 scalar field for gfx950, behind the
// EvaluationDomain boundary (proofs/src/poly/domain.rs:54-414) and best_fft /
// best_fft_with_twiddles / fft_coeff_to_extended (curves/src/fft.rs:31-186).
//
// A transform of size N = N1 * N2 (N1 = 2^ceil(logN/2), N2 = 2^floor(logN/2),
// both <= 2^11) runs as two passes of the Bailey four-step method:
//   pass 1: for each n1 < N1, a size-N2 DFT of the column a[n1 + N1 n2], times the
//           twist omega_N^(n1 k2), stored at tmp[n1 + N1 k2];
//   pass 2: for each k2 < N2, a size-N1 DFT of the row tmp[k2 N1 + n1], stored at
//           out[k2 + N2 k1] -- natural order, so no bit-reversal pass exists.
// Above 2^22, N = A B C (each <= 2^11, input index nA + A nB + A B nC) runs as three
// passes, the column DFT of size M = B C itself a four-step:
//   pass 1: size-C DFTs over nC for each (nA, nB), twist omega_M^(nB kC), into tmp
//           at the same positions (nC -> kC);
//   pass 2: size-B DFTs over nB for each (nA, kC), in place in tmp (nB -> kB), twist
//           omega_N^(nA kM) with kM = kC + C kB;
//   pass 3: size-A DFTs over nA (contiguous rows of tmp) for each kM, stored at
//           out[kM + M kA].
// Each sub-DFT is staged through LDS (up to 2^11 elements = 72 KiB per workgroup):
// loaded in bit-reversed order, log2(L) Cooley-Tukey stages, stored naturally.
// The zeta-coset distribution and zero padding of coeff_to_extended are fused into
// the first pass's load; the 1/N scaling and zeta^-(i mod 3) of the inverse transforms
// into the last pass's store.  Results are the unique DFT values, hence identical to
// the reference's bit-reverse + DIT and pruned-DIF schedules.
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "field.hpp"
#include "fr29.hpp"
#include "host_ec.hpp"
#include "ntt.hpp"
#include "tune.hpp"

namespace mzk {

static constexpr int NTT_LDS_ELEMS = 1 << NTT_MAX_LOG_L;
static constexpr int TWIST_LO_BITS = 11;
static constexpr int TW29 = NTT_TW29;
static constexpr int ELEM29 = 9;  // words per element staged in LDS

__device__ __forceinline__ fr_t to_fr(const FrWords& w) {
  fr_t r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = w.v[i];
  return r;
}

__device__ __forceinline__ uint32_t bitrev(uint32_t j, uint32_t bits) {
  return bits == 0 ? 0u : (__brev(j) >> (32 - bits));
}

// 512 threads per workgroup: two 72 KiB workgroups fit a CU's LDS, so 16 waves per CU
// (4 per SIMD) instead of 8 with 256 threads; 2^22 forward 0.78 -> 0.74 ms
static constexpr uint32_t NTT_THREADS = 512;
// stages two per LDS round trip once a launch holds at least 2^21 points (measured: 2^22
// 0.75 -> 0.71 ms, seven coset 2^20 outputs 1.21 -> 1.18 ms; a lone 2^18 0.055 -> 0.069,
// 2^20 unchanged), for transforms of at least 2^16 points
static constexpr uint32_t NTT_PAIRS_MIN_LAUNCH_LOG = 21;
static constexpr uint32_t NTT_PAIRS_MIN_LOG = 16;

__device__ __forceinline__ void ntt_pass_body(const PassArgs& a, uint32_t block, uint32_t* lds);

// Workgroups go to the 8 XCDs round robin (workgroup i to XCD i mod 8), each XCD with its own
// L2.  Neighbouring sub-DFTs of the column pass read neighbouring 32-byte pieces of the same
// 128-byte lines, so the logical index is remapped to put consecutive sub-DFTs on one XCD
// (a permutation of [0, G) when 8 divides G): the other order fetched every line from HBM
// once per XCD touching it.
__device__ __forceinline__ uint32_t xcd_local_block(uint32_t b, uint32_t G) {
  return (G & 7u) ? b : (b & 7u) * (G >> 3) + (b >> 3);
}

__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass(PassArgs a) {
  extern __shared__ uint32_t lds[];   // C * L elements of ELEM29 words
  ntt_pass_body(a, xcd_local_block(blockIdx.x, gridDim.x), lds);
}

// Many transforms in one launch: workgroups [first[i], first[i+1]) run pass table[i].
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass_batch(const PassArgs* __restrict__ table,
                                                        const uint32_t* __restrict__ first, uint32_t njobs) {
  extern __shared__ uint32_t lds[];
  const uint32_t blk = xcd_local_block(blockIdx.x, gridDim.x);
  uint32_t lo = 0, hi = njobs;   // first[lo] <= blk < first[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= blk) lo = mid; else hi = mid;
  }
  const PassArgs a = table[lo];
  ntt_pass_body(a, blk - first[lo], lds);
}

// Element arithmetic of the passes is the radix-2^29 form (fr29.hpp): each butterfly's
// product is one v_mad_u64_u32 per limb product with no carry instruction.  Values enter
// a pass from canonical blst words (below 2 r after reduce_small), grow by at most 2 r
// per stage (x + t and x - t + 2 r with t below 2 r: a product of a value below 25 r and
// a table entry below r, or reduce_small of the operand where the twiddle is 1), so they
// stay below 25 r (< 2^261) through 11 stages; the store multiplies by the twist / scale
// factors where there are any and leaves in canonical blst words.
// LDS slot of element idx: ELEM29 words each, plus one word of padding per 2^NTT_LDS_SWIZZLE
// elements.  With 9-word elements alone, the load phase's bit-reversed stores put the 32 lanes
// of a half wave (indices 64 apart) in one bank, and the paired stages' stride-2 accesses
// conflicted 4 ways; the padding spreads both over the banks.
static constexpr uint32_t NTT_LDS_SWIZZLE = 5;
__device__ __forceinline__ uint32_t* elem29(uint32_t* lds, uint32_t idx) {
  return lds + ELEM29 * idx + (idx >> NTT_LDS_SWIZZLE);
}
__host__ __device__ constexpr size_t ntt_lds_bytes(size_t elems) {
  return (elems * ELEM29 + (elems >> NTT_LDS_SWIZZLE)) * 4;
}

__device__ __forceinline__ fr29 load29(const uint32_t* __restrict__ p) {
  fr29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.l[i] = p[i];
  return r;
}

__device__ __forceinline__ void store29(uint32_t* __restrict__ p, const fr29& v) {
#pragma unroll
  for (int i = 0; i < 9; i++) p[i] = v.l[i];
}

// a table record (TW29 words, 16-byte aligned): the value below r
__device__ __forceinline__ fr29 load_rec29(const uint32_t* __restrict__ p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  const uint4 a = q[0], b = q[1], c = q[2];
  fr29 r;
  r.l[0] = a.x; r.l[1] = a.y; r.l[2] = a.z; r.l[3] = a.w;
  r.l[4] = b.x; r.l[5] = b.y; r.l[6] = b.z; r.l[7] = b.w;
  r.l[8] = c.x;
  return r;
}

__device__ __forceinline__ fr29 arg29(const FrWords& w) { return fr29_reduce_full(fr29_from_std(w.v)); }

__device__ __forceinline__ void ntt_pass_body(const PassArgs& a, uint32_t block, uint32_t* lds) {
  const uint32_t L = 1u << a.log_l;
  const uint32_t CL = a.C * L;
  const uint64_t b0 = (uint64_t)block * a.C;
  const uint32_t nt = blockDim.x;

  // load (bit-reversed within each sub-DFT)
  fr29 z1, z2;
  if (a.in_coset) {
    z1 = arg29(a.zeta);
    z2 = arg29(a.zeta2);
  }
  const uint32_t lc = 31u - __clz(a.C);   // C and L are powers of two: shifts, not divisions
  for (uint32_t e = threadIdx.x; e < CL; e += nt) {
    const uint32_t c = e & (a.C - 1u), j = e >> lc;
    const uint64_t b = b0 + c;
    const uint64_t gi = (b & a.in_mask) * a.in_lo + (b >> a.in_shift) * a.in_hi + (uint64_t)j * a.in_j;
    fr29 x;
    if (gi < a.n_real) {
      const uint4* q = reinterpret_cast<const uint4*>(a.src + 8 * gi);
      const uint4 u0 = q[0], u1 = q[1];
      const uint32_t w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
      x = fr29_reduce_small(fr29_from_std(w));
      if (a.in_coset) {
        const uint32_t m = (uint32_t)(gi % 3);
        if (m == 1) x = fr29_mul(x, z1);
        else if (m == 2) x = fr29_mul(x, z2);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 9; i++) x.l[i] = 0;
    }
    store29(elem29(lds, c * L + bitrev(j, a.log_l)), x);
  }
  __syncthreads();

  // log_l radix-2 DIT stages.  For large launches they are taken two at a time in
  // registers (four elements per thread per LDS round trip; an odd first stage, twiddle 1
  // throughout, runs alone); on smaller launches the halved thread parallelism of the
  // longer per-thread chains costs more than the round trips save, so they keep one stage
  // per round trip.  Either way each butterfly is the radix-2 one -- t = y w (or
  // reduce_small(y) where w = 1), x + t, x + 2r - t -- so the values are those of
  // stage-by-stage order.
  uint32_t s = 0;
  if (!a.pairs) {
    const uint32_t halfL = L >> 1;
    for (; s < a.log_l; s++) {
      const uint32_t half = 1u << s;
      for (uint32_t u = threadIdx.x; u < a.C * halfL; u += nt) {
        const uint32_t c = u >> (a.log_l - 1), v = u & (halfL - 1u);
        const uint32_t pos = v & (half - 1);
        const uint32_t i = ((v >> s) << (s + 1)) + pos;
        uint32_t* pa = elem29(lds, c * L + i);
        uint32_t* pb = elem29(lds, c * L + i + half);
        const fr29 x = load29(pa);
        const fr29 y = load29(pb);
        const fr29 t = pos != 0 ? fr29_mul(y, load_rec29(a.tw + TW29 * (pos << (a.log_l - 1 - s))))
                                : fr29_reduce_small(y);
        store29(pa, fr29_add(x, t));
        store29(pb, fr29_sub_kr<2>(x, t));
      }
      __syncthreads();
    }
  } else if (a.log_l & 1) {
    const uint32_t halfL = L >> 1;
    for (uint32_t u = threadIdx.x; u < a.C * halfL; u += nt) {
      const uint32_t c = u >> (a.log_l - 1), v = u & (halfL - 1u);
      uint32_t* pa = elem29(lds, c * L + 2 * v);
      uint32_t* pb = elem29(lds, c * L + 2 * v + 1);
      const fr29 x = load29(pa);
      const fr29 t = fr29_reduce_small(load29(pb));
      store29(pa, fr29_add(x, t));
      store29(pb, fr29_sub_kr<2>(x, t));
    }
    __syncthreads();
    s = 1;
  }
  // Register pairs of stages.  The first stage's outputs are left unnormalised where the
  // second stage can take them (fr29.hpp: add_lazy, sub3r_lazy): y0 and y2 as limb sums, y3
  // as x2 + 3 r - t3; y1 feeds sub_kr<2> as its first operand and stays normalised.  A pair
  // starting at stage 0 has twiddle 1 in its first stage and in half of its second
  // (reduce_small where the twiddle is 1, uniformly across the wave); later pairs always
  // multiply by their table entry, whose entry 0 is 1, so no lane takes another path.
  // Values grow by at most 4 r per pair, as with stage-by-stage order.
  const uint32_t quarterL = L >> 2;
  for (; a.pairs && s + 1 < a.log_l; s += 2) {
    const uint32_t half = 1u << s;
    for (uint32_t u = threadIdx.x; u < a.C * quarterL; u += nt) {
      const uint32_t c = u >> (a.log_l - 2), v = u & (quarterL - 1u);
      const uint32_t pos = v & (half - 1);
      const uint32_t i = ((v >> s) << (s + 2)) + pos;
      uint32_t* p0 = elem29(lds, c * L + i);
      uint32_t* p1 = elem29(lds, c * L + i + half);
      uint32_t* p2 = elem29(lds, c * L + i + 2 * half);
      uint32_t* p3 = elem29(lds, c * L + i + 3 * half);
      const fr29 x0 = load29(p0), x1 = load29(p1), x2 = load29(p2), x3 = load29(p3);
      // stage s: (x0, x1) and (x2, x3), both with w = omega_(2 half)^pos; stage s + 1:
      // (y0, y2) with omega_(4 half)^pos, (y1, y3) with omega_(4 half)^(pos + half)
      fr29 y0, y1, t2, t4;
      if (s == 0) {
        const fr29 t1 = fr29_reduce_small(x1), t3 = fr29_reduce_small(x3);
        y0 = fr29_add_lazy(x0, t1);
        y1 = fr29_sub_kr<2>(x0, t1);
        t2 = fr29_reduce_small(fr29_add(x2, t3));
        t4 = fr29_mul(fr29_sub3r_lazy(x2, t3), load_rec29(a.tw + TW29 * (1u << (a.log_l - 2))));
      } else {
        const fr29 w = load_rec29(a.tw + TW29 * (pos << (a.log_l - 1 - s)));
        const fr29 t1 = fr29_mul(x1, w), t3 = fr29_mul(x3, w);
        y0 = fr29_add_lazy(x0, t1);
        y1 = fr29_sub_kr<2>(x0, t1);
        t2 = fr29_mul(fr29_add_lazy(x2, t3), load_rec29(a.tw + TW29 * (pos << (a.log_l - 2 - s))));
        t4 = fr29_mul(fr29_sub3r_lazy(x2, t3), load_rec29(a.tw + TW29 * ((pos + half) << (a.log_l - 2 - s))));
      }
      store29(p0, fr29_add(y0, t2));
      store29(p2, fr29_sub_kr<2>(y0, t2));
      store29(p1, fr29_add(y1, t4));
      store29(p3, fr29_sub_kr<2>(y1, t4));
    }
    __syncthreads();
  }

  // store (natural order), fused twist / scaling, back to canonical blst words
  fr29 p0, p1, p2;
  if (a.out_mode != 0) {
    p0 = arg29(a.post[0]);
    p1 = arg29(a.post[1]);
    p2 = arg29(a.post[2]);
  }
  for (uint32_t e = threadIdx.x; e < CL; e += nt) {
    const uint32_t c = e & (a.C - 1u), k = e >> lc;
    const uint64_t b = b0 + c;
    fr29 x = load29(elem29(lds, c * L + k));
    if (a.do_twist) {
      const uint64_t ex = (b & a.tw_mask) * ((b >> a.tw_shift) + ((uint64_t)k << a.tw_kshift));   // < N
      if (ex != 0) {
        const fr29 w = fr29_mul(load_rec29(a.twist_lo + TW29 * (ex & ((1u << TWIST_LO_BITS) - 1))),
                                load_rec29(a.twist_hi + TW29 * (ex >> TWIST_LO_BITS)));
        x = fr29_mul(x, w);
      }
    }
    const uint64_t go = (b & a.out_mask) * a.out_lo + (b >> a.out_shift) * a.out_hi + (uint64_t)k * a.out_k;
    if (a.out_mode == 1) {
      x = fr29_mul(x, p0);
    } else if (a.out_mode == 2) {
      const uint32_t m = (uint32_t)(go % 3);
      x = fr29_mul(x, m == 0 ? p0 : (m == 1 ? p1 : p2));
    }
    uint32_t o[8];
    fr29_to_std(x, o);
    uint4* q = reinterpret_cast<uint4*>(a.dst + 8 * go);
    q[0] = make_uint4(o[0], o[1], o[2], o[3]);
    q[1] = make_uint4(o[4], o[5], o[6], o[7]);
  }
}

// out[i] = rec29(g^i) for i < count, with g^i = prod_{j : bit j of i} pows[j]
// (pows[j] = g^(2^j)): table records of TW29 words, the value below r in the radix-2^29
// form, then zero padding.
struct PowArgs {
  FrWords pows[32];
};
__global__ void k_pow_rec29(uint32_t* __restrict__ out, uint64_t count, PowArgs p) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  fr_t r = fr_t::one();
  for (int j = 0; j < 32; j++)
    if ((i >> j) & 1) r = mul(r, to_fr(p.pows[j]));
  const fr29 v = fr29_reduce_full(fr29_from_std(r.v));
  uint32_t* o = out + TW29 * i;
#pragma unroll
  for (int k = 0; k < 9; k++) o[k] = v.l[k];
#pragma unroll
  for (int k = 9; k < TW29; k++) o[k] = 0;
}

// out[i] = g^i in blst words (compute_twiddles)
__global__ void k_pow_table(uint32_t* __restrict__ out, uint64_t count, PowArgs p) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  fr_t r = fr_t::one();
  for (int j = 0; j < 32; j++)
    if ((i >> j) & 1) r = mul(r, to_fr(p.pows[j]));
  store<FrP>(out + 8 * i, r);
}

// a[i] *= t[i & mask]
__global__ void k_mul_periodic(uint32_t* __restrict__ a, size_t n, const uint32_t* __restrict__ t, uint32_t mask) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  store<FrP>(a + 8 * i, mul(load<FrP>(a + 8 * i), load<FrP>(t + 8 * (i & mask))));
}

// ----------------------------------------------------------------------- host
static FrWords words(const host::Fr& f) {
  FrWords w;
  memcpy(w.v, f.l, 32);
  return w;
}

static PowArgs pow_args(const host::Fr& g) {
  PowArgs p;
  host::Fr x = g;
  for (int j = 0; j < 32; j++) {
    p.pows[j] = words(x);
    x = host::fr_mul(x, x);
  }
  return p;
}

static int pow_records(hipStream_t st, uint32_t* out, uint64_t count, const host::Fr& g) {
  hipLaunchKernelGGL(k_pow_rec29, dim3(grid_for(count, 256)), dim3(256), 0, st, out, count, pow_args(g));
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

NttPlan::~NttPlan() {
  if (mem) (void)hipFree(mem);
}

void NttPlan::split(uint32_t log_n, uint32_t* np_out, uint32_t lp_out[3]) {
  lp_out[0] = lp_out[1] = lp_out[2] = 0;
  if (log_n <= (uint32_t)NTT_MAX_LOG_L) {
    *np_out = 1;
    lp_out[0] = log_n;
  } else if (log_n <= 2u * NTT_MAX_LOG_L) {
    // rows (pass 2) of 2^l1, columns (pass 1) of 2^(log_n - l1)
    uint32_t l1 = (log_n + 1) / 2;
    const NttTune tn = tune_ntt_get(log_n);
    if (tn.l1 && tn.l1 <= (uint32_t)NTT_MAX_LOG_L && log_n > tn.l1 && log_n - tn.l1 <= (uint32_t)NTT_MAX_LOG_L)
      l1 = tn.l1;
    *np_out = 2;
    lp_out[0] = log_n - l1;
    lp_out[1] = l1;
  } else {
    // A (rows, pass 3) >= B (pass 2) >= C (columns, pass 1), as even as the sizes allow
    const uint32_t la = (log_n + 2) / 3;
    const uint32_t lb = (log_n - la + 1) / 2;
    *np_out = 3;
    lp_out[0] = log_n - la - lb;
    lp_out[1] = lb;
    lp_out[2] = la;
  }
}

size_t NttPlan::records(uint32_t log_n, uint32_t np_, const uint32_t lp_[3]) {
  size_t r = 0;
  for (uint32_t p = 0; p < np_; p++) r += std::max<size_t>(((size_t)1 << lp_[p]) / 2, 1);
  if (np_ >= 2) r += ((size_t)1 << std::min<uint32_t>(TWIST_LO_BITS, log_n)) + std::max<size_t>(1, ((size_t)1 << log_n) >> TWIST_LO_BITS);
  return r;
}

int NttPlan::init_at(hipStream_t st, uint32_t log_n_, const host::Fr& omega, uint32_t* tables, size_t cap_records) {
  log_n = log_n_;
  split(log_n, &np, lp);
  const size_t need = records(log_n, np, lp);
  if (need > cap_records) {
    set_error("ntt plan 2^%u: %zu table records needed, %zu available", log_n, need, cap_records);
    return MZK_ERR_INVALID;
  }
  uint32_t* p = tables;
  for (uint32_t q = 0; q < np; q++) {
    // omega_{L_q} = omega^(N / L_q)
    host::Fr w = omega;
    for (uint32_t i = lp[q]; i < log_n; i++) w = host::fr_mul(w, w);
    const size_t cnt = std::max<size_t>(((size_t)1 << lp[q]) / 2, 1);
    MZK_TRY(pow_records(st, p, cnt, w));
    tw[q] = p;
    p += TW29 * cnt;
  }
  if (np >= 2) {
    const size_t lo_cnt = (size_t)1 << std::min<uint32_t>(TWIST_LO_BITS, log_n);
    const size_t hi_cnt = std::max<size_t>(1, ((size_t)1 << log_n) >> TWIST_LO_BITS);
    host::Fr whi = omega;
    for (int i = 0; i < TWIST_LO_BITS; i++) whi = host::fr_mul(whi, whi);
    MZK_TRY(pow_records(st, p, lo_cnt, omega));
    twist_lo = p;
    p += TW29 * lo_cnt;
    MZK_TRY(pow_records(st, p, hi_cnt, whi));
    twist_hi = p;
  }
  return MZK_OK;
}

int NttPlan::init(hipStream_t st, uint32_t log_n_, const host::Fr& omega) {
  uint32_t np_, lp_[3];
  split(log_n_, &np_, lp_);
  const size_t recs = records(log_n_, np_, lp_);
  if (hipMalloc(&mem, recs * TW29 * 4) != hipSuccess) {
    mem = nullptr;
    set_error("ntt plan 2^%u: out of device memory for twiddle tables", log_n_);
    return MZK_ERR_NOMEM;
  }
  return init_at(st, log_n_, omega, (uint32_t*)mem, recs);
}

// 2^11 staged elements of 9 words take 72 KiB of LDS, above the 64 KiB a kernel gets
// without asking (gfx950 gives a workgroup up to 160 KiB)
static int ntt_lds_attr() {
  static std::once_flag once;
  static int rc = MZK_OK;
  std::call_once(once, [] {
    const int bytes = (int)ntt_lds_bytes(NTT_LDS_ELEMS);
    if (hipFuncSetAttribute((const void*)k_ntt_pass, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
            hipSuccess ||
        hipFuncSetAttribute((const void*)k_ntt_pass_batch, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
            hipSuccess) {
      set_error("ntt: cannot raise the LDS limit to %d bytes", bytes);
      rc = MZK_ERR_HIP;
    }
  });
  return rc;
}

static uint32_t choose_c(uint32_t L, uint64_t nbatch, uint32_t groups) {
  uint32_t cmax = NTT_LDS_ELEMS / L;
  uint64_t want = nbatch / (groups ? groups : 1024);   // keep >= ~1024 workgroups when the size allows
  uint32_t c = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cmax, want));
  while (nbatch % c) c >>= 1;
  return std::max(c, 1u);
}

static uint32_t pass_threads(const PassArgs& a) {
  const uint32_t L = 1u << a.log_l;
  return std::min<uint32_t>(NTT_THREADS, std::max<uint32_t>(64, a.C * L / 2));
}

static int launch_pass(mzk_ctx* c, const char* name, const PassArgs& a, uint64_t groups) {
  const size_t lds = ntt_lds_bytes((size_t)a.C << a.log_l);
  MZK_TRY(ntt_lds_attr());
  MZK_LAUNCH(c, name, k_ntt_pass, dim3((unsigned)groups), dim3(pass_threads(a)), lds, a);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

int ntt_launch_pass_raw(hipStream_t st, const PassArgs& a, uint64_t groups) {
  const size_t lds = ntt_lds_bytes((size_t)a.C << a.log_l);
  MZK_TRY(ntt_lds_attr());
  hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)groups), dim3(pass_threads(a)), lds, st, a);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

// contiguous batches: sub-DFT b starts at b * lo (the masks keep all of b, the shifts none)
static void linear_addressing(PassArgs& a, uint64_t in_lo, uint64_t in_j, uint64_t out_lo, uint64_t out_k) {
  a.in_mask = ~0ull;
  a.in_shift = 63;
  a.in_hi = 0;
  a.in_lo = in_lo;
  a.in_j = in_j;
  a.out_mask = ~0ull;
  a.out_shift = 63;
  a.out_hi = 0;
  a.out_lo = out_lo;
  a.out_k = out_k;
  a.tw_mask = ~0ull;
  a.tw_shift = 63;
  a.tw_kshift = 0;
}

int ntt_build_passes(const NttPlan& plan, const uint32_t* src, uint32_t* dst, uint32_t* tmp, uint64_t n_real,
                     int in_coset, int out_mode, const host::Fr& scale, uint64_t launch_elems, PassArgs pa[3],
                     uint64_t groups[3]) {
  const uint32_t log_n = plan.log_n;
  const uint64_t N = 1ull << log_n;
  host::Fr z = host::fr_const(host::kFrZeta);
  host::Fr z2 = host::fr_mul(z, z);
  PassArgs a;
  memset(&a, 0, sizeof a);
  a.zeta = words(z);
  a.zeta2 = words(z2);
  a.post[0] = words(scale);
  a.post[1] = words(host::fr_mul(scale, z2));
  a.post[2] = words(host::fr_mul(scale, z));
  a.pairs = log_n >= NTT_PAIRS_MIN_LOG && launch_elems >= (1ull << NTT_PAIRS_MIN_LAUNCH_LOG) ? 1 : 0;
  const uint32_t groups_target = plan.np == 2 ? tune_ntt_get(log_n).groups : 0;
  if (plan.np == 1) {
    a.src = src;
    a.dst = dst;
    a.tw = plan.tw[0];
    linear_addressing(a, 0, 1, 0, 1);
    a.n_real = n_real;
    a.in_coset = in_coset;
    a.log_l = log_n;
    a.C = 1;
    a.out_mode = out_mode;
    pa[0] = a;
    groups[0] = 1;
    return 1;
  }
  if (plan.np == 2) {
    const uint64_t N1 = 1ull << plan.lp[1], N2 = 1ull << plan.lp[0];
    // pass 1: columns (size N2), twist omega_N^(n1 k2), out-of-place into tmp
    a.src = src;
    a.dst = tmp;
    a.tw = plan.tw[0];
    a.twist_lo = plan.twist_lo;
    a.twist_hi = plan.twist_hi;
    a.do_twist = 1;
    linear_addressing(a, 1, N1, 1, N1);
    a.n_real = n_real;
    a.in_coset = in_coset;
    a.log_l = plan.lp[0];
    a.C = choose_c((uint32_t)N2, N1, groups_target);
    a.out_mode = 0;
    pa[0] = a;
    groups[0] = N1 / a.C;
    // pass 2: rows (size N1), natural-order output
    PassArgs b = a;
    b.src = tmp;
    b.dst = dst;
    b.tw = plan.tw[1];
    b.do_twist = 0;
    b.in_coset = 0;
    b.n_real = N;
    linear_addressing(b, N1, 1, 1, N2);
    b.log_l = plan.lp[1];
    b.C = choose_c((uint32_t)N1, N2, groups_target);
    b.out_mode = out_mode;
    pa[1] = b;
    groups[1] = N2 / b.C;
    return 2;
  }
  // three passes: N = A B C, input index nA + A nB + A B nC
  const uint32_t lc = plan.lp[0], lb = plan.lp[1], la = plan.lp[2];
  const uint64_t A = 1ull << la, B = 1ull << lb, Cs = 1ull << lc, AB = A * B, M = B * Cs;
  // pass 1: size-C DFTs over nC, sub-DFT b = nA + A nB; twist omega_M^(nB kC) =
  // omega_N^(A nB kC), i.e. e = (b & ~(A - 1)) k; out to tmp at the same positions
  a.src = src;
  a.dst = tmp;
  a.tw = plan.tw[0];
  a.twist_lo = plan.twist_lo;
  a.twist_hi = plan.twist_hi;
  a.do_twist = 1;
  linear_addressing(a, 1, AB, 1, AB);
  a.tw_mask = ~(A - 1);
  a.n_real = n_real;
  a.in_coset = in_coset;
  a.log_l = lc;
  a.C = choose_c((uint32_t)Cs, AB, 0);
  a.out_mode = 0;
  pa[0] = a;
  groups[0] = AB / a.C;
  // pass 2: size-B DFTs over nB, sub-DFT b = nA + A kC at nA + A B kC + A j, in place;
  // twist omega_N^(nA (kC + C kB)), e = (b & (A - 1)) ((b >> la) + (k << lc))
  PassArgs b = a;
  b.src = tmp;
  b.dst = tmp;
  b.tw = plan.tw[1];
  b.in_mask = A - 1;
  b.in_lo = 1;
  b.in_shift = la;
  b.in_hi = AB;
  b.in_j = A;
  b.out_mask = A - 1;
  b.out_lo = 1;
  b.out_shift = la;
  b.out_hi = AB;
  b.out_k = A;
  b.tw_mask = A - 1;
  b.tw_shift = la;
  b.tw_kshift = lc;
  b.n_real = N;
  b.in_coset = 0;
  b.log_l = lb;
  b.C = choose_c((uint32_t)B, A * Cs, 0);
  pa[1] = b;
  groups[1] = A * Cs / b.C;
  // pass 3: size-A DFTs over nA, sub-DFT b = kM = kC + C kB, a contiguous row of tmp at
  // A B kC + A kB; natural-order output out[kM + M kA]
  PassArgs d = b;
  d.src = tmp;
  d.dst = dst;
  d.tw = plan.tw[2];
  d.do_twist = 0;
  linear_addressing(d, 0, 1, 1, M);
  d.in_mask = Cs - 1;
  d.in_lo = AB;
  d.in_shift = lc;
  d.in_hi = A;
  d.log_l = la;
  d.C = choose_c((uint32_t)A, M, 0);
  d.out_mode = out_mode;
  pa[2] = d;
  groups[2] = M / d.C;
  return 3;
}

static const char* pass_name(int np, int i) {
  static const char* names[6] = {"k_ntt_pass_single", "k_ntt_pass1", "k_ntt_pass2",
                                 "k_ntt_pass3a", "k_ntt_pass3b", "k_ntt_pass3c"};
  return np == 1 ? names[0] : (np == 2 ? names[1 + i] : names[3 + i]);
}

int ntt_run(mzk_ctx* c, const NttPlan& plan, const uint32_t* src, uint32_t* dst, uint64_t n_real, int in_coset,
            int out_mode, const host::Fr& scale) {
  const uint64_t N = 1ull << plan.log_n;
  uint32_t* tmp = nullptr;
  if (plan.np > 1) {
    MZK_TRY(c->scratch[SCR_NTT].reserve(N * 32));
    tmp = (uint32_t*)c->scratch[SCR_NTT].ptr;
  }
  PassArgs pa[3];
  uint64_t groups[3];
  const int np = ntt_build_passes(plan, src, dst, tmp, n_real, in_coset, out_mode, scale, N, pa, groups);
  for (int i = 0; i < np; i++) MZK_TRY(launch_pass(c, pass_name(np, i), pa[i], groups[i]));
  return MZK_OK;
}

// Cached plans on a context, keyed by size, split and root.  Trimmed only at the start
// of a call (ntt_plans_trim), so the plans a call looked up stay valid until it returns.
static constexpr size_t MAX_CTX_PLANS = 64;

static void ntt_plans_trim(mzk_ctx* c) {
  if (c->ntt_plans.size() <= MAX_CTX_PLANS) return;
  (void)hipStreamSynchronize(c->stream);   // queued passes may still read the tables
  free_ntt_plans(c);
}

static const NttPlan* ctx_plan_root(mzk_ctx* c, uint32_t log_n, const host::Fr& omega) {
  uint32_t np, lp[3];
  NttPlan::split(log_n, &np, lp);
  GraphKey k;
  k.add(log_n).add(lp[0]).add(lp[1]).add(lp[2]).add(omega.l);
  auto it = c->ntt_plans.find(k.s);
  if (it != c->ntt_plans.end()) return it->second;
  NttPlan* p = new NttPlan();
  if (p->init(c->stream, log_n, omega) != MZK_OK) {
    delete p;
    return nullptr;
  }
  c->ntt_plans[k.s] = p;
  return p;
}

// omega = ROOT^(2^(32 - log_n)) or its inverse
const NttPlan* ctx_plan(mzk_ctx* c, uint32_t log_n, int inverse) {
  host::Fr w = host::fr_root_of_unity(log_n);
  if (inverse) w = host::fr_inv(w);
  return ctx_plan_root(c, log_n, w);
}

void free_ntt_plans(mzk_ctx* c) {
  for (auto& kv : c->ntt_plans) delete kv.second;
  c->ntt_plans.clear();
}

// Runs many transforms with up to three launches in all: every job's first pass, then
// every multi-pass job's second pass, then every three-pass job's third pass (each
// workgroup finds its job in a table).
int ntt_run_batch(mzk_ctx* c, const mzk_ntt_job* jobs, size_t njobs) {
  ntt_plans_trim(c);
  size_t tmp_elems = 0, all_elems = 0;
  for (size_t j = 0; j < njobs; j++) {
    if (jobs[j].log_n > (uint32_t)NTT_MAX_LOG_L) tmp_elems += 1ull << jobs[j].log_n;
    all_elems += 1ull << jobs[j].log_n;
  }
  MZK_TRY(c->scratch[SCR_NTT].reserve(tmp_elems * 32 + 64));
  uint32_t* tmp = (uint32_t*)c->scratch[SCR_NTT].ptr;
  std::vector<PassArgs> t[3];
  std::vector<uint32_t> f[3];
  for (int p = 0; p < 3; p++) f[p].push_back(0);
  for (size_t j = 0; j < njobs; j++) {
    const mzk_ntt_job& jb = jobs[j];
    const int inverse = (jb.flags & MZK_NTT_INVERSE) != 0;
    const int coset = (jb.flags & MZK_NTT_COSET) != 0;
    const NttPlan* plan = ctx_plan(c, jb.log_n, inverse);
    MZK_REQUIRE(plan != nullptr, "ntt batch: plan for 2^%u failed", jb.log_n);
    const uint64_t N = 1ull << jb.log_n;
    host::Fr scale = inverse ? host::fr_inv(host::fr_from_u64(N)) : host::fr_one();
    PassArgs pa[3];
    uint64_t groups[3];
    const int np = ntt_build_passes(*plan, (const uint32_t*)jb.d_in, (uint32_t*)jb.d_out, tmp,
                                    inverse ? N : jb.n_real, inverse ? 0 : coset, inverse ? (coset ? 2 : 1) : 0,
                                    scale, all_elems, pa, groups);
    if (np > 1) tmp += 8 * N;
    for (int p = 0; p < np; p++) {
      MZK_REQUIRE((uint64_t)f[p].back() + groups[p] < (1ull << 31), "ntt batch: too many workgroups");
      t[p].push_back(pa[p]);
      f[p].push_back(f[p].back() + (uint32_t)groups[p]);
    }
  }
  // tables: PassArgs and first-workgroup arrays, copied once into the batch scratch
  size_t bytes = 64;
  for (int p = 0; p < 3; p++) bytes += t[p].size() * sizeof(PassArgs) + f[p].size() * 4 + 32;
  MZK_TRY(c->scratch[SCR_IN1].reserve(bytes));
  char* d = (char*)c->scratch[SCR_IN1].ptr;
  std::vector<char> h(bytes);
  size_t o = 0;
  auto put = [&](const void* src, size_t n) {
    memcpy(h.data() + o, src, n);
    size_t at = o;
    o = (o + n + 15) & ~(size_t)15;
    return d + at;
  };
  PassArgs* d_t[3];
  uint32_t* d_f[3];
  for (int p = 0; p < 3; p++) {
    d_t[p] = (PassArgs*)put(t[p].data(), t[p].size() * sizeof(PassArgs));
    d_f[p] = (uint32_t*)put(f[p].data(), f[p].size() * 4);
  }
  MZK_HIP(hipMemcpyAsync(d, h.data(), o, hipMemcpyHostToDevice, c->stream));
  const size_t lds = ntt_lds_bytes(NTT_LDS_ELEMS);
  MZK_TRY(ntt_lds_attr());
  // the launches as one cached graph (the pass tables are copied before it runs)
  uint32_t n[3], g[3];
  GraphKey gk;
  gk.add('N').add(lds);
  for (int p = 0; p < 3; p++) {
    n[p] = (uint32_t)t[p].size();
    g[p] = f[p].back();
    gk.add(d_t[p]).add(d_f[p]).add(n[p]).add(g[p]);
  }
  static const char* names[3] = {"k_ntt_pass_batch1", "k_ntt_pass_batch2", "k_ntt_pass_batch3"};
  return run_graph(c, gk.s, [&]() -> int {
    for (int p = 0; p < 3; p++)
      if (n[p]) MZK_LAUNCH(c, names[p], k_ntt_pass_batch, dim3(g[p]), dim3(NTT_THREADS), lds, d_t[p], d_f[p], n[p]);
    MZK_HIP(hipGetLastError());
    return MZK_OK;
  });
}

}  // namespace mzk

// ----------------------------------------------------------------------- domain
struct mzk_domain {
  mzk_ctx* ctx = nullptr;   // runs the transforms; never read at destroy
  int device = 0;
  uint32_t k = 0, ext_k = 0;
  mzk::NttPlan fwd, inv, ext_fwd, ext_inv;
  mzk::host::Fr n_inv, ext_n_inv;
  uint32_t* t_evals = nullptr;   // 2^(ext_k - k) values of 1 / ((zeta w_ext^i)^n - 1)
  uint32_t t_count = 0;
};

using namespace mzk;

static int domain_check(mzk_domain* d) {
  MZK_REQUIRE(d != nullptr && d->ctx != nullptr, "null domain");
  return MZK_OK;
}

extern "C" {

int mzk_domain_create(mzk_ctx* c, uint32_t j, uint32_t k, mzk_domain** out) {
  MZK_REQUIRE(c != nullptr && out != nullptr, "mzk_domain_create: null argument");
  MZK_REQUIRE(j >= 1, "mzk_domain_create: j must be >= 1");
  MZK_REQUIRE(k >= 1 && k <= 32, "mzk_domain_create: k out of range");
  // extended_k (domain.rs:64-67)
  const uint64_t n = 1ull << k;
  const uint64_t qd = (uint64_t)(j > 1 ? j - 1 : 0);
  uint32_t ek = k;
  while (ek < 40 && (1ull << ek) < n * qd) ek++;
  MZK_REQUIRE(ek <= (uint32_t)NTT_MAX_LOG_N, "mzk_domain_create: extended_k %u exceeds S = 32", ek);
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  mzk_domain* d = new mzk_domain();
  d->ctx = c;
  d->device = c->device;
  d->k = k;
  d->ext_k = ek;
  host::Fr w_ext = host::fr_root_of_unity(ek);
  host::Fr w = w_ext;
  for (uint32_t i = k; i < ek; i++) w = host::fr_mul(w, w);
  host::Fr w_inv = host::fr_inv(w), w_ext_inv = host::fr_inv(w_ext);
  int rc = d->fwd.init(c->stream, k, w);
  if (rc == MZK_OK) rc = d->inv.init(c->stream, k, w_inv);
  if (rc == MZK_OK) rc = d->ext_fwd.init(c->stream, ek, w_ext);
  if (rc == MZK_OK) rc = d->ext_inv.init(c->stream, ek, w_ext_inv);
  if (rc != MZK_OK) {
    delete d;
    return rc;
  }
  d->n_inv = host::fr_inv(host::fr_from_u64(n));
  d->ext_n_inv = host::fr_inv(host::fr_from_u64(1ull << ek));
  // t_evaluations (domain.rs:104-140)
  d->t_count = 1u << (ek - k);
  std::vector<uint64_t> t(4 * d->t_count);
  host::Fr zeta = host::fr_const(host::kFrZeta);
  host::Fr orig = host::fr_pow(zeta, n), step = host::fr_pow(w_ext, n), cur = orig;
  for (uint32_t i = 0; i < d->t_count; i++) {
    host::Fr v = host::fr_inv(host::fr_sub(cur, host::fr_one()));
    memcpy(&t[4 * i], v.l, 32);
    cur = host::fr_mul(cur, step);
  }
  hipError_t e = hipMalloc(&d->t_evals, 32 * (size_t)d->t_count);
  if (e != hipSuccess) {
    delete d;
    set_error("hipMalloc: %s", hipGetErrorString(e));
    return MZK_ERR_NOMEM;
  }
  MZK_HIP(hipMemcpyAsync(d->t_evals, t.data(), 32 * (size_t)d->t_count, hipMemcpyHostToDevice, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  *out = d;
  return MZK_OK;
}

// The creating context may already be destroyed: wait for the device, touching no context.
void mzk_domain_destroy(mzk_domain* d) {
  if (!d) return;
  (void)hipSetDevice(d->device);
  (void)hipDeviceSynchronize();
  if (d->t_evals) (void)hipFree(d->t_evals);
  delete d;
}

int mzk_domain_info(const mzk_domain* d, uint32_t* k, uint32_t* ek) {
  MZK_REQUIRE(d != nullptr, "null domain");
  if (k) *k = d->k;
  if (ek) *ek = d->ext_k;
  return MZK_OK;
}

// Transform bodies; callers hold the context lock.
static int l2c_impl(mzk_domain* d, const void* in, void* out) {
  return ntt_run(d->ctx, d->inv, (const uint32_t*)in, (uint32_t*)out, 1ull << d->k, 0, 1, d->n_inv);
}
static int c2l_impl(mzk_domain* d, const void* in, void* out) {
  return ntt_run(d->ctx, d->fwd, (const uint32_t*)in, (uint32_t*)out, 1ull << d->k, 0, 0, host::fr_one());
}
static int c2e_impl(mzk_domain* d, const void* in, void* out) {
  return ntt_run(d->ctx, d->ext_fwd, (const uint32_t*)in, (uint32_t*)out, 1ull << d->k, 1, 0, host::fr_one());
}
static int e2c_impl(mzk_domain* d, const void* in, void* out) {
  return ntt_run(d->ctx, d->ext_inv, (const uint32_t*)in, (uint32_t*)out, 1ull << d->ext_k, 0, 2, d->ext_n_inv);
}

typedef int (*transform_fn)(mzk_domain*, const void*, void*);

static int run_dev(mzk_domain* d, const void* in, void* out, transform_fn fn) {
  MZK_TRY(domain_check(d));
  MZK_REQUIRE(in != nullptr && out != nullptr, "null data");
  std::lock_guard<std::mutex> g(d->ctx->mu);
  MZK_HIP(hipSetDevice(d->ctx->device));
  return fn(d, in, out);
}

// Host-pointer form: stage through context scratch, run, copy back, synchronise,
// all under the context lock.
static int run_host(mzk_domain* d, const uint64_t* in, uint64_t* out, size_t in_elems, size_t out_elems,
                    transform_fn fn) {
  MZK_TRY(domain_check(d));
  MZK_REQUIRE(in != nullptr && out != nullptr, "null data");
  mzk_ctx* c = d->ctx;
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  MZK_TRY(c->scratch[SCR_IN0].reserve(in_elems * 32));
  MZK_TRY(c->scratch[SCR_IN1].reserve(out_elems * 32));
  MZK_HIP(hipMemcpyAsync(c->scratch[SCR_IN0].ptr, in, in_elems * 32, hipMemcpyHostToDevice, c->stream));
  MZK_TRY(fn(d, c->scratch[SCR_IN0].ptr, c->scratch[SCR_IN1].ptr));
  MZK_HIP(hipMemcpyAsync(out, c->scratch[SCR_IN1].ptr, out_elems * 32, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  return MZK_OK;
}

int mzk_lagrange_to_coeff_dev(mzk_domain* d, void* a) { return run_dev(d, a, a, l2c_impl); }
int mzk_coeff_to_lagrange_dev(mzk_domain* d, void* a) { return run_dev(d, a, a, c2l_impl); }
int mzk_coeff_to_extended_dev(mzk_domain* d, const void* in, void* out) {
  MZK_REQUIRE(in != out, "coeff_to_extended: input and output must differ");
  return run_dev(d, in, out, c2e_impl);
}
int mzk_extended_to_coeff_dev(mzk_domain* d, void* a) { return run_dev(d, a, a, e2c_impl); }

int mzk_divide_by_vanishing_poly_dev(mzk_domain* d, void* a) {
  MZK_TRY(domain_check(d));
  MZK_REQUIRE(a != nullptr, "null data");
  std::lock_guard<std::mutex> g(d->ctx->mu);
  MZK_HIP(hipSetDevice(d->ctx->device));
  const size_t N = 1ull << d->ext_k;
  mzk_ctx* c = d->ctx;
  MZK_LAUNCH(c, "k_mul_periodic", k_mul_periodic, dim3(grid_for(N, 256)), dim3(256), 0, (uint32_t*)a, N,
             d->t_evals, d->t_count - 1);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

int mzk_lagrange_to_coeff(mzk_domain* d, uint64_t* a) {
  MZK_TRY(domain_check(d));
  return run_host(d, a, a, 1ull << d->k, 1ull << d->k, l2c_impl);
}
int mzk_extended_to_coeff(mzk_domain* d, uint64_t* a) {
  MZK_TRY(domain_check(d));
  return run_host(d, a, a, 1ull << d->ext_k, 1ull << d->ext_k, e2c_impl);
}
int mzk_coeff_to_extended(mzk_domain* d, const uint64_t* in, uint64_t* out) {
  MZK_TRY(domain_check(d));
  return run_host(d, in, out, 1ull << d->k, 1ull << d->ext_k, c2e_impl);
}

}  // extern "C"

namespace mzk {
int ntt_check_job(const mzk_ntt_job& j, size_t i) {
  MZK_REQUIRE(j.d_in != nullptr && j.d_out != nullptr, "ntt job %zu: null data", i);
  MZK_REQUIRE(j.log_n >= 1 && j.log_n <= (uint32_t)NTT_MAX_LOG_N, "ntt job %zu: log_n %u outside [1, %d]", i,
              j.log_n, NTT_MAX_LOG_N);
  MZK_REQUIRE((j.flags & ~(uint32_t)(MZK_NTT_INVERSE | MZK_NTT_COSET)) == 0, "ntt job %zu: unknown flags", i);
  MZK_REQUIRE(j.n_real <= (1ull << j.log_n), "ntt job %zu: n_real above 2^log_n", i);
  MZK_REQUIRE(!(j.flags & MZK_NTT_INVERSE) || j.n_real == (1ull << j.log_n),
              "ntt job %zu: inverse transforms take 2^log_n inputs", i);
  // in place (d_in == d_out) is fine: a multi-pass transform reads d_in only in its first
  // pass and writes d_out only in its last; a single pass stages all of a sub-DFT in LDS
  return MZK_OK;
}
}  // namespace mzk

extern "C" {

// byte ranges of a job's input and output
static void job_spans(const mzk_ntt_job& j, uintptr_t in[2], uintptr_t out[2]) {
  const uint64_t n_in = (j.flags & MZK_NTT_INVERSE) ? (1ull << j.log_n) : j.n_real;
  in[0] = (uintptr_t)j.d_in;
  in[1] = in[0] + 32 * n_in;
  out[0] = (uintptr_t)j.d_out;
  out[1] = out[0] + 32 * (1ull << j.log_n);
}

static bool spans_meet(const uintptr_t a[2], const uintptr_t b[2]) { return a[0] < b[1] && b[0] < a[1]; }

// A batch runs every job's first pass in one launch (single-pass jobs write their outputs
// there) and its later passes in later launches, so no job's output may overlap another
// job's input or output; a job's own d_in == d_out is allowed.  Sorted sweep: O(n log n).
static int check_batch_overlaps(const mzk_ntt_job* jobs, size_t njobs) {
  struct Span {
    uintptr_t lo, hi;
    size_t job;
    bool out;
  };
  std::vector<Span> v;
  v.reserve(2 * njobs);
  for (size_t i = 0; i < njobs; i++) {
    uintptr_t in[2], out[2];
    job_spans(jobs[i], in, out);
    MZK_REQUIRE(in[0] == out[0] || !spans_meet(in, out), "ntt batch: job %zu's input and output overlap", i);
    if (in[1] > in[0] && in[0] != out[0]) v.push_back({in[0], in[1], i, false});
    v.push_back({out[0], out[1], i, true});
  }
  std::sort(v.begin(), v.end(), [](const Span& x, const Span& y) { return x.lo < y.lo; });
  // every pair of overlapping spans meets some span that starts earlier and is still open;
  // track the open span reaching furthest (and, if it is an input, the furthest output)
  size_t far = SIZE_MAX, far_out = SIZE_MAX;
  for (size_t k = 0; k < v.size(); k++) {
    const Span& s = v[k];
    auto clash = [&](size_t o) {
      return o != SIZE_MAX && v[o].hi > s.lo && v[o].job != s.job && (v[o].out || s.out);
    };
    if (clash(far) || clash(far_out)) {
      const Span& o = clash(far) ? v[far] : v[far_out];
      set_error("ntt batch: job %zu's %s overlaps job %zu's %s", s.job, s.out ? "output" : "input", o.job,
                o.out ? "output" : "input");
      return MZK_ERR_INVALID;
    }
    if (far == SIZE_MAX || s.hi > v[far].hi) far = k;
    if (s.out && (far_out == SIZE_MAX || s.hi > v[far_out].hi)) far_out = k;
  }
  return MZK_OK;
}

int mzk_ntt_dev(mzk_ctx* c, const mzk_ntt_job* job) {
  MZK_REQUIRE(c != nullptr && job != nullptr, "mzk_ntt_dev: null argument");
  MZK_TRY(ntt_check_job(*job, 0));
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  ntt_plans_trim(c);
  const int inverse = (job->flags & MZK_NTT_INVERSE) != 0, coset = (job->flags & MZK_NTT_COSET) != 0;
  const NttPlan* plan = ctx_plan(c, job->log_n, inverse);
  MZK_REQUIRE(plan != nullptr, "mzk_ntt_dev: plan for 2^%u failed", job->log_n);
  const uint64_t N = 1ull << job->log_n;
  if (inverse)
    return ntt_run(c, *plan, (const uint32_t*)job->d_in, (uint32_t*)job->d_out, N, 0, coset ? 2 : 1,
                   host::fr_inv(host::fr_from_u64(N)));
  return ntt_run(c, *plan, (const uint32_t*)job->d_in, (uint32_t*)job->d_out, job->n_real, coset, 0, host::fr_one());
}

int mzk_ntt_batch_dev(mzk_ctx* c, const mzk_ntt_job* jobs, size_t njobs) {
  MZK_REQUIRE(c != nullptr && (jobs != nullptr || njobs == 0), "mzk_ntt_batch_dev: null argument");
  for (size_t i = 0; i < njobs; i++) MZK_TRY(ntt_check_job(jobs[i], i));
  MZK_TRY(check_batch_overlaps(jobs, njobs));
  if (njobs == 0) return MZK_OK;
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  return ntt_run_batch(c, jobs, njobs);
}

// The root a twiddle table [1, w, w^2, ..., w^(n/2 - 1)] (compute_twiddles, fft.rs:31-40)
// was built from: w = twiddles[1] (n >= 4), -1 for n = 2.  The table is checked where
// it is cheap: twiddles[0] = 1, w^(n/2) = -1 (the radix-2 butterflies of fft.rs:56-101
// compute the DFT by w exactly when this holds) and the last entry = w^(n/2 - 1).
static int root_from_twiddles(const uint64_t* tw, uint32_t log_n, host::Fr* w, const char* who) {
  MZK_REQUIRE(tw != nullptr, "%s: null twiddles", who);
  const host::Fr one = host::fr_one();
  const host::Fr minus_one = host::fr_sub(host::fr_zero(), one);
  host::Fr t0;
  memcpy(t0.l, tw, 32);
  MZK_REQUIRE(host::fr_eq(t0, one), "%s: twiddles[0] is not 1", who);
  if (log_n == 1) {
    *w = minus_one;
    return MZK_OK;
  }
  memcpy(w->l, tw + 4, 32);
  host::Fr h = *w;
  for (uint32_t i = 1; i < log_n; i++) h = host::fr_mul(h, h);   // w^(n/2)
  MZK_REQUIRE(host::fr_eq(h, minus_one), "%s: twiddles[1] is not a primitive 2^%u-th root of unity", who, log_n);
  const uint64_t last = (1ull << (log_n - 1)) - 1;
  host::Fr tl;
  memcpy(tl.l, tw + 4 * last, 32);
  MZK_REQUIRE(host::fr_eq(tl, host::fr_pow(*w, last)), "%s: twiddles[%llu] is not twiddles[1]^%llu", who,
              (unsigned long long)last, (unsigned long long)last);
  return MZK_OK;
}

// DFT by omega of the first n_real host elements of a (zero beyond), in place, all of
// 2^log_n elements written back.
static int host_fft(mzk_ctx* c, uint64_t* a, const host::Fr& omega, uint32_t log_n, uint64_t n_real) {
  const size_t N = 1ull << log_n;
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  ntt_plans_trim(c);
  const NttPlan* plan = ctx_plan_root(c, log_n, omega);
  MZK_REQUIRE(plan != nullptr, "fft: plan for 2^%u failed", log_n);
  MZK_TRY(c->scratch[SCR_IN0].reserve(N * 32));
  uint32_t* d = (uint32_t*)c->scratch[SCR_IN0].ptr;
  if (n_real) MZK_HIP(hipMemcpyAsync(d, a, n_real * 32, hipMemcpyHostToDevice, c->stream));
  MZK_TRY(ntt_run(c, *plan, d, d, n_real, 0, 0, host::fr_one()));
  MZK_HIP(hipMemcpyAsync(a, d, N * 32, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  return MZK_OK;
}

// best_fft (fft.rs:49): DFT with a caller-supplied omega, in place on host memory.
int mzk_best_fft(mzk_ctx* c, uint64_t* a, const uint64_t omega[4], uint32_t log_n) {
  MZK_REQUIRE(c != nullptr && a != nullptr && omega != nullptr, "mzk_best_fft: null argument");
  MZK_REQUIRE(log_n <= (uint32_t)NTT_MAX_LOG_N, "mzk_best_fft: log_n %u above %d", log_n, NTT_MAX_LOG_N);
  if (log_n == 0) return MZK_OK;   // a DFT of length 1 is the identity
  host::Fr om;
  memcpy(om.l, omega, 32);
  return host_fft(c, a, om, log_n, 1ull << log_n);
}

// compute_twiddles (fft.rs:31-40): out[i] = omega^i, i < 2^(log_n - 1), host memory.
int mzk_compute_twiddles(mzk_ctx* c, const uint64_t omega[4], uint32_t log_n, uint64_t* out) {
  MZK_REQUIRE(c != nullptr && omega != nullptr && out != nullptr, "mzk_compute_twiddles: null argument");
  MZK_REQUIRE(log_n >= 1 && log_n <= (uint32_t)NTT_MAX_LOG_N, "mzk_compute_twiddles: log_n %u outside [1, %d]",
              log_n, NTT_MAX_LOG_N);
  const size_t half = 1ull << (log_n - 1);
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  MZK_TRY(c->scratch[SCR_IN0].reserve(half * 32));
  host::Fr om;
  memcpy(om.l, omega, 32);
  hipLaunchKernelGGL(k_pow_table, dim3(grid_for(half, 256)), dim3(256), 0, c->stream,
                     (uint32_t*)c->scratch[SCR_IN0].ptr, (uint64_t)half, pow_args(om));
  MZK_HIP(hipGetLastError());
  MZK_HIP(hipMemcpyAsync(out, c->scratch[SCR_IN0].ptr, half * 32, hipMemcpyDeviceToHost, c->stream));
  MZK_HIP(hipStreamSynchronize(c->stream));
  return MZK_OK;
}

// best_fft_with_twiddles (fft.rs:56-101): the DFT by the root the table was built from.
int mzk_best_fft_with_twiddles(mzk_ctx* c, uint64_t* a, const uint64_t* twiddles, uint32_t log_n) {
  MZK_REQUIRE(c != nullptr && a != nullptr, "mzk_best_fft_with_twiddles: null argument");
  MZK_REQUIRE(log_n <= (uint32_t)NTT_MAX_LOG_N, "mzk_best_fft_with_twiddles: log_n %u above %d", log_n,
              NTT_MAX_LOG_N);
  if (log_n == 0) return MZK_OK;
  host::Fr w;
  MZK_TRY(root_from_twiddles(twiddles, log_n, &w, "mzk_best_fft_with_twiddles"));
  return host_fft(c, a, w, log_n, 1ull << log_n);
}

// fft_coeff_to_extended (fft.rs:110-124): the same DFT of a whose entries from n_real on
// are zero (the reference's pruned DIF relies on it; only a[0, n_real) is read here).
int mzk_fft_coeff_to_extended(mzk_ctx* c, uint64_t* a, const uint64_t* twiddles, uint32_t log_n, size_t n_real) {
  MZK_REQUIRE(c != nullptr && a != nullptr, "mzk_fft_coeff_to_extended: null argument");
  MZK_REQUIRE(log_n >= 1 && log_n <= (uint32_t)NTT_MAX_LOG_N, "mzk_fft_coeff_to_extended: log_n %u outside [1, %d]",
              log_n, NTT_MAX_LOG_N);
  MZK_REQUIRE(n_real <= (1ull << log_n), "mzk_fft_coeff_to_extended: n_real %zu above 2^%u", n_real, log_n);
  host::Fr w;
  MZK_TRY(root_from_twiddles(twiddles, log_n, &w, "mzk_fft_coeff_to_extended"));
  return host_fft(c, a, w, log_n, n_real);
}

// best_fft on device memory with a caller's root: d_out = DFT by omega of d_in's first
// n_real elements, zero-padded to 2^log_n (in place allowed).
int mzk_best_fft_dev(mzk_ctx* c, const void* d_in, void* d_out, const uint64_t omega[4], uint32_t log_n,
                     uint64_t n_real) {
  MZK_REQUIRE(c != nullptr && d_in != nullptr && d_out != nullptr && omega != nullptr,
              "mzk_best_fft_dev: null argument");
  MZK_REQUIRE(log_n >= 1 && log_n <= (uint32_t)NTT_MAX_LOG_N, "mzk_best_fft_dev: log_n %u outside [1, %d]", log_n,
              NTT_MAX_LOG_N);
  MZK_REQUIRE(n_real <= (1ull << log_n), "mzk_best_fft_dev: n_real above 2^log_n");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  ntt_plans_trim(c);
  host::Fr om;
  memcpy(om.l, omega, 32);
  const NttPlan* plan = ctx_plan_root(c, log_n, om);
  MZK_REQUIRE(plan != nullptr, "mzk_best_fft_dev: plan for 2^%u failed", log_n);
  return ntt_run(c, *plan, (const uint32_t*)d_in, (uint32_t*)d_out, n_real, 0, 0, host::fr_one());
}

}  // extern "C"
