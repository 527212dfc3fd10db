// Lagrange and delta bases of a resident SRS, built on the device.
//
//  mzk_srs_to_lagrange  g_to_lagrange (proofs/src/utils/arithmetic.rs:34-54): the
//                       inverse DFT of the monomial bases over G1 (best_fft with
//                       omega^-1 on projective points), times 1/n, normalised to
//                       affine -- g_lagrange_i = [L_i(tau)] G without knowing tau.
//  mzk_srs_suffix_sum   suffix_sum (proofs/src/poly/kzg/params.rs:47-57):
//                       out_i = sum_(j >= i) in_j (the g_lagrange_delta and
//                       g_lagrange_double_delta bases of unsafe_setup, :196-197).
//
// The DFT is radix-2 decimation in time over bit-reversed input; each butterfly
// multiplies one point by its twiddle (a 255-bit double-and-add) and writes the sum
// and the difference.  Points live in HBM as radix-2^28 XYZZ (g1_28.hpp) between
// stages.  The suffix sum is a three-phase scan: chunk totals, one workgroup's suffix
// scan over the totals, and each chunk again from its carry-in.
#include <vector>

#include "ctx.hpp"
#include "g1_28.hpp"
#include "host_ec.hpp"
#include "srs.hpp"

namespace mzk {

static constexpr int W28 = MZK_X28_WORDS;

// -P: Y -> (16p - Y) 2^392 2^-392, brought below 2p by one product (Y < 8p here)
__device__ __forceinline__ g1x28 g1x28_neg(const g1x28& p) {
  if (g1x28_is_inf(p)) return p;
  g1x28 r = p;
  r.Y = f28_mul(f28_sub<16, 1>(f28_zero(), p.Y), f28_one());
  return r;
}

// [s] P for a canonical scalar (8 x 32-bit limbs), left-to-right double-and-add
__device__ g1x28 g1x28_mul_scalar(const g1x28& p, const uint32_t s[8]) {
  g1x28 acc = g1x28_inf();
  int top = 255;
  while (top >= 0 && !((s[top >> 5] >> (top & 31)) & 1u)) top--;
  for (int b = top; b >= 0; b--) {
    if (!g1x28_is_inf(acc)) acc = g1x28_dbl(acc);
    if ((s[b >> 5] >> (b & 31)) & 1u) acc = g1x28_add(acc, p);
  }
  return acc;
}

__device__ __forceinline__ g1x28 load_affine_x28(const uint32_t* __restrict__ a) {
  uint32_t any = 0;
  for (int i = 0; i < 24; i++) any |= a[i];
  if (!any) return g1x28_inf();
  return g1x28_from_affine(f28_from_std(a), f28_from_std(a + 12), false);
}

__device__ __forceinline__ uint32_t bit_reverse(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

// pts[rev(i)] = base i
__global__ void __launch_bounds__(64) k_g1fft_load(const uint32_t* __restrict__ aff, size_t n, int log_n,
                                                   uint32_t* __restrict__ pts) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = log_n ? bit_reverse((uint32_t)i, log_n) : 0u;
  store_x28_raw(pts + W28 * (size_t)r, load_affine_x28(aff + 24 * i));
}

// One DIT stage with half-size h: butterflies (i0, i0 + h), i0 = 2h g + j, twiddle
// w^(j n / (2h)) from tw (canonical, w^k for k < n/2).
__global__ void __launch_bounds__(64) k_g1fft_stage(uint32_t* __restrict__ pts, size_t n, size_t h,
                                                    const uint32_t* __restrict__ tw) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n / 2) return;
  const size_t j = t % h, g = t / h;
  const size_t i0 = 2 * h * g + j, i1 = i0 + h;
  const g1x28 a = load_x28(pts + W28 * i0);
  g1x28 b = load_x28(pts + W28 * i1);
  if (j != 0) {
    uint32_t s[8];
    const uint32_t* w = tw + 8 * (j * (n / (2 * h)));
#pragma unroll
    for (int k = 0; k < 8; k++) s[k] = w[k];
    b = g1x28_mul_scalar(b, s);
  }
  store_x28_raw(pts + W28 * i0, g1x28_add(a, b));
  store_x28_raw(pts + W28 * i1, g1x28_add(a, g1x28_neg(b)));
}

// out_std[i] = [s] pts[i] in the blst XYZZ layout
struct Scalar8 {
  uint32_t v[8];
};
__global__ void __launch_bounds__(64) k_g1fft_scale(const uint32_t* __restrict__ pts, size_t n, Scalar8 s,
                                                    uint32_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t o[48];
  g1x28_to_std(g1x28_mul_scalar(load_x28(pts + W28 * i), s.v), o);
  uint4* q = reinterpret_cast<uint4*>(out + 48 * i);
#pragma unroll
  for (int k = 0; k < 12; k++) q[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
}

// ------------------------------------------------------------------ suffix sums
static constexpr int SUF_C = 64;      // points per thread
static constexpr int SUF_T = 256;     // threads of the carry workgroup

__global__ void __launch_bounds__(64) k_suffix_chunks(const uint32_t* __restrict__ aff, size_t n,
                                                      uint32_t* __restrict__ tot) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t lo = c * SUF_C;
  if (lo >= n) return;
  const size_t hi = lo + SUF_C < n ? lo + SUF_C : n;
  g1x28 acc = g1x28_inf();
  for (size_t i = lo; i < hi; i++) {
    const uint32_t* a = aff + 24 * i;
    uint32_t any = 0;
    for (int k = 0; k < 24; k++) any |= a[k];
    if (any) acc = g1x28_madd(acc, f28_from_std(a), f28_from_std(a + 12), false);
  }
  store_x28_raw(tot + W28 * c, acc);
}

// carry[c] = sum of the totals of chunks above c.  Thread t owns chunks [t S, t S + S).
__global__ void __launch_bounds__(SUF_T) k_suffix_carries(const uint32_t* __restrict__ tot, size_t C,
                                                          uint32_t* __restrict__ carry) {
  extern __shared__ uint32_t lds[];
  const int t = threadIdx.x;
  const size_t S = (C + SUF_T - 1) / SUF_T;
  const size_t c0 = t * S, c1 = c0 + S < C ? c0 + S : C;
  g1x28 mine = g1x28_inf();
  for (size_t c = c0; c < c1; c++) mine = g1x28_add(mine, load_x28(tot + W28 * c));
  // inclusive suffix sums over threads: after the loop lds[t] = sum of threads >= t
  store_x28_raw(lds + W28 * t, mine);
  __syncthreads();
  g1x28 x = mine;
  for (int d = 1; d < SUF_T; d <<= 1) {
    g1x28 y = g1x28_inf();
    if (t + d < SUF_T) y = load_x28(lds + W28 * (t + d));
    __syncthreads();
    x = g1x28_add(x, y);
    store_x28_raw(lds + W28 * t, x);
    __syncthreads();
  }
  g1x28 in = t + 1 < SUF_T ? load_x28(lds + W28 * (t + 1)) : g1x28_inf();
  for (size_t c = c1; c-- > c0;) {
    store_x28_raw(carry + W28 * c, in);
    in = g1x28_add(in, load_x28(tot + W28 * c));
  }
}

__global__ void __launch_bounds__(64) k_suffix_final(const uint32_t* __restrict__ aff, size_t n,
                                                     const uint32_t* __restrict__ carry, uint32_t* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t lo = c * SUF_C;
  if (lo >= n) return;
  const size_t hi = lo + SUF_C < n ? lo + SUF_C : n;
  g1x28 acc = load_x28(carry + W28 * c);
  for (size_t i = hi; i-- > lo;) {
    const uint32_t* a = aff + 24 * i;
    uint32_t any = 0;
    for (int k = 0; k < 24; k++) any |= a[k];
    if (any) acc = g1x28_madd(acc, f28_from_std(a), f28_from_std(a + 12), false);
    uint32_t o[48];
    g1x28_to_std(acc, o);
    uint4* q = reinterpret_cast<uint4*>(out + 48 * i);
#pragma unroll
    for (int k = 0; k < 12; k++) q[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
  }
}

static void fr_canonical(const host::Fr& a, uint32_t out[8]) {
  host::Fr one_raw = host::fr_zero();
  one_raw.l[0] = 1;
  const host::Fr v = host::fr_mul(a, one_raw);
  memcpy(out, v.l, 32);
}

}  // namespace mzk

using namespace mzk;

extern "C" {

int mzk_srs_to_lagrange(mzk_ctx* c, const mzk_srs* g, uint32_t k, mzk_srs** out) {
  MZK_REQUIRE(c != nullptr && g != nullptr && out != nullptr, "mzk_srs_to_lagrange: null argument");
  MZK_REQUIRE(k <= 24 && ((size_t)1 << k) <= g->n, "mzk_srs_to_lagrange: 2^%u bases needed, SRS has %zu", k, g->n);
  const size_t n = (size_t)1 << k;
  mzk_srs* s = nullptr;
  int rc = MZK_OK;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    do {
      if (hipSetDevice(c->device) != hipSuccess) {
        rc = MZK_ERR_HIP;
        set_error("mzk_srs_to_lagrange: hipSetDevice failed");
        break;
      }
      if ((rc = srs_alloc_bases(c, n, &s)) != MZK_OK) break;
      // twiddles omega^-k, k < n/2 (canonical), omega = ROOT^(2^(32 - k)) (domain.rs:72-91)
      const size_t half = n / 2 ? n / 2 : 1;
      std::vector<uint32_t> tw(8 * half);
      const host::Fr winv = host::fr_inv(host::fr_root_of_unity(k));
      host::Fr cur = host::fr_one();
      for (size_t i = 0; i < half; i++) {
        fr_canonical(cur, &tw[8 * i]);
        cur = host::fr_mul(cur, winv);
      }
      if ((rc = c->scratch[SCR_IN0].reserve(tw.size() * 4 + 256)) != MZK_OK) break;
      if ((rc = c->scratch[SCR_MSM_C].reserve(n * W28 * 4 + 256)) != MZK_OK) break;
      if ((rc = c->scratch[SCR_MSM_B].reserve(n * 48 * 4 + 256)) != MZK_OK) break;
      uint32_t* d_tw = (uint32_t*)c->scratch[SCR_IN0].ptr;
      uint32_t* pts = (uint32_t*)c->scratch[SCR_MSM_C].ptr;
      uint32_t* xyzz = (uint32_t*)c->scratch[SCR_MSM_B].ptr;
      if (hipMemcpyAsync(d_tw, tw.data(), tw.size() * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        rc = MZK_ERR_HIP;
        set_error("mzk_srs_to_lagrange: twiddle upload failed");
        break;
      }
      MZK_LAUNCH(c, "k_g1fft_load", k_g1fft_load, dim3(grid_for(n, 64)), dim3(64), 0, (const uint32_t*)g->bases, n,
                 (int)k, pts);
      for (size_t h = 1; h < n; h <<= 1)
        MZK_LAUNCH(c, "k_g1fft_stage", k_g1fft_stage, dim3(grid_for(n / 2, 64)), dim3(64), 0, pts, n, h, d_tw);
      Scalar8 ninv;
      fr_canonical(host::fr_inv(host::fr_from_u64(n)), ninv.v);
      MZK_LAUNCH(c, "k_g1fft_scale", k_g1fft_scale, dim3(grid_for(n, 64)), dim3(64), 0, pts, n, ninv, xyzz);
      if (hipGetLastError() != hipSuccess) {
        rc = MZK_ERR_HIP;
        set_error("mzk_srs_to_lagrange: launch failed");
        break;
      }
      if ((rc = batch_normalize(c, xyzz, n, (uint32_t*)s->bases)) != MZK_OK) break;
      if (hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = MZK_ERR_HIP;
        set_error("mzk_srs_to_lagrange: device error");
        break;
      }
      rc = srs_finish_bases(c, s);
    } while (0);
  }
  if (rc != MZK_OK) {
    if (s) mzk_srs_destroy(s);
    return rc;
  }
  *out = s;
  return MZK_OK;
}

int mzk_srs_suffix_sum(mzk_ctx* c, const mzk_srs* in, mzk_srs** out) {
  MZK_REQUIRE(c != nullptr && in != nullptr && out != nullptr, "mzk_srs_suffix_sum: null argument");
  const size_t n = in->n;
  const size_t C = (n + SUF_C - 1) / SUF_C;
  mzk_srs* s = nullptr;
  int rc = MZK_OK;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    do {
      if (hipSetDevice(c->device) != hipSuccess) {
        rc = MZK_ERR_HIP;
        set_error("mzk_srs_suffix_sum: hipSetDevice failed");
        break;
      }
      if ((rc = srs_alloc_bases(c, n, &s)) != MZK_OK) break;
      if ((rc = c->scratch[SCR_MSM_C].reserve(2 * C * W28 * 4 + 256)) != MZK_OK) break;
      if ((rc = c->scratch[SCR_MSM_B].reserve(n * 48 * 4 + 256)) != MZK_OK) break;
      uint32_t* tot = (uint32_t*)c->scratch[SCR_MSM_C].ptr;
      uint32_t* carry = tot + C * W28;
      uint32_t* xyzz = (uint32_t*)c->scratch[SCR_MSM_B].ptr;
      MZK_LAUNCH(c, "k_suffix_chunks", k_suffix_chunks, dim3(grid_for(C, 64)), dim3(64), 0,
                 (const uint32_t*)in->bases, n, tot);
      MZK_LAUNCH(c, "k_suffix_carries", k_suffix_carries, dim3(1), dim3(SUF_T), SUF_T * W28 * 4, tot, C, carry);
      MZK_LAUNCH(c, "k_suffix_final", k_suffix_final, dim3(grid_for(C, 64)), dim3(64), 0, (const uint32_t*)in->bases,
                 n, carry, xyzz);
      if (hipGetLastError() != hipSuccess) {
        rc = MZK_ERR_HIP;
        set_error("mzk_srs_suffix_sum: launch failed");
        break;
      }
      if ((rc = batch_normalize(c, xyzz, n, (uint32_t*)s->bases)) != MZK_OK) break;
      if (hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = MZK_ERR_HIP;
        set_error("mzk_srs_suffix_sum: device error");
        break;
      }
      rc = srs_finish_bases(c, s);
    } while (0);
  }
  if (rc != MZK_OK) {
    if (s) mzk_srs_destroy(s);
    return rc;
  }
  *out = s;
  return MZK_OK;
}

}  // extern "C"
