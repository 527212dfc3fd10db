// BLS12-381 scalar field r in a redundant radix-2^29 form, for the NTT passes.
//
// An element is 9 limbs l_i (value V = sum l_i 2^(29 i), up to 2^261) representing x
// with V = x * 2^261 (mod r): Montgomery form with R = 2^261.  Every 29x29-bit limb
// product is below 2^58 (2^60 for 30-bit limbs), so one 64-bit column accumulator takes
// a whole column of a Montgomery product with no carry handling: each limb product is one
// v_mad_u64_u32.  r = 1 mod 2^29, so the Montgomery digit of a column is its negation
// and the digit's product with r's low limb is the digit itself.
//
// Bounds (callers keep them; the NTT's are stated where it uses these):
//  * mul(a, b): limbs < 2^30; output normalised (limbs 0..7 < 2^29) and below
//    a b / 2^261 + r, i.e. below 2 r whenever a b < 70 r^2 (r / 2^261 < 1/70.9);
//  * add / sub_kr: normalised inputs, output normalised, value a + b or a + k r - b
//    (b < k r), all below 2^261;
//  * from_std: a canonical blst value x 2^256 shifted to x 2^261 (below 32 r);
//    reduce_small brings any V < 2^261 below 3 r;
//  * to_std: V < 33 r back to canonical x 2^256 (exact division by 32 mod r gives a value
//    below 2 r, then at most one subtraction of r); reduce_full gives the value below r.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MZK_HD29 __host__ __device__ __forceinline__
#else
#define MZK_HD29 inline
#endif

namespace mzk {
namespace r29c {

constexpr int N = 9;
constexpr uint32_t M29 = (1u << 29) - 1;

// r in radix 2^29
MZK_HD29 constexpr uint32_t R(int i) {
  constexpr uint32_t v[N] = {0x1,       0x1ffffff8, 0x1f96ffbf, 0x1b4805ff, 0x1d80553b,
                             0xc0404d0, 0x1520cce7, 0xa6533af,  0x73eda7};
  return v[i];
}
// 32-bit words of r (canonical comparisons in to_std)
MZK_HD29 constexpr uint32_t RW(int i) {
  constexpr uint32_t v[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                             0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
  return v[i];
}

}  // namespace r29c

struct fr29 {
  uint32_t l[r29c::N];
};

// (a b) 2^-261 mod r, product scanning with the reduction terms in a second accumulator
MZK_HD29 fr29 fr29_mul(const fr29& a, const fr29& b) {
  using namespace r29c;
  uint32_t m[N];
  fr29 o;
  uint64_t acc = 0;   // one chain per column: products and reduction terms alike
#pragma unroll
  for (int k = 0; k < N; k++) {
#pragma unroll
    for (int j = 0; j < k; j++) {
      acc += (uint64_t)a.l[j] * b.l[k - j];
      acc += (uint64_t)m[j] * R(k - j);
    }
    acc += (uint64_t)a.l[k] * b.l[0];
    m[k] = (0u - (uint32_t)acc) & M29;   // -acc / r mod 2^29 (r = 1 mod 2^29)
    acc += m[k];
    acc >>= 29;
  }
#pragma unroll
  for (int k = N; k < 2 * N - 1; k++) {
#pragma unroll
    for (int j = k - N + 1; j < N; j++) {
      acc += (uint64_t)a.l[j] * b.l[k - j];
      acc += (uint64_t)m[j] * R(k - j);
    }
    o.l[k - N] = (uint32_t)acc & M29;
    acc >>= 29;
  }
  o.l[N - 1] = (uint32_t)acc;
  return o;
}

MZK_HD29 fr29 fr29_add(const fr29& a, const fr29& b) {
  fr29 o;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < r29c::N - 1; i++) {
    const uint32_t t = a.l[i] + b.l[i] + c;
    o.l[i] = t & r29c::M29;
    c = t >> 29;
  }
  o.l[r29c::N - 1] = a.l[r29c::N - 1] + b.l[r29c::N - 1] + c;
  return o;
}

// a + K r - b for b < K r (signed limb differences, carries propagated).  K <= 2: every
// limb sum a_i + K r_i - b_i + c lies in (-2^29 - 2, 2^29 + 2^30 + 4), so 32-bit signed
// arithmetic holds it (four instructions a limb instead of a 64-bit chain).
template <int K>
MZK_HD29 fr29 fr29_sub_kr(const fr29& a, const fr29& b) {
  static_assert(K >= 0 && K <= 2, "sub_kr: K r must keep the limb sums within 32 bits");
  fr29 o;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < r29c::N - 1; i++) {
    const int32_t t = (int32_t)(a.l[i] + (uint32_t)K * r29c::R(i) - b.l[i]) + c;
    o.l[i] = (uint32_t)t & r29c::M29;
    c = t >> 29;   // arithmetic: the borrow is -1, 0 or a small carry
  }
  o.l[r29c::N - 1] = a.l[r29c::N - 1] + (uint32_t)K * r29c::R(r29c::N - 1) - b.l[r29c::N - 1] + (uint32_t)c;
  return o;
}

// Limb-wise sums with no carry propagation, for values that only feed a product or an
// add / sub_kr (the NTT's first stage of a register pair):
//  * add_lazy(a, b) = a + b limb by limb: normalised inputs give limbs <= 2^30 - 2,
//    which mul takes (limbs < 2^30) and sub_kr<2> takes as its first operand (its limb
//    sum a_i + 2 r_i - b_i + carry stays <= 2^31 - 1);
//  * sub3r_lazy(a, b) = a + 3 r - b, 3 r written with its lower limbs raised by 2^29
//    (and the next limb lowered by 1) so every limb difference is >= 0 for a normalised
//    b < 2 r; limbs < 2^31, which mul takes when its other operand is normalised and
//    below 2^29 per limb (a column then sums to < 9 2^60 + 9 2^58 < 2^64).
namespace r29c {
MZK_HD29 constexpr uint32_t R3RAISED(int i) {
  constexpr uint32_t v[N] = {0x20000003, 0x3fffffe7, 0x3ec4ff3e, 0x31d811fe, 0x3880ffb2,
                             0x240c0e71, 0x3f6266b5, 0x3f2f9b0d, 0x15bc8f4};
  return v[i];
}
}  // namespace r29c

MZK_HD29 fr29 fr29_add_lazy(const fr29& a, const fr29& b) {
  fr29 o;
#pragma unroll
  for (int i = 0; i < r29c::N; i++) o.l[i] = a.l[i] + b.l[i];
  return o;
}

MZK_HD29 fr29 fr29_sub3r_lazy(const fr29& a, const fr29& b) {
  fr29 o;
#pragma unroll
  for (int i = 0; i < r29c::N; i++) o.l[i] = a.l[i] + r29c::R3RAISED(i) - b.l[i];
  return o;
}

// blst fr (8 x 32-bit, canonical x 2^256) -> x 2^261: the value shifted left by 5 bits,
// regrouped into 29-bit limbs (below 32 r)
MZK_HD29 fr29 fr29_from_std(const uint32_t s[8]) {
  fr29 o;
#pragma unroll
  for (int i = 0; i < r29c::N; i++) {
    const int lo = 29 * i - 5;   // bit of s that lands at bit 29 i
    uint32_t d;
    if (lo < 0) {
      d = (s[0] << (-lo)) & r29c::M29;
    } else {
      const int w = lo >> 5, sh = lo & 31;
      uint64_t two = (uint64_t)s[w];
      if (w + 1 < 8) two |= (uint64_t)s[w + 1] << 32;
      d = (uint32_t)(two >> sh);
      if (i < r29c::N - 1) d &= r29c::M29;
    }
    o.l[i] = d;
  }
  return o;
}

// V (normalised, below 2^261) -> V - q r below 3 r, q = floor(V_8 / (r_8 + 1)) <= V / r
MZK_HD29 fr29 fr29_reduce_small(const fr29& a) {
  const uint32_t q = a.l[r29c::N - 1] / (r29c::R(r29c::N - 1) + 1);
  fr29 o;
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < r29c::N - 1; i++) {
    const int64_t t = (int64_t)a.l[i] - (int64_t)q * r29c::R(i) + c;
    o.l[i] = (uint32_t)t & r29c::M29;
    c = t >> 29;
  }
  o.l[r29c::N - 1] = (uint32_t)((int64_t)a.l[r29c::N - 1] - (int64_t)q * r29c::R(r29c::N - 1) + c);
  return o;
}

// V (normalised, below 2^261) -> the representative below r (reduce_small, then r once)
MZK_HD29 fr29 fr29_reduce_full(const fr29& a) {
  const fr29 v = fr29_reduce_small(a);   // below 2 r
  bool ge = true;
#pragma unroll
  for (int i = r29c::N - 1; i >= 0; i--) {
    if (v.l[i] != r29c::R(i)) {
      ge = v.l[i] > r29c::R(i);
      break;
    }
  }
  return ge ? fr29_sub_kr<0>(v, fr29{{r29c::R(0), r29c::R(1), r29c::R(2), r29c::R(3), r29c::R(4), r29c::R(5),
                                     r29c::R(6), r29c::R(7), r29c::R(8)}})
            : v;
}

// V (normalised, below 33 r, representing x 2^261) -> canonical x 2^256 as 8 words:
// (V + k r) / 32 with k = -V mod 32 (r = 1 mod 32), then one conditional subtraction of r
MZK_HD29 void fr29_to_std(const fr29& a, uint32_t out[8]) {
  using namespace r29c;
  const uint32_t k = (0u - a.l[0]) & 31u;
  // W = V + k r, limbs normalised
  uint32_t w[N];
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const uint64_t t = (uint64_t)a.l[i] + (uint64_t)k * R(i) + c;
    w[i] = (i < N - 1) ? (uint32_t)t & M29 : (uint32_t)t;
    c = t >> 29;
  }
  // W / 32 in 32-bit words: bit 32 j + b of the quotient is bit 32 j + b + 5 of W
  uint32_t q[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int lo = 32 * j + 5;
    const int li = lo / 29, sh = lo % 29;
    uint64_t v = (uint64_t)w[li] >> sh;
    int have = 29 - sh;
    if (li + 1 < N) {
      v |= (uint64_t)w[li + 1] << have;
      have += 29;
    }
    if (li + 2 < N && have < 32) v |= (uint64_t)w[li + 2] << have;
    q[j] = (uint32_t)v;
  }
  // below 2 r: subtract r once if q >= r
  bool ge = true;
#pragma unroll
  for (int j = 7; j >= 0; j--) {
    if (q[j] != RW(j)) {
      ge = q[j] > RW(j);
      break;
    }
  }
  uint32_t borrow = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const uint64_t t = (uint64_t)q[j] - (ge ? RW(j) : 0u) - borrow;
    out[j] = (uint32_t)t;
    borrow = (uint32_t)(t >> 63);
  }
}

}  // namespace mzk
