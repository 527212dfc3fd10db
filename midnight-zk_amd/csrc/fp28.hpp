// BLS12-381 base field in a redundant radix-2^28 form, for the MSM's bucket
// accumulation.
//
// An element is 14 limbs l_i (value V = sum l_i 2^(28 i)) representing x with
// V = x * 2^392 (mod p): Montgomery form with R = 2^392.  Limbs are not kept
// below 2^28 between operations; every 28x28-bit limb product is below 2^60, so a
// 64-bit column accumulator absorbs a whole column of a Montgomery product (<= 28
// products plus the carry) without any carry handling: each limb product is one
// v_mad_u64_u32 and nothing else.  The 32-bit form of field.hpp needs a second
// instruction per product for its carry word.
//
// Bounds (checked by the callers, stated at each use):
//  * mul(a, b) needs a*b < 2520 p^2 (so the output is < 2p, R/p = 2520.2) and limbs
//    small enough that a column sum stays below 2^64: both operands' limbs < 2^30,
//    or one operand's limbs < 2^28 and the other's < 2^31.
//  * mul outputs are normalised (limbs 0..12 < 2^28) and < 2p.
//  * mul_sum(a, b, c, d) = (a b + c d) 2^-392 needs a b + c d < 2520 p^2 and the
//    summed columns below 2^64 (both pairs' limbs < 2^29.6, or one pair's < 2^28 and
//    2^29); it saves one reduction where a formula subtracts two products.
//  * sub(a, b) adds a multiple k p of p written with lower limbs raised by
//    m 2^28 (and the next limb lowered by m), so every limb difference is >= 0
//    when b's lower limbs are <= m (2^28 - 1) and b's top limb is <= that of k p
//    minus m.
//
// Points cross HBM in the blst layout (field.hpp's 12 x 32-bit limbs, R = 2^384):
// from_std shifts a canonical value left by 8 bits (x 2^384 -> x 2^392, below 2^389,
// no reduction needed) and regroups the bits; to_std multiplies by 2^384 mod p,
// reduces to [0, p) and regroups.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MZK_HD __host__ __device__ __forceinline__
#else
#define MZK_HD inline
#endif

namespace mzk {
namespace f28c {

constexpr int N = 14;
constexpr uint32_t M28 = (1u << 28) - 1;
constexpr uint32_t PINV = 0xfffcfffdu;   // -p^-1 mod 2^32 (its low 28 bits: -p^-1 mod 2^28)
constexpr uint32_t P0INV = 0x30003u;     // p^-1 mod 2^28

// p in radix 2^28
MZK_HD constexpr uint32_t P(int i) {
  constexpr uint32_t v[N] = {0xfffaaab, 0xfefffff, 0x3ffffb9, 0xfffeb15, 0x6241eab, 0xa0f6b0f, 0xf6730d2,
                             0xf38512b, 0x4774b84, 0x4bacd76, 0xba7b643, 0xe69a4b1, 0x1ea397f, 0x1a011};
  return v[i];
}
// 2^392 mod p (one in this Montgomery form) and 2^384 mod p (the factor taking this
// form to the 2^384 form), radix 2^28
MZK_HD constexpr uint32_t ONE(int i) {
  constexpr uint32_t v[N] = {0x347fcb8, 0xd800000, 0x2b119, 0xcde6d2, 0xc7212e0, 0x83a2090, 0x37669f,
                             0xda0f73e, 0x9b09b42, 0x1297bb0, 0x515d98f, 0x12ca7c, 0x659fcfa, 0x577a};
  return v[i];
}
MZK_HD constexpr uint32_t C384(int i) {
  constexpr uint32_t v[N] = {0x2fffd, 0x900000, 0xc000276, 0xbc40, 0x8baebf4, 0x5753c75, 0x55f4898,
                             0x7052574, 0x7ce5853, 0x56ec6d7, 0x71a97a2, 0xe4935c0, 0xec3fa80, 0x15f65};
  return v[i];
}
// digit i of k p (k < 32), radix 2^28 (exact digits: limbs 0..12 < 2^28)
MZK_HD constexpr uint32_t KP_DIGIT(int k, int i) {
  // k p = sum_j k P(j) 2^(28 j): propagate the carries of the small multiples
  uint64_t c = 0;
  uint32_t d = 0;
  for (int j = 0; j <= i; j++) {
    c += (uint64_t)k * P(j);
    d = (j == N - 1) ? (uint32_t)c : (uint32_t)(c & M28);
    c >>= 28;
  }
  return d;
}
// limb i of k p with lower limbs raised by m 2^28 and the next limb lowered by m
MZK_HD constexpr uint32_t KPB(int k, int m, int i) {
  return KP_DIGIT(k, i) + (i < N - 1 ? (uint32_t)m << 28 : 0u) - (i > 0 ? (uint32_t)m : 0u);
}

}  // namespace f28c

struct f28 {
  uint32_t l[f28c::N];
};

MZK_HD f28 f28_zero() {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = 0;
  return r;
}

MZK_HD f28 f28_one() {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = f28c::ONE(i);
  return r;
}

// literal zero (the encoding of the identity's ZZ); not a test for V = 0 mod p
MZK_HD bool f28_is_literal_zero(const f28& a) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) acc |= a.l[i];
  return acc == 0;
}

// Montgomery product a b 2^-392 mod p by product scanning (Koc, Acar, Kaliski 1996):
// column k accumulates a_j b_(k-j) and m_j p_(k-j) in one 64-bit word, a single chain of
// v_mad_u64_u32 with no merge (the column's final sum, which bounds every partial one,
// is what the header's bounds keep below 2^64).
MZK_HD f28 f28_mul(const f28& a, const f28& b) {
  using namespace f28c;
  uint32_t m[N];
  f28 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
#pragma unroll
    for (int j = 0; j < k; j++) {
      acc += (uint64_t)a.l[j] * b.l[k - j];
      acc += (uint64_t)m[j] * P(k - j);
    }
    acc += (uint64_t)a.l[k] * b.l[0];
    m[k] = ((uint32_t)acc * PINV) & M28;
    acc += (uint64_t)m[k] * P(0);
    acc >>= 28;
  }
#pragma unroll
  for (int k = N; k < 2 * N - 1; k++) {
#pragma unroll
    for (int j = k - N + 1; j < N; j++) {
      acc += (uint64_t)a.l[j] * b.l[k - j];
      acc += (uint64_t)m[j] * P(k - j);
    }
    r.l[k - N] = (uint32_t)acc & M28;
    acc >>= 28;
  }
  r.l[N - 1] = (uint32_t)acc;
  return r;
}

// (a b + c d) 2^-392 mod p with one reduction: the two products share the column
// accumulator.  Needs a b + c d < 2520 p^2 and column sums below 2^64.
MZK_HD f28 f28_mul_sum(const f28& a, const f28& b, const f28& c, const f28& d) {
  using namespace f28c;
  uint32_t m[N];
  f28 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
#pragma unroll
    for (int j = 0; j < k; j++) {
      acc += (uint64_t)a.l[j] * b.l[k - j];
      acc += (uint64_t)c.l[j] * d.l[k - j];
      acc += (uint64_t)m[j] * P(k - j);
    }
    acc += (uint64_t)a.l[k] * b.l[0];
    acc += (uint64_t)c.l[k] * d.l[0];
    m[k] = ((uint32_t)acc * PINV) & M28;
    acc += (uint64_t)m[k] * P(0);
    acc >>= 28;
  }
#pragma unroll
  for (int k = N; k < 2 * N - 1; k++) {
#pragma unroll
    for (int j = k - N + 1; j < N; j++) {
      acc += (uint64_t)a.l[j] * b.l[k - j];
      acc += (uint64_t)c.l[j] * d.l[k - j];
      acc += (uint64_t)m[j] * P(k - j);
    }
    r.l[k - N] = (uint32_t)acc & M28;
    acc >>= 28;
  }
  r.l[N - 1] = (uint32_t)acc;
  return r;
}

// a^2: each cross product a_i a_j (i < j) once, as a_i (2 a_j), plus the squares, and the
// reduction terms, in one accumulator chain as in f28_mul (the column sums are those of
// f28_mul(a, a))
MZK_HD f28 f28_sqr(const f28& a) {
  using namespace f28c;
  uint32_t m[N], d[N];
  f28 r;
#pragma unroll
  for (int i = 0; i < N; i++) d[i] = a.l[i] << 1;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
#pragma unroll
    for (int j = 0; 2 * j < k; j++) acc += (uint64_t)a.l[j] * d[k - j];
#pragma unroll
    for (int j = 0; j < k; j++) acc += (uint64_t)m[j] * P(k - j);
    if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
    m[k] = ((uint32_t)acc * PINV) & M28;
    acc += (uint64_t)m[k] * P(0);
    acc >>= 28;
  }
#pragma unroll
  for (int k = N; k < 2 * N - 1; k++) {
#pragma unroll
    for (int j = k - N + 1; 2 * j < k; j++) acc += (uint64_t)a.l[j] * d[k - j];
#pragma unroll
    for (int j = k - N + 1; j < N; j++) acc += (uint64_t)m[j] * P(k - j);
    if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
    r.l[k - N] = (uint32_t)acc & M28;
    acc >>= 28;
  }
  r.l[N - 1] = (uint32_t)acc;
  return r;
}

MZK_HD f28 f28_add(const f28& a, const f28& b) {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = a.l[i] + b.l[i];
  return r;
}

// a + k p - b, k p written with lower limbs raised by m 2^28 (see the header)
template <int K, int Mg>
MZK_HD f28 f28_sub(const f28& a, const f28& b) {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = a.l[i] + f28c::KPB(K, Mg, i) - b.l[i];
  return r;
}

// k p - a - b (the subtraction of two values at once)
template <int K, int Mg>
MZK_HD f28 f28_neg_sum(const f28& a, const f28& b) {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = f28c::KPB(K, Mg, i) - a.l[i] - b.l[i];
  return r;
}

// carries propagated: limbs 0..12 < 2^28, the rest in limb 13 (value unchanged)
MZK_HD f28 f28_norm(const f28& a) {
  f28 r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < f28c::N - 1; i++) {
    const uint32_t t = a.l[i] + c;
    r.l[i] = t & f28c::M28;
    c = t >> 28;
  }
  r.l[f28c::N - 1] = a.l[f28c::N - 1] + c;
  return r;
}

// V = k p for some 0 < k < KMAX?  V below KMAX p with limbs < 2^31.  The low limb
// decides for all but a 2^-28 fraction of inputs; the rest are normalised and
// compared digit by digit.
template <int KMAX>
MZK_HD bool f28_is_zero_mod_p(const f28& a) {
  // the low digit of k p is k p_0 mod 2^28 (p_0 odd): k = lo p_0^-1 mod 2^28
  const uint32_t lo = a.l[0] & f28c::M28;
  const uint32_t kk = (lo * f28c::P0INV) & f28c::M28;
  // (the likely exit: the slow path below is placed out of the callers' loops)
  if (__builtin_expect(kk - 1u >= (uint32_t)(KMAX - 1), 1)) return false;
  const f28 n = f28_norm(a);
  // unrolled: the digits of k p are then constants (a runtime k or i would index the
  // constant tables in scratch memory)
  bool hit = false;
#pragma unroll
  for (int k = 1; k < KMAX; k++) {
    uint32_t diff = 0;
#pragma unroll
    for (int i = 0; i < f28c::N; i++) diff |= n.l[i] ^ f28c::KP_DIGIT(k, i);
    hit = hit || diff == 0;
  }
  return hit;
}

// blst Montgomery (12 x 32-bit limbs, canonical x 2^384) -> this form (x 2^392): the
// value shifted left by 8 bits, regrouped into 28-bit limbs (value < 2^389).
MZK_HD f28 f28_from_std(const uint32_t s[12]) {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) {
    const int lo = 28 * i - 8;   // bit of s that lands at bit 28 i
    uint32_t d;
    if (lo < 0) {
      d = (s[0] << (-lo)) & f28c::M28;
    } else {
      const int w = lo >> 5, sh = lo & 31;
      uint64_t two = (uint64_t)s[w];
      if (w + 1 < 12) two |= (uint64_t)s[w + 1] << 32;
      d = (uint32_t)(two >> sh);
      if (i < f28c::N - 1) d &= f28c::M28;
    }
    r.l[i] = d;
  }
  return r;
}

// this form -> canonical blst Montgomery (12 x 32-bit limbs)
MZK_HD void f28_to_std(const f28& a, uint32_t s[12]) {
  f28 c;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) c.l[i] = f28c::C384(i);
  f28 v = f28_mul(a, c);   // < 2p, normalised
  // v - p with signed carries; keep v when that is negative
  f28 d;
  int32_t carry = 0;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) {
    int32_t t = (int32_t)v.l[i] - (int32_t)f28c::P(i) + carry;
    if (i < f28c::N - 1) {
      d.l[i] = (uint32_t)t & f28c::M28;
      carry = t >> 28;   // arithmetic shift: -1 on a borrow
    } else {
      d.l[i] = (uint32_t)t;
      carry = t;
    }
  }
  const bool keep = carry < 0;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) v.l[i] = keep ? v.l[i] : d.l[i];
  // regroup 14 x 28 bits -> 12 x 32 bits (value < p < 2^381)
#pragma unroll
  for (int w = 0; w < 12; w++) {
    const int bit = 32 * w;
    const int i = bit / 28, sh = bit % 28;
    uint64_t two = (uint64_t)v.l[i] | ((i + 1 < f28c::N) ? (uint64_t)v.l[i + 1] << 28 : 0ull);
    if (sh > 24 && i + 2 < f28c::N) two |= (uint64_t)v.l[i + 2] << 56;
    s[w] = (uint32_t)(two >> sh);
  }
}

}  // namespace mzk
