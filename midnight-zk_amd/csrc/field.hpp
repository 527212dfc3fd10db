// BLS12-381 field arithmetic for gfx950 on 32-bit limbs.
//
// fr: the 255-bit scalar field r (the reference's `Fq`, curves/src/bls12_381/fq.rs:29),
//     8 x u32 little-endian, Montgomery form with R = 2^256 -- the same bytes as
//     blst_fr's 4 x u64 (fq.rs:27-29), so host arrays cross the boundary unchanged.
// fp: the 381-bit base field p (the reference's `Fp`, fp.rs:98), 12 x u32, R = 2^384,
//     the same bytes as blst_fp's 6 x u64.
//
// Products are Montgomery multiplications (Koc, Acar, Kaliski 1996) with every limb
// product a 32x32+64 -> 64-bit multiply-add, which gfx950 executes as
// v_mad_u64_u32.  Outputs are fully reduced to [0, m).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mzk {

struct FrP {
  static constexpr int N = 8;
  static constexpr uint32_t INV = 0xffffffffu;   // -r^-1 mod 2^32
  __device__ static constexpr uint32_t m(int i) {
    constexpr uint32_t P[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                               0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
    return P[i];
  }
  __device__ static constexpr uint32_t one(int i) {   // R mod r
    constexpr uint32_t O[8] = {0xfffffffeu, 0x00000001u, 0x00034802u, 0x5884b7fau,
                               0xecbc4ff5u, 0x998c4fefu, 0xacc5056fu, 0x1824b159u};
    return O[i];
  }
};

struct FpP {
  static constexpr int N = 12;
  static constexpr uint32_t INV = 0xfffcfffdu;   // -p^-1 mod 2^32
  __device__ static constexpr uint32_t m(int i) {
    constexpr uint32_t P[12] = {0xffffaaabu, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu,
                                0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu, 0x64774b84u,
                                0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
    return P[i];
  }
  __device__ static constexpr uint32_t one(int i) {   // R mod p
    constexpr uint32_t O[12] = {0x0002fffdu, 0x76090000u, 0xc40c0002u, 0xebf4000bu,
                                0x53c758bau, 0x5f489857u, 0x70525745u, 0x77ce5853u,
                                0xa256ec6du, 0x5c071a97u, 0xfa80e493u, 0x15f65ec3u};
    return O[i];
  }
};

template <class P>
struct Field {
  static constexpr int N = P::N;
  uint32_t v[N];

  __device__ __forceinline__ static Field zero() {
    Field r;
#pragma unroll
    for (int i = 0; i < N; i++) r.v[i] = 0;
    return r;
  }
  __device__ __forceinline__ static Field one() {
    Field r;
#pragma unroll
    for (int i = 0; i < N; i++) r.v[i] = P::one(i);
    return r;
  }
  __device__ __forceinline__ bool is_zero() const {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < N; i++) acc |= v[i];
    return acc == 0;
  }
  __device__ __forceinline__ bool operator==(const Field& o) const {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < N; i++) acc |= v[i] ^ o.v[i];
    return acc == 0;
  }
};

using fr_t = Field<FrP>;
using fp_t = Field<FpP>;

// r = a - m if a >= m (a < 2m assumed), branch-free select.
template <class P>
__device__ __forceinline__ void reduce_once(Field<P>& a) {
  constexpr int N = P::N;
  uint32_t d[N];
  uint64_t borrow = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    uint64_t t = (uint64_t)a.v[i] - P::m(i) - borrow;
    d[i] = (uint32_t)t;
    borrow = (t >> 32) & 1;
  }
  // borrow == 0  <=>  a >= m
  const bool ge = borrow == 0;
#pragma unroll
  for (int i = 0; i < N; i++) a.v[i] = ge ? d[i] : a.v[i];
}

template <class P>
__device__ __forceinline__ Field<P> add(const Field<P>& a, const Field<P>& b) {
  constexpr int N = P::N;
  Field<P> r;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    c = (uint64_t)a.v[i] + b.v[i] + (c >> 32);
    r.v[i] = (uint32_t)c;
  }
  // both moduli leave the top bit free, so a + b < 2^(32N) never carries out
  reduce_once(r);
  return r;
}

template <class P>
__device__ __forceinline__ Field<P> sub(const Field<P>& a, const Field<P>& b) {
  constexpr int N = P::N;
  Field<P> r;
  uint64_t borrow = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    uint64_t t = (uint64_t)a.v[i] - b.v[i] - borrow;
    r.v[i] = (uint32_t)t;
    borrow = (t >> 32) & 1;
  }
  // add m back when the subtraction borrowed
  const uint32_t mask = 0u - (uint32_t)borrow;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    c = (uint64_t)r.v[i] + (P::m(i) & mask) + (c >> 32);
    r.v[i] = (uint32_t)c;
  }
  return r;
}

template <class P>
__device__ __forceinline__ Field<P> dbl(const Field<P>& a) { return add(a, a); }

template <class P>
__device__ __forceinline__ Field<P> neg(const Field<P>& a) {
  return sub(Field<P>::zero(), a);
}

// CIOS Montgomery product a*b*2^(-32N) mod m.  Requires a*b < m * 2^(32N), which
// holds for a < 2^(32N) and b < m; output fully reduced.
template <class P>
__device__ __forceinline__ Field<P> mul_cios(const Field<P>& a, const Field<P>& b) {
  constexpr int N = P::N;
  uint32_t t[N + 2];
#pragma unroll
  for (int i = 0; i < N + 2; i++) t[i] = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < N; j++) {
      c = (uint64_t)a.v[j] * b.v[i] + (uint64_t)t[j] + (c >> 32);
      t[j] = (uint32_t)c;
    }
    uint64_t s = (uint64_t)t[N] + (c >> 32);
    t[N] = (uint32_t)s;
    t[N + 1] = (uint32_t)(s >> 32);
    const uint32_t mq = t[0] * P::INV;
    c = (uint64_t)mq * P::m(0) + (uint64_t)t[0];
#pragma unroll
    for (int j = 1; j < N; j++) {
      c = (uint64_t)mq * P::m(j) + (uint64_t)t[j] + (c >> 32);
      t[j - 1] = (uint32_t)c;
    }
    s = (uint64_t)t[N] + (c >> 32);
    t[N - 1] = (uint32_t)s;
    t[N] = t[N + 1] + (uint32_t)(s >> 32);
  }
  Field<P> r;
#pragma unroll
  for (int i = 0; i < N; i++) r.v[i] = t[i];
  // t < 2m (and t[N] == 0 because 2m < 2^(32N) for both moduli)
  reduce_once(r);
  return r;
}

// (hi:acc) += a * b, with the 64-bit carry-out of v_mad_u64_u32 added into hi.
__device__ __forceinline__ void mac3(uint64_t& acc, uint32_t& hi, uint32_t a, uint32_t b) {
  uint64_t cc;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(acc), "=s"(cc) : "v"(a), "v"(b), "0"(acc));
  asm("v_addc_co_u32_e64 %0, %1, 0, %2, %3" : "=v"(hi), "=s"(cc) : "v"(hi), "s"(cc));
}

// Same with a wave-uniform second factor (a modulus limb) read from an SGPR.
__device__ __forceinline__ void mac3_s(uint64_t& acc, uint32_t& hi, uint32_t a, uint32_t b_uniform) {
  uint64_t cc;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(acc), "=s"(cc) : "v"(a), "s"(b_uniform), "0"(acc));
  asm("v_addc_co_u32_e64 %0, %1, 0, %2, %3" : "=v"(hi), "=s"(cc) : "v"(hi), "s"(cc));
}

// Montgomery product by finely integrated product scanning (Koc, Acar, Kaliski
// 1996, FIPS): column k accumulates a_j b_(k-j) and m_j m(k-j) in a three-word
// accumulator, each product one v_mad_u64_u32 plus one carry add.
template <class P>
__device__ __forceinline__ Field<P> mul_fips(const Field<P>& a, const Field<P>& b) {
  constexpr int N = P::N;
  uint32_t m[N];
  uint32_t r[N];
  uint64_t acc = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
#pragma unroll
    for (int j = 0; j < k; j++) {
      mac3(acc, hi, a.v[j], b.v[k - j]);
      mac3_s(acc, hi, m[j], P::m(k - j));
    }
    mac3(acc, hi, a.v[k], b.v[0]);
    m[k] = (uint32_t)acc * P::INV;
    mac3_s(acc, hi, m[k], P::m(0));
    acc = (acc >> 32) | ((uint64_t)hi << 32);
    hi = 0;
  }
#pragma unroll
  for (int k = N; k < 2 * N - 1; k++) {
#pragma unroll
    for (int j = k - N + 1; j < N; j++) {
      mac3(acc, hi, a.v[j], b.v[k - j]);
      mac3_s(acc, hi, m[j], P::m(k - j));
    }
    r[k - N] = (uint32_t)acc;
    acc = (acc >> 32) | ((uint64_t)hi << 32);
    hi = 0;
  }
  r[N - 1] = (uint32_t)acc;
  Field<P> out;
#pragma unroll
  for (int i = 0; i < N; i++) out.v[i] = r[i];
  // result < 2m < 2^(32N), and acc >> 32 is 0 for both moduli
  reduce_once(out);
  return out;
}

// The product used everywhere: FIPS, measured 1.44x the CIOS form's throughput on
// gfx950 (723 against 1290 VALU instructions per fp product; CIOS's 64-bit carry
// chains compile to a v_mov and a v_lshl_add_u64 per limb product).
template <class P>
__device__ __forceinline__ Field<P> mul(const Field<P>& a, const Field<P>& b) { return mul_fips(a, b); }

template <class P>
__device__ __forceinline__ Field<P> sqr(const Field<P>& a) { return mul(a, a); }

// Montgomery -> canonical (multiply by 1).
template <class P>
__device__ __forceinline__ Field<P> from_mont(const Field<P>& a) {
  Field<P> one_raw = Field<P>::zero();
  one_raw.v[0] = 1;
  return mul(a, one_raw);
}

// a^e for a small public exponent.
template <class P>
__device__ __forceinline__ Field<P> pow_u32(Field<P> a, uint32_t e) {
  Field<P> r = Field<P>::one();
  while (e) {
    if (e & 1) r = mul(r, a);
    a = sqr(a);
    e >>= 1;
  }
  return r;
}

// Loads / stores of AoS elements (N u32 per element) with 16-byte vector accesses.
template <class P>
__device__ __forceinline__ Field<P> load(const uint32_t* __restrict__ p) {
  Field<P> r;
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int i = 0; i < P::N / 4; i++) {
    uint4 w = q[i];
    r.v[4 * i + 0] = w.x;
    r.v[4 * i + 1] = w.y;
    r.v[4 * i + 2] = w.z;
    r.v[4 * i + 3] = w.w;
  }
  return r;
}

template <class P>
__device__ __forceinline__ void store(uint32_t* __restrict__ p, const Field<P>& a) {
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < P::N / 4; i++) q[i] = make_uint4(a.v[4 * i], a.v[4 * i + 1], a.v[4 * i + 2], a.v[4 * i + 3]);
}

}  // namespace mzk
