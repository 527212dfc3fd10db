// Host-side fp / G1 arithmetic for the short serial tails of the device pipeline
// (the last few dozen point operations of an MSM and the one-time tables of the
// SRS generator).  A dependent chain of point operations costs ~30 us on one GPU
// wave (2.1 us per 381-bit product, measured) against well under 1 us on one host
// core, so chains shorter than a few hundred operations finish here.
//
// Limbs: 6 x u64 little-endian Montgomery (R = 2^384), the same bytes as the
// device's 12 x u32 limbs.  Points: XYZZ (x = X/ZZ, y = Y/ZZZ), identity ZZ = 0.
#pragma once
#include <stdint.h>
#include <string.h>

namespace mzk {
namespace host {

typedef unsigned __int128 u128;

struct Fp {
  uint64_t l[6];
};

static const uint64_t kP[6] = {0xb9feffffffffaaabull, 0x1eabfffeb153ffffull, 0x6730d2a0f6b0f624ull,
                               0x64774b84f38512bfull, 0x4b1ba7b6434bacd7ull, 0x1a0111ea397fe69aull};
static const uint64_t kPInv = 0x89f3fffcfffcfffdull;   // -p^-1 mod 2^64
static const uint64_t kOne[6] = {0x760900000002fffdull, 0xebf4000bc40c0002ull, 0x5f48985753c758baull,
                                 0x77ce585370525745ull, 0x5c071a97a256ec6dull, 0x15f65ec3fa80e493ull};

inline Fp fp_zero() { Fp r; memset(r.l, 0, sizeof r.l); return r; }
inline Fp fp_one() { Fp r; memcpy(r.l, kOne, sizeof r.l); return r; }
inline bool fp_is_zero(const Fp& a) { return !(a.l[0] | a.l[1] | a.l[2] | a.l[3] | a.l[4] | a.l[5]); }
inline bool fp_eq(const Fp& a, const Fp& b) { return memcmp(a.l, b.l, sizeof a.l) == 0; }

inline bool ge_p(const uint64_t* a) {
  for (int i = 5; i >= 0; i--) {
    if (a[i] != kP[i]) return a[i] > kP[i];
  }
  return true;
}

inline void sub_p(uint64_t* a) {
  uint64_t br = 0;
  for (int i = 0; i < 6; i++) {
    u128 d = (u128)a[i] - kP[i] - br;
    a[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
}

inline Fp fp_add(const Fp& a, const Fp& b) {
  Fp r;
  uint64_t c = 0;
  for (int i = 0; i < 6; i++) {
    u128 s = (u128)a.l[i] + b.l[i] + c;
    r.l[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  if (ge_p(r.l)) sub_p(r.l);   // 2p < 2^384: no carry out
  return r;
}

inline Fp fp_sub(const Fp& a, const Fp& b) {
  Fp r;
  uint64_t br = 0;
  for (int i = 0; i < 6; i++) {
    u128 d = (u128)a.l[i] - b.l[i] - br;
    r.l[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  if (br) {
    uint64_t c = 0;
    for (int i = 0; i < 6; i++) {
      u128 s = (u128)r.l[i] + kP[i] + c;
      r.l[i] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
  }
  return r;
}

inline Fp fp_mul(const Fp& a, const Fp& b) {   // CIOS
  uint64_t t[8] = {0};
  for (int i = 0; i < 6; i++) {
    u128 c = 0;
    for (int j = 0; j < 6; j++) {
      c = (u128)a.l[j] * b.l[i] + t[j] + (uint64_t)(c >> 64);
      t[j] = (uint64_t)c;
    }
    u128 s = (u128)t[6] + (uint64_t)(c >> 64);
    t[6] = (uint64_t)s;
    t[7] = (uint64_t)(s >> 64);
    uint64_t m = t[0] * kPInv;
    c = (u128)m * kP[0] + t[0];
    for (int j = 1; j < 6; j++) {
      c = (u128)m * kP[j] + t[j] + (uint64_t)(c >> 64);
      t[j - 1] = (uint64_t)c;
    }
    s = (u128)t[6] + (uint64_t)(c >> 64);
    t[5] = (uint64_t)s;
    t[6] = t[7] + (uint64_t)(s >> 64);
  }
  Fp r;
  memcpy(r.l, t, sizeof r.l);
  if (ge_p(r.l)) sub_p(r.l);
  return r;
}

inline Fp fp_sqr(const Fp& a) { return fp_mul(a, a); }
inline Fp fp_dbl(const Fp& a) { return fp_add(a, a); }

// a^(p-2) by square-and-multiply.
inline Fp fp_inv(const Fp& a) {
  uint64_t e[6];
  memcpy(e, kP, sizeof e);
  e[0] -= 2;
  Fp r = fp_one();
  for (int i = 383; i >= 0; i--) {
    r = fp_sqr(r);
    if ((e[i / 64] >> (i % 64)) & 1) r = fp_mul(r, a);
  }
  return r;
}

struct Xyzz {
  Fp X, Y, ZZ, ZZZ;
  bool is_inf() const { return fp_is_zero(ZZ); }
  static Xyzz inf() { Xyzz r; r.X = r.Y = r.ZZ = r.ZZZ = fp_zero(); return r; }
};

struct Affine {
  Fp x, y;
  bool is_inf() const { return fp_is_zero(x) && fp_is_zero(y); }
};

inline Xyzz xyzz_dbl(const Xyzz& p) {
  if (p.is_inf()) return p;
  Fp U = fp_dbl(p.Y), V = fp_sqr(U), W = fp_mul(U, V), S = fp_mul(p.X, V);
  Fp x2 = fp_sqr(p.X), M = fp_add(fp_dbl(x2), x2);
  Xyzz r;
  r.X = fp_sub(fp_sqr(M), fp_dbl(S));
  r.Y = fp_sub(fp_mul(M, fp_sub(S, r.X)), fp_mul(W, p.Y));
  r.ZZ = fp_mul(V, p.ZZ);
  r.ZZZ = fp_mul(W, p.ZZZ);
  return r;
}

inline Xyzz xyzz_add(const Xyzz& p, const Xyzz& q) {
  if (p.is_inf()) return q;
  if (q.is_inf()) return p;
  Fp U1 = fp_mul(p.X, q.ZZ), U2 = fp_mul(q.X, p.ZZ);
  Fp S1 = fp_mul(p.Y, q.ZZZ), S2 = fp_mul(q.Y, p.ZZZ);
  Fp P = fp_sub(U2, U1), R = fp_sub(S2, S1);
  if (fp_is_zero(P)) {
    if (fp_is_zero(R)) return xyzz_dbl(p);
    return Xyzz::inf();
  }
  Fp PP = fp_sqr(P), PPP = fp_mul(P, PP), Q = fp_mul(U1, PP);
  Xyzz r;
  r.X = fp_sub(fp_sub(fp_sqr(R), PPP), fp_dbl(Q));
  r.Y = fp_sub(fp_mul(R, fp_sub(Q, r.X)), fp_mul(S1, PPP));
  r.ZZ = fp_mul(fp_mul(p.ZZ, q.ZZ), PP);
  r.ZZZ = fp_mul(fp_mul(p.ZZZ, q.ZZZ), PPP);
  return r;
}

inline Xyzz xyzz_from_affine(const Affine& a) {
  Xyzz r;
  r.X = a.x;
  r.Y = a.y;
  if (a.is_inf()) { r.ZZ = r.ZZZ = fp_zero(); } else { r.ZZ = r.ZZZ = fp_one(); }
  return r;
}

// Jacobian (X', Y', Z') with Z' = ZZ ZZZ (see g1_to_jacobian on the device).
inline void xyzz_to_jacobian(const Xyzz& p, uint64_t out[18]) {
  if (p.is_inf()) { memset(out, 0, 18 * 8); return; }
  Fp zzz2 = fp_sqr(p.ZZZ), zz2 = fp_sqr(p.ZZ);
  Fp Z = fp_mul(p.ZZ, p.ZZZ);
  Fp X = fp_mul(fp_mul(p.X, p.ZZ), zzz2);
  Fp Y = fp_mul(fp_mul(p.Y, fp_mul(zz2, p.ZZ)), zzz2);
  memcpy(out, X.l, 48);
  memcpy(out + 6, Y.l, 48);
  memcpy(out + 12, Z.l, 48);
}

inline Affine xyzz_to_affine(const Xyzz& p) {
  Affine a;
  if (p.is_inf()) { a.x = a.y = fp_zero(); return a; }
  a.x = fp_mul(p.X, fp_inv(p.ZZ));
  a.y = fp_mul(p.Y, fp_inv(p.ZZZ));
  return a;
}


// ---------------------------------------------------------------- host fr (r, 4 x u64, R = 2^256)
struct Fr {
  uint64_t l[4];
};

static const uint64_t kR[4] = {0xffffffff00000001ull, 0x53bda402fffe5bfeull, 0x3339d80809a1d805ull,
                               0x73eda753299d7d48ull};
static const uint64_t kRInv = 0xfffffffeffffffffull;
static const uint64_t kFrOne[4] = {0x00000001fffffffeull, 0x5884b7fa00034802ull, 0x998c4fefecbc4ff5ull,
                                   0x1824b159acc5056full};
static const uint64_t kFrR2[4] = {0xc999e990f3f29c6dull, 0x2b6cedcb87925c23ull, 0x05d314967254398full,
                                  0x0748d9d99f59ff11ull};
// 2^32-th root of unity 7^((r-1)/2^32) (fq.rs:70-77) and zeta = 7^((r-1)/3) (fq.rs:700-710), Montgomery
static const uint64_t kFrRoot[4] = {0xb9b58d8c5f0e466aull, 0x5b1b4c801819d7ecull, 0x0af53ae352a31e64ull,
                                    0x5bf3adda19e9b27bull};
static const uint64_t kFrZeta[4] = {0x92d9090b093011d2ull, 0xfc9cbd719d6aa073ull, 0xc1f14ef0cd65a1a6ull,
                                    0x017f6d35e72fcdebull};

inline Fr fr_zero() { Fr r; memset(r.l, 0, sizeof r.l); return r; }
inline Fr fr_one() { Fr r; memcpy(r.l, kFrOne, sizeof r.l); return r; }
inline Fr fr_const(const uint64_t* v) { Fr r; memcpy(r.l, v, sizeof r.l); return r; }
inline bool fr_eq(const Fr& a, const Fr& b) { return memcmp(a.l, b.l, sizeof a.l) == 0; }

inline bool fr_ge(const uint64_t* a) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] != kR[i]) return a[i] > kR[i];
  }
  return true;
}
inline void fr_sub_mod(uint64_t* a) {
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a[i] - kR[i] - br;
    a[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
}

inline Fr fr_mul(const Fr& a, const Fr& b) {
  uint64_t t[6] = {0};
  for (int i = 0; i < 4; i++) {
    u128 c = 0;
    for (int j = 0; j < 4; j++) {
      c = (u128)a.l[j] * b.l[i] + t[j] + (uint64_t)(c >> 64);
      t[j] = (uint64_t)c;
    }
    u128 s = (u128)t[4] + (uint64_t)(c >> 64);
    t[4] = (uint64_t)s;
    t[5] = (uint64_t)(s >> 64);
    uint64_t m = t[0] * kRInv;
    c = (u128)m * kR[0] + t[0];
    for (int j = 1; j < 4; j++) {
      c = (u128)m * kR[j] + t[j] + (uint64_t)(c >> 64);
      t[j - 1] = (uint64_t)c;
    }
    s = (u128)t[4] + (uint64_t)(c >> 64);
    t[3] = (uint64_t)s;
    t[4] = t[5] + (uint64_t)(s >> 64);
  }
  Fr r;
  memcpy(r.l, t, sizeof r.l);
  if (t[4] || fr_ge(r.l)) fr_sub_mod(r.l);
  return r;
}

inline Fr fr_sub(const Fr& a, const Fr& b) {
  Fr r;
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a.l[i] - b.l[i] - br;
    r.l[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  if (br) {
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) {
      u128 s = (u128)r.l[i] + kR[i] + c;
      r.l[i] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
  }
  return r;
}

inline Fr fr_add(const Fr& a, const Fr& b) {   // r < 2^255: the sum fits four limbs
  Fr r;
  uint64_t c = 0;
  for (int i = 0; i < 4; i++) {
    u128 s = (u128)a.l[i] + b.l[i] + c;
    r.l[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  if (fr_ge(r.l)) fr_sub_mod(r.l);
  return r;
}

inline Fr fr_pow(Fr a, uint64_t e) {
  Fr r = fr_one();
  while (e) {
    if (e & 1) r = fr_mul(r, a);
    a = fr_mul(a, a);
    e >>= 1;
  }
  return r;
}

// a^(r-2)
inline Fr fr_inv(const Fr& a) {
  uint64_t e[4];
  memcpy(e, kR, sizeof e);
  e[0] -= 2;
  Fr r = fr_one();
  for (int i = 255; i >= 0; i--) {
    r = fr_mul(r, r);
    if ((e[i / 64] >> (i % 64)) & 1) r = fr_mul(r, a);
  }
  return r;
}

inline Fr fr_from_u64(uint64_t v) {
  Fr t;
  memset(t.l, 0, sizeof t.l);
  t.l[0] = v;
  return fr_mul(t, fr_const(kFrR2));
}

// omega = ROOT^(2^(32 - log_n)) (domain.rs:72-91)
inline Fr fr_root_of_unity(unsigned log_n) {
  Fr w = fr_const(kFrRoot);
  for (unsigned i = log_n; i < 32; i++) w = fr_mul(w, w);
  return w;
}

}  // namespace host
}  // namespace mzk
