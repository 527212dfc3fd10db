// G1 of BLS12-381 (y^2 = x^3 + 4 over fp, curves/src/bls12_381/g1.rs:40-50) on gfx950.
//
// Device points use XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2), the
// cheapest system for mixed additions into buckets: 8M + 2S per affine addition
// against 7M + 4S for Jacobian.  Formulas are the EFD "xyzz" set for a = 0
// (madd-2008-s, mdbl-2008-s-1, add-2008-s, dbl-2008-s-1), with the exceptional
// cases (equal points, P + (-P), identity operands) handled exactly so results
// match the textbook group law bit for bit after normalisation.
//
// Affine points use blst_p1_affine's layout (x[12] || y[12] u32, Montgomery;
// identity = all zero, g1.rs:32-34); Jacobian output uses blst_p1's layout
// (X || Y || Z, identity Z = 0, g1.rs:493-495).
#pragma once
#include "field.hpp"

namespace mzk {

struct g1_aff_t {
  fp_t x, y;
  __device__ __forceinline__ bool is_inf() const { return x.is_zero() && y.is_zero(); }
};

struct g1_xyzz_t {
  fp_t X, Y, ZZ, ZZZ;
  __device__ __forceinline__ bool is_inf() const { return ZZ.is_zero(); }
  __device__ __forceinline__ static g1_xyzz_t inf() {
    g1_xyzz_t r;
    r.X = fp_t::zero();
    r.Y = fp_t::zero();
    r.ZZ = fp_t::zero();
    r.ZZZ = fp_t::zero();
    return r;
  }
  __device__ __forceinline__ static g1_xyzz_t from_affine(const g1_aff_t& a) {
    g1_xyzz_t r;
    r.X = a.x;
    r.Y = a.y;
    if (a.is_inf()) {
      r.ZZ = fp_t::zero();
      r.ZZZ = fp_t::zero();
    } else {
      r.ZZ = fp_t::one();
      r.ZZZ = fp_t::one();
    }
    return r;
  }
};

__device__ __forceinline__ g1_aff_t g1_neg(const g1_aff_t& a) {
  g1_aff_t r;
  r.x = a.x;
  r.y = neg(a.y);   // neg(0) = 0 keeps the identity encoding
  return r;
}

// 2 * (x, y) for an affine point (mdbl-2008-s-1).
__device__ __forceinline__ g1_xyzz_t g1_dbl_affine(const g1_aff_t& q) {
  g1_xyzz_t r;
  fp_t U = dbl(q.y);
  fp_t V = sqr(U);
  fp_t W = mul(U, V);
  fp_t S = mul(q.x, V);
  fp_t x2 = sqr(q.x);
  fp_t M = add(dbl(x2), x2);
  r.X = sub(sqr(M), dbl(S));
  r.Y = sub(mul(M, sub(S, r.X)), mul(W, q.y));
  r.ZZ = V;     // zero iff y == 0, i.e. the result is the identity
  r.ZZZ = W;
  return r;
}

// 2 * P in XYZZ (dbl-2008-s-1).
__device__ __forceinline__ g1_xyzz_t g1_dbl(const g1_xyzz_t& p) {
  g1_xyzz_t r;
  fp_t U = dbl(p.Y);
  fp_t V = sqr(U);
  fp_t W = mul(U, V);
  fp_t S = mul(p.X, V);
  fp_t x2 = sqr(p.X);
  fp_t M = add(dbl(x2), x2);
  r.X = sub(sqr(M), dbl(S));
  r.Y = sub(mul(M, sub(S, r.X)), mul(W, p.Y));
  r.ZZ = mul(V, p.ZZ);
  r.ZZZ = mul(W, p.ZZZ);
  return r;
}

// P + Q with Q affine (madd-2008-s); Q may be the identity.
__device__ __forceinline__ g1_xyzz_t g1_add_affine(const g1_xyzz_t& p, const g1_aff_t& q) {
  if (q.is_inf()) return p;
  if (p.is_inf()) return g1_xyzz_t::from_affine(q);
  fp_t U2 = mul(q.x, p.ZZ);
  fp_t S2 = mul(q.y, p.ZZZ);
  fp_t P = sub(U2, p.X);
  fp_t R = sub(S2, p.Y);
  if (P.is_zero()) {
    if (R.is_zero()) return g1_dbl_affine(q);
    return g1_xyzz_t::inf();
  }
  fp_t PP = sqr(P);
  fp_t PPP = mul(P, PP);
  fp_t Q = mul(p.X, PP);
  g1_xyzz_t r;
  r.X = sub(sub(sqr(R), PPP), dbl(Q));
  r.Y = sub(mul(R, sub(Q, r.X)), mul(p.Y, PPP));
  r.ZZ = mul(p.ZZ, PP);
  r.ZZZ = mul(p.ZZZ, PPP);
  return r;
}

// P + Q, both XYZZ (add-2008-s).
__device__ __forceinline__ g1_xyzz_t g1_add(const g1_xyzz_t& p, const g1_xyzz_t& q) {
  if (p.is_inf()) return q;
  if (q.is_inf()) return p;
  fp_t U1 = mul(p.X, q.ZZ);
  fp_t U2 = mul(q.X, p.ZZ);
  fp_t S1 = mul(p.Y, q.ZZZ);
  fp_t S2 = mul(q.Y, p.ZZZ);
  fp_t P = sub(U2, U1);
  fp_t R = sub(S2, S1);
  if (P.is_zero()) {
    if (R.is_zero()) return g1_dbl(p);
    return g1_xyzz_t::inf();
  }
  fp_t PP = sqr(P);
  fp_t PPP = mul(P, PP);
  fp_t Q = mul(U1, PP);
  g1_xyzz_t r;
  r.X = sub(sub(sqr(R), PPP), dbl(Q));
  r.Y = sub(mul(R, sub(Q, r.X)), mul(S1, PPP));
  r.ZZ = mul(mul(p.ZZ, q.ZZ), PP);
  r.ZZZ = mul(mul(p.ZZZ, q.ZZZ), PPP);
  return r;
}

// XYZZ -> Jacobian (X', Y', Z') with Z' = ZZ * ZZZ:
//   X' = x Z'^2 = X ZZ ZZZ^2,  Y' = y Z'^3 = Y ZZ^3 ZZZ^2.
__device__ __forceinline__ void g1_to_jacobian(const g1_xyzz_t& p, fp_t& X, fp_t& Y, fp_t& Z) {
  if (p.is_inf()) {
    X = fp_t::zero();
    Y = fp_t::zero();
    Z = fp_t::zero();
    return;
  }
  fp_t zzz2 = sqr(p.ZZZ);
  fp_t zz2 = sqr(p.ZZ);
  Z = mul(p.ZZ, p.ZZZ);
  X = mul(mul(p.X, p.ZZ), zzz2);
  Y = mul(mul(p.Y, mul(zz2, p.ZZ)), zzz2);
}

__device__ __forceinline__ g1_aff_t load_affine(const uint32_t* __restrict__ p) {
  g1_aff_t a;
  a.x = load<FpP>(p);
  a.y = load<FpP>(p + 12);
  return a;
}

__device__ __forceinline__ g1_xyzz_t load_xyzz(const uint32_t* __restrict__ p) {
  g1_xyzz_t a;
  a.X = load<FpP>(p);
  a.Y = load<FpP>(p + 12);
  a.ZZ = load<FpP>(p + 24);
  a.ZZZ = load<FpP>(p + 36);
  return a;
}

__device__ __forceinline__ void store_xyzz(uint32_t* __restrict__ p, const g1_xyzz_t& a) {
  store<FpP>(p, a.X);
  store<FpP>(p + 12, a.Y);
  store<FpP>(p + 24, a.ZZ);
  store<FpP>(p + 36, a.ZZZ);
}

}  // namespace mzk
