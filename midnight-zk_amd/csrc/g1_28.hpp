// XYZZ bucket accumulation and bucket reduction in the radix-2^28 field form of
// fp28.hpp.
//
// Same formulas as g1.hpp (madd-2008-s with the exceptional cases handled exactly,
// mdbl-2008-s-1 for the doubling case), restated on f28 with the value bounds the
// form needs.  Invariants of an accumulator entering or leaving g1_28_madd:
//   X < 10p, Y < 6p (normalised; the formulas below leave Y < 2p), ZZ, ZZZ < 2p
//   (product outputs), or the identity (ZZ literally zero).
// Affine inputs come straight from HBM in the blst layout; from_std leaves them
// below 2^389 < 256 p.
#pragma once
#include "fp28.hpp"

// words of one point in this form (4 coordinates x 14 limbs; 224 bytes, 16-byte aligned)
#define MZK_X28_WORDS 56

namespace mzk {

struct g1x28 {
  f28 X, Y, ZZ, ZZZ;
};

MZK_HD g1x28 g1x28_inf() {
  g1x28 r;
  r.X = f28_zero();
  r.Y = f28_zero();
  r.ZZ = f28_zero();
  r.ZZZ = f28_zero();
  return r;
}

MZK_HD bool g1x28_is_inf(const g1x28& p) { return f28_is_literal_zero(p.ZZ); }

// An affine input as an accumulator: coordinates brought below 2p (one product each
// by 2^392 mod p), y negated when neg.
MZK_HD g1x28 g1x28_from_affine(const f28& x, const f28& y, bool neg) {
  g1x28 r;
  const f28 one = f28_one();
  r.X = f28_mul(x, one);                      // < 2p
  f28 yy = f28_mul(y, one);                   // < 2p
  if (neg) yy = f28_norm(f28_sub<4, 1>(f28_zero(), yy));   // 4p - y < 4p
  r.Y = yy;
  r.ZZ = one;
  r.ZZZ = one;
  return r;
}

// The same for an affine input already below 2p and normalised (the resident SRS's
// accumulation records): no product, so a lane that opens a bucket costs its wave a
// few selects instead of two products.
MZK_HD g1x28 g1x28_from_reduced(const f28& x, const f28& y, bool neg) {
  g1x28 r;
  const f28 one = f28_one();
  r.X = x;
  f28 yy = y;
  if (neg) yy = f28_norm(f28_sub<4, 1>(f28_zero(), y));   // 4p - y < 4p
  r.Y = yy;
  r.ZZ = one;
  r.ZZZ = one;
  return r;
}

// 2 (x, +-y) for an affine point (mdbl-2008-s-1); coordinates as from_affine makes
// them (x < 2p, y < 4p, normalised).
MZK_HD g1x28 g1x28_dbl_affine(const f28& x, const f28& y) {
  g1x28 r;
  const f28 U = f28_add(y, y);                     // < 8p, limbs < 2^29
  const f28 V = f28_sqr(U);                        // 64 p^2 -> < 2p
  const f28 W = f28_mul(U, V);
  const f28 S = f28_mul(x, V);
  const f28 x2 = f28_sqr(x);
  const f28 M = f28_add(f28_add(x2, x2), x2);      // < 6p, limbs < 3 2^28
  const f28 S2 = f28_add(S, S);                    // < 4p, limbs < 2^29
  // X3 = M^2 + 8p - 2S: lower limbs of 8p raised by 2 2^28 cover 2S's (< 2^29)
  r.X = f28_norm(f28_sub<8, 2>(f28_sqr(M), S2));   // < 10p
  const f28 T = f28_sub<16, 1>(S, r.X);            // S + 16p - X3 < 18p
  // Y3 = M T + W (8p - y): 108 p^2 + 16 p^2, one reduction -> < 2p
  r.Y = f28_mul_sum(M, T, W, f28_sub<8, 1>(f28_zero(), y));
  r.ZZ = V;
  r.ZZZ = W;
  return r;
}

// acc + (x, +-y) (madd-2008-s).  x, y from from_std (< 2^389), or below 2p and
// normalised when REDUCED; the affine identity is the caller's to skip.
template <bool REDUCED = false>
MZK_HD g1x28 g1x28_madd(const g1x28& p, const f28& x, const f28& y, bool neg) {
  if (g1x28_is_inf(p)) {
    if constexpr (REDUCED) return g1x28_from_reduced(x, y, neg);
    else return g1x28_from_affine(x, y, neg);
  }
  const f28 U2 = f28_mul(x, p.ZZ);                 // 512 p^2 -> < 2p
  const f28 S2 = f28_mul(y, p.ZZZ);                // < 2p
  // P = U2 - X < 18p (X < 10p; 16p's top limb minus 1 covers X's)
  const f28 P = f28_sub<16, 1>(U2, p.X);
  // R = +-S2 - Y: S2 + 8p - Y < 10p, or 16p - S2 - Y < 16p (lower limbs raised by 2 2^28)
  const f28 R = neg ? f28_neg_sum<16, 2>(S2, p.Y) : f28_sub<8, 1>(S2, p.Y);
  if (__builtin_expect(f28_is_zero_mod_p<18>(P), 0)) {
    if (f28_is_zero_mod_p<16>(R)) {
      if constexpr (REDUCED) {
        const g1x28 q = g1x28_from_reduced(x, y, neg);
        return g1x28_dbl_affine(q.X, q.Y);
      } else {
        const g1x28 q = g1x28_from_affine(x, y, neg);
        return g1x28_dbl_affine(q.X, q.Y);
      }
    }
    return g1x28_inf();
  }
  const f28 PP = f28_sqr(P);                       // 324 p^2 -> < 2p
  const f28 PPP = f28_mul(P, PP);
  const f28 Q = f28_mul(p.X, PP);
  g1x28 r;
  // X3 = R^2 + 8p - PPP - 2Q: lower limbs of 8p raised by 3 2^28 cover PPP + 2Q
  const f28 Q2 = f28_add(Q, Q);
  r.X = f28_norm(f28_sub<8, 3>(f28_sqr(R), f28_add(PPP, Q2)));   // < 10p
  // Y3 = R (Q + 16p - X3) + Y (4p - PPP): 288 p^2 + 24 p^2, one reduction -> < 2p
  const f28 T = f28_sub<16, 1>(Q, r.X);            // < 18p
  r.Y = f28_mul_sum(R, T, p.Y, f28_sub<4, 1>(f28_zero(), PPP));
  r.ZZ = f28_mul(p.ZZ, PP);
  r.ZZZ = f28_mul(p.ZZZ, PPP);
  return r;
}

// 2 P (dbl-2008-s-1) for P within the accumulator invariants
MZK_HD g1x28 g1x28_dbl(const g1x28& p) {
  g1x28 r;
  const f28 U = f28_add(p.Y, p.Y);                 // < 12p, limbs < 2^29
  const f28 V = f28_sqr(U);                        // 144 p^2 -> < 2p
  const f28 W = f28_mul(U, V);
  const f28 S = f28_mul(p.X, V);
  const f28 x2 = f28_sqr(p.X);                     // 100 p^2
  const f28 M = f28_add(f28_add(x2, x2), x2);      // < 6p, limbs < 3 2^28
  r.X = f28_norm(f28_sub<8, 2>(f28_sqr(M), f28_add(S, S)));      // < 10p
  const f28 T = f28_sub<16, 1>(S, r.X);            // < 18p
  // Y3 = M T + W (8p - Y): 108 p^2 + 16 p^2, one reduction -> < 2p
  r.Y = f28_mul_sum(M, T, W, f28_sub<8, 1>(f28_zero(), p.Y));
  r.ZZ = f28_mul(V, p.ZZ);
  r.ZZZ = f28_mul(W, p.ZZZ);
  return r;
}

// P + Q (add-2008-s), both within the accumulator invariants
MZK_HD g1x28 g1x28_add(const g1x28& p, const g1x28& q) {
  if (g1x28_is_inf(p)) return q;
  if (g1x28_is_inf(q)) return p;
  const f28 U1 = f28_mul(p.X, q.ZZ);              // 20 p^2 -> < 2p
  const f28 U2 = f28_mul(q.X, p.ZZ);
  const f28 S1 = f28_mul(p.Y, q.ZZZ);
  const f28 S2 = f28_mul(q.Y, p.ZZZ);
  const f28 P = f28_sub<4, 1>(U2, U1);            // < 6p
  const f28 R = f28_sub<4, 1>(S2, S1);            // < 6p
  if (f28_is_zero_mod_p<6>(P)) {
    if (f28_is_zero_mod_p<6>(R)) return g1x28_dbl(p);
    return g1x28_inf();
  }
  const f28 PP = f28_sqr(P);
  const f28 PPP = f28_mul(P, PP);
  const f28 Q = f28_mul(U1, PP);
  g1x28 r;
  r.X = f28_norm(f28_sub<8, 3>(f28_sqr(R), f28_add(PPP, f28_add(Q, Q))));   // < 10p
  const f28 T = f28_sub<16, 1>(Q, r.X);            // < 18p
  // Y3 = R T + S1 (4p - PPP): 108 p^2 + 8 p^2, one reduction -> < 2p
  r.Y = f28_mul_sum(R, T, S1, f28_sub<4, 1>(f28_zero(), PPP));
  r.ZZ = f28_mul(f28_mul(p.ZZ, q.ZZ), PP);
  r.ZZZ = f28_mul(f28_mul(p.ZZZ, q.ZZZ), PPP);
  return r;
}

// accumulator -> blst-layout XYZZ (4 x 12 canonical Montgomery limbs)
MZK_HD void g1x28_to_std(const g1x28& p, uint32_t out[48]) {
  if (g1x28_is_inf(p)) {
    for (int i = 0; i < 48; i++) out[i] = 0;
    return;
  }
  f28_to_std(p.X, out);
  f28_to_std(p.Y, out + 12);
  f28_to_std(p.ZZ, out + 24);
  f28_to_std(p.ZZZ, out + 36);
}

}  // namespace mzk

#if defined(__HIPCC__)
namespace mzk {
__device__ __forceinline__ g1x28 load_x28(const uint32_t* __restrict__ p) {
  g1x28 r;
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint32_t w[MZK_X28_WORDS];
#pragma unroll
  for (int i = 0; i < MZK_X28_WORDS / 4; i++) {
    const uint4 v = q[i];
    w[4 * i] = v.x;
    w[4 * i + 1] = v.y;
    w[4 * i + 2] = v.z;
    w[4 * i + 3] = v.w;
  }
#pragma unroll
  for (int i = 0; i < 14; i++) {
    r.X.l[i] = w[i];
    r.Y.l[i] = w[14 + i];
    r.ZZ.l[i] = w[28 + i];
    r.ZZZ.l[i] = w[42 + i];
  }
  return r;
}

// ---- quad-cooperative addition: the four lanes of an aligned quad hold the same P and Q
// and compute P + Q together, each lane one of the formula's independent products per
// step (lane q's operands chosen by selects, so the quad runs one instruction stream), the
// products shared by DPP quad broadcasts.  The dependent chain is 4 product steps instead
// of the 13 reductions of g1x28_add: for reductions whose waves are too few to fill the
// SIMDs, where each addition's latency is the cost.  Same formulas and bounds as
// g1x28_add, so the result is the same point (the same representative where the
// products' operands coincide: Montgomery outputs depend only on the product's value).
template <int K>   // lane K of the quad, to every lane of the quad (DPP quad_perm K,K,K,K)
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, K * 0x55, 0xF, 0xF, false);
}

template <int K>
__device__ __forceinline__ f28 f28_quad_bcast(const f28& a) {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = quad_bcast<K>(a.l[i]);
  return r;
}

// a or b, limb by limb, by a per-lane condition: one v_cndmask per limb.  Both operands
// are registers here; a select between two loaded values could turn into a load from a
// selected address (the loads of Q below take their address by lane instead).
__device__ __forceinline__ f28 f28_pick(bool c, const f28& a, const f28& b) {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = c ? a.l[i] : b.l[i];
  return r;
}

__device__ __forceinline__ f28 load_f28(const uint32_t* __restrict__ p) {
  f28 r;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) r.l[i] = p[i];
  return r;
}

// P + Q on a quad (every lane of the quad active, holding the same P; Q read from q, a
// point in the layout of store_x28_raw -- global memory or LDS -- one coordinate per lane
// and step as the step needs it, so Q never occupies registers whole).  The four product
// steps and their broadcasts run unconditionally (straight-line code around the DPP
// moves; branching around them made the compiler pass the result through scratch
// memory); the identity and equal-x cases then choose the result.
__device__ __forceinline__ g1x28 g1x28_add_quad(const g1x28 p, const uint32_t* __restrict__ q) {
  uint32_t q_any = 0;
#pragma unroll
  for (int i = 0; i < f28c::N; i++) q_any |= q[28 + i];
  const bool q_inf = q_any == 0;                  // Q is the identity (ZZ literally zero)
  const bool p_inf = g1x28_is_inf(p);
  const int ql = (int)(threadIdx.x & 3u);
  const bool o1 = (ql & 1) != 0, o2 = (ql & 2) != 0;   // lane bits: operands by two picks
  // step 1: U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1
  const f28 qa = load_f28(q + (ql == 1 ? 0 : 14));      // X2 (lane 1) or Y2 (lane 3)
  const f28 qb = load_f28(q + (ql == 0 ? 28 : 42));     // ZZ2 (lane 0) or ZZZ2 (lane 2)
  f28 m = f28_mul(f28_pick(o1, qa, f28_pick(o2, p.Y, p.X)), f28_pick(o1, f28_pick(o2, p.ZZZ, p.ZZ), qb));
  const f28 U1 = f28_quad_bcast<0>(m), U2 = f28_quad_bcast<1>(m);
  const f28 S1 = f28_quad_bcast<2>(m), S2 = f28_quad_bcast<3>(m);
  const f28 P = f28_sub<4, 1>(U2, U1);            // < 6p
  const f28 R = f28_sub<4, 1>(S2, S1);            // < 6p
  // step 2: PP = P^2, R^2, ZZ1 ZZ2, ZZZ1 ZZZ2
  const f28 qc = load_f28(q + (ql == 2 ? 28 : 42));     // ZZ2 (lane 2) or ZZZ2 (lane 3)
  const f28 PR = f28_pick(o1, R, P);
  m = f28_mul(f28_pick(o2, f28_pick(o1, p.ZZZ, p.ZZ), PR), f28_pick(o2, qc, PR));
  const f28 PP = f28_quad_bcast<0>(m), RR = f28_quad_bcast<1>(m);
  const f28 ZZ12 = f28_quad_bcast<2>(m), ZZZ12 = f28_quad_bcast<3>(m);
  // step 3: PPP = P PP, Q = U1 PP, ZZ3 = ZZ1 ZZ2 PP (lane 3 repeats lane 2)
  m = f28_mul(f28_pick(o2, ZZ12, f28_pick(o1, U1, P)), PP);
  const f28 PPP = f28_quad_bcast<0>(m), Qv = f28_quad_bcast<1>(m);
  g1x28 r;
  r.ZZ = f28_quad_bcast<2>(m);
  r.X = f28_norm(f28_sub<8, 3>(RR, f28_add(PPP, f28_add(Qv, Qv))));   // < 10p
  const f28 T = f28_sub<16, 1>(Qv, r.X);                               // < 18p
  // step 4: ZZZ3 = ZZZ1 ZZZ2 PPP (lane 0, as ZZZ12 PPP + 0 0) and
  // Y3 = R T + S1 (4p - PPP) (lanes 1..3), one two-product reduction each
  const f28 z = f28_zero();
  const f28 nPPP = f28_sub<4, 1>(z, PPP);
  const bool l0 = ql == 0;
  m = f28_mul_sum(f28_pick(l0, ZZZ12, R), f28_pick(l0, PPP, T), f28_pick(l0, z, S1), f28_pick(l0, z, nPPP));
  r.ZZZ = f28_quad_bcast<0>(m);
  r.Y = f28_quad_bcast<1>(m);
  // An identity operand also gives P = R = 0; the selects below replace the result then,
  // so only two finite points take the doubling, which would otherwise run on the whole
  // wave whenever one quad adds an identity.
  const bool equal_x = !p_inf && !q_inf && f28_is_zero_mod_p<6>(P);
  if (__builtin_expect(equal_x, 0)) {   // no struct-valued ?: here: it would go through memory
    if (f28_is_zero_mod_p<6>(R)) r = g1x28_dbl(p);
    else r = g1x28_inf();
  }
  if (q_inf) r = p;
  if (p_inf) r = load_x28(q);
  return r;
}

__device__ __forceinline__ void store_x28_raw(uint32_t* __restrict__ p, const g1x28& a) {
  uint32_t w[MZK_X28_WORDS];
#pragma unroll
  for (int i = 0; i < 14; i++) {
    w[i] = a.X.l[i];
    w[14 + i] = a.Y.l[i];
    w[28 + i] = a.ZZ.l[i];
    w[42 + i] = a.ZZZ.l[i];
  }
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < MZK_X28_WORDS / 4; i++) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}
}  // namespace mzk
#endif
