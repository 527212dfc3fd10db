// The SRS file format of ParamsKZG::write_custom / read_custom
// (proofs/src/poly/kzg/params.rs:259-349), decoded and encoded on the device.
//
// A file is k (u32 LE), then 2^k points g, 2^k points g_lagrange, then g2 and s_g2.
// G1 points are in the ZCash serialisation (the encoding of blst and of
// curves/src/bls12_381/g1.rs:337-390, 448-489):
//   SerdeFormat::Processed          48 B compressed: big-endian x with flag bits in its
//                                   top byte (0x80 compressed, 0x40 identity, 0x20 y is
//                                   the larger of y and p - y); decoded with a square
//                                   root and checked to lie in G1 (from_compressed);
//   SerdeFormat::RawBytes           96 B uncompressed: big-endian x || y, flags 0x40 for
//                                   the identity; checked on the curve and in G1
//                                   (from_uncompressed);
//   SerdeFormat::RawBytesUnchecked  the same bytes, checked on the curve only
//                                   (from_uncompressed_unchecked).
// The G2 points stay with the caller: the prover's MSMs never read them.
//
// One thread decodes one point: byte swap into 12 x 32-bit limbs, flag checks, x < p
// (and y < p), Montgomery form by one product with R^2, the curve equation
// y^2 = x^3 + 4, for compressed input y = (x^3 + 4)^((p+1)/4) (p = 3 mod 4) with the
// sign chosen by the flag, and, where the format asks for it, [r] P = O in the
// radix-2^28 XYZZ form (r prime and the cofactor prime to r, so this holds exactly for
// the points of G1).  The first failing point (lowest index) is reported.
#include <string.h>

#include <vector>

#include "ctx.hpp"
#include "field.hpp"
#include "g1_28.hpp"
#include "host_ec.hpp"
#include "srs.hpp"

namespace mzk {

// per-launch constants, computed on the host
struct SerdeArgs {
  uint32_t r2[12];       // R^2 mod p (Montgomery form of R)
  uint32_t sqrt_e[12];   // (p + 1) / 4
  uint32_t half[12];     // (p - 1) / 2, canonical
};

enum : uint32_t { PT_OK = 0, PT_BAD_ENCODING = 1, PT_NOT_ON_CURVE = 2, PT_NOT_IN_GROUP = 3 };

__device__ __forceinline__ bool lt_p(const uint32_t a[12]) {
  for (int i = 11; i >= 0; i--) {
    if (a[i] != FpP::m(i)) return a[i] < FpP::m(i);
  }
  return false;
}

__device__ __forceinline__ bool gt_words(const uint32_t a[12], const uint32_t b[12]) {
  for (int i = 11; i >= 0; i--) {
    if (a[i] != b[i]) return a[i] > b[i];
  }
  return false;
}

// 48 big-endian bytes at p (16-byte aligned) -> 12 little-endian 32-bit limbs
__device__ __forceinline__ void load_be48(const uint8_t* __restrict__ p, uint32_t w[12]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint32_t raw[12];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const uint4 v = q[i];
    raw[4 * i] = v.x;
    raw[4 * i + 1] = v.y;
    raw[4 * i + 2] = v.z;
    raw[4 * i + 3] = v.w;
  }
#pragma unroll
  for (int i = 0; i < 12; i++) w[i] = __builtin_bswap32(raw[11 - i]);
}

__device__ __forceinline__ void store_be48(uint8_t* __restrict__ p, const uint32_t w[12]) {
  uint32_t raw[12];
#pragma unroll
  for (int i = 0; i < 12; i++) raw[11 - i] = __builtin_bswap32(w[i]);
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < 3; i++) q[i] = make_uint4(raw[4 * i], raw[4 * i + 1], raw[4 * i + 2], raw[4 * i + 3]);
}

__device__ __forceinline__ fp_t fp_of_words(const uint32_t w[12]) {
  fp_t r;
#pragma unroll
  for (int i = 0; i < 12; i++) r.v[i] = w[i];
  return r;
}

__device__ fp_t fp_pow_words(const fp_t& a, const uint32_t e[12]) {
  fp_t acc = fp_t::one();
  int top = 383;
  while (top >= 0 && !((e[top >> 5] >> (top & 31)) & 1u)) top--;
  for (int b = top; b >= 0; b--) {
    acc = sqr(acc);
    if ((e[b >> 5] >> (b & 31)) & 1u) acc = mul(acc, a);
  }
  return acc;
}

// [r] P == O for an affine point (Montgomery 12-limb coordinates, not the identity)
__device__ bool in_g1(const uint32_t xm[12], const uint32_t ym[12]) {
  const f28 x = f28_from_std(xm), y = f28_from_std(ym);
  g1x28 acc = g1x28_inf();
  for (int b = 254; b >= 0; b--) {   // r < 2^255, bit 254 set
    if (!g1x28_is_inf(acc)) acc = g1x28_dbl(acc);
    if ((FrP::m(b >> 5) >> (b & 31)) & 1u) acc = g1x28_madd(acc, x, y, false);
  }
  return g1x28_is_inf(acc);
}

// Decode count points (rec bytes each: 48 compressed, 96 uncompressed) into affine
// Montgomery points; *first_bad = min over failing points of (index << 8 | reason).
__global__ void __launch_bounds__(64) k_g1_decode(const uint8_t* __restrict__ in, size_t count, int compressed,
                                                  int subgroup, SerdeArgs A, size_t index_base,
                                                  uint32_t* __restrict__ out,
                                                  unsigned long long* __restrict__ first_bad) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const size_t rec = compressed ? 48 : 96;
  uint32_t xw[12], yw[12];
  load_be48(in + rec * i, xw);
  if (!compressed) load_be48(in + rec * i + 48, yw);
  const uint32_t flags = xw[11] >> 29;   // C, I, S from the top bits
  xw[11] &= 0x1fffffffu;
  uint32_t reason = PT_OK;
  uint32_t* o = out + 24 * i;
  if (((flags >> 2) & 1u) != (compressed ? 1u : 0u)) reason = PT_BAD_ENCODING;
  if (reason == PT_OK && (flags & 2u)) {   // the identity: every other bit zero
    uint32_t any = flags & 1u;
    for (int j = 0; j < 12; j++) any |= xw[j];
    if (!compressed)
      for (int j = 0; j < 12; j++) any |= yw[j];
    if (any) {
      reason = PT_BAD_ENCODING;
    } else {
      for (int j = 0; j < 24; j++) o[j] = 0;
      return;
    }
  }
  if (reason == PT_OK && !compressed && (flags & 1u)) reason = PT_BAD_ENCODING;
  if (reason == PT_OK && (!lt_p(xw) || (!compressed && !lt_p(yw)))) reason = PT_BAD_ENCODING;
  fp_t X, Y;
  if (reason == PT_OK) {
    const fp_t r2 = fp_of_words(A.r2);
    X = mul(fp_of_words(xw), r2);
    fp_t four = fp_t::one();
    four = add(four, four);
    four = add(four, four);
    const fp_t rhs = add(mul(sqr(X), X), four);
    if (compressed) {
      Y = fp_pow_words(rhs, A.sqrt_e);
      if (!(sqr(Y) == rhs)) {
        reason = PT_NOT_ON_CURVE;   // x^3 + 4 has no square root
      } else {
        const fp_t yc = from_mont(Y);
        const bool larger = gt_words(yc.v, A.half);
        if (larger != (bool)(flags & 1u)) Y = neg(Y);
      }
    } else {
      Y = mul(fp_of_words(yw), r2);
      if (!(sqr(Y) == rhs)) reason = PT_NOT_ON_CURVE;
    }
  }
  if (reason == PT_OK && subgroup && !in_g1(X.v, Y.v)) reason = PT_NOT_IN_GROUP;
  if (reason != PT_OK) {
    for (int j = 0; j < 24; j++) o[j] = 0;
    atomicMin(first_bad, ((unsigned long long)(index_base + i) << 8) | reason);
    return;
  }
  store<FpP>(o, X);
  store<FpP>(o + 12, Y);
}

// Affine Montgomery points -> ZCash bytes (48 compressed / 96 uncompressed each)
__global__ void __launch_bounds__(64) k_g1_encode(const uint32_t* __restrict__ aff, size_t count, int compressed,
                                                  SerdeArgs A, uint8_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const size_t rec = compressed ? 48 : 96;
  const fp_t x = load<FpP>(aff + 24 * i), y = load<FpP>(aff + 24 * i + 12);
  uint32_t xw[12], yw[12];
  if (x.is_zero() && y.is_zero()) {
    for (int j = 0; j < 12; j++) xw[j] = yw[j] = 0;
    xw[11] = compressed ? 0xc0000000u : 0x40000000u;
  } else {
    const fp_t xc = from_mont(x), yc = from_mont(y);
    for (int j = 0; j < 12; j++) {
      xw[j] = xc.v[j];
      yw[j] = yc.v[j];
    }
    if (compressed) xw[11] |= 0x80000000u | (gt_words(yc.v, A.half) ? 0x20000000u : 0u);
  }
  store_be48(out + rec * i, xw);
  if (!compressed) store_be48(out + rec * i + 48, yw);
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

static SerdeArgs serde_args() {
  SerdeArgs A;
  // R^2 mod p: R mod p doubled 384 times
  host::Fp r2 = host::fp_one();
  for (int i = 0; i < 384; i++) r2 = host::fp_add(r2, r2);
  uint64_t e[6], h[6];
  // (p + 1) / 4 and (p - 1) / 2 (p odd, p + 1 has no carry out of the top limb)
  uint64_t carry = 1;
  for (int i = 0; i < 6; i++) {
    const uint64_t s = host::kP[i] + carry;
    carry = s < carry;
    e[i] = s;
  }
  for (int i = 0; i < 6; i++) {
    const uint64_t nxt = i < 5 ? e[i + 1] : 0;
    e[i] = (e[i] >> 2) | (nxt << 62);
    const uint64_t pm1 = i == 0 ? host::kP[0] - 1 : host::kP[i];
    const uint64_t pn = i < 5 ? host::kP[i + 1] : 0;
    h[i] = (pm1 >> 1) | (pn << 63);
  }
  for (int i = 0; i < 6; i++) {
    A.r2[2 * i] = (uint32_t)r2.l[i];
    A.r2[2 * i + 1] = (uint32_t)(r2.l[i] >> 32);
    A.sqrt_e[2 * i] = (uint32_t)e[i];
    A.sqrt_e[2 * i + 1] = (uint32_t)(e[i] >> 32);
    A.half[2 * i] = (uint32_t)h[i];
    A.half[2 * i + 1] = (uint32_t)(h[i] >> 32);
  }
  return A;
}

static const char* reason_text(uint32_t r) {
  switch (r) {
    case PT_BAD_ENCODING: return "invalid encoding (flags, or a coordinate not below p)";
    case PT_NOT_ON_CURVE: return "not on the curve";
    case PT_NOT_IN_GROUP: return "not in the prime-order subgroup";
    default: return "unknown";
  }
}

static constexpr size_t SERDE_CHUNK = 1u << 18;   // points per staged chunk

}  // namespace mzk

using namespace mzk;

extern "C" {

int mzk_srs_read_points(mzk_ctx* c, const uint8_t* bytes, size_t count, int format, mzk_srs** out) {
  MZK_REQUIRE(c != nullptr && out != nullptr && bytes != nullptr, "mzk_srs_read_points: null argument");
  MZK_REQUIRE(count >= 1 && count < (1ull << 31), "mzk_srs_read_points: count out of range");
  MZK_REQUIRE(format >= MZK_SERDE_PROCESSED && format <= MZK_SERDE_RAW_BYTES_UNCHECKED,
              "mzk_srs_read_points: unknown format %d", format);
  *out = nullptr;
  const int compressed = format == MZK_SERDE_PROCESSED;
  const int subgroup = format != MZK_SERDE_RAW_BYTES_UNCHECKED;
  const size_t rec = compressed ? 48 : 96;
  mzk_srs* s = nullptr;
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipSetDevice(c->device));
    MZK_TRY(srs_alloc_bases(c, count, &s));
    int rc = c->scratch[SCR_IN0].reserve(SERDE_CHUNK * rec);
    if (rc == MZK_OK) rc = c->scratch[SCR_MSM_D].reserve(256);
    if (rc != MZK_OK) {
      mzk_srs_destroy(s);
      return rc;
    }
    const SerdeArgs A = serde_args();
    unsigned long long* bad = (unsigned long long*)c->scratch[SCR_MSM_D].ptr;
    hipError_t e = hipSuccess;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, c->stream, bad, ~0ull);
    for (size_t off = 0; off < count && e == hipSuccess; off += SERDE_CHUNK) {
      const size_t m = count - off < SERDE_CHUNK ? count - off : SERDE_CHUNK;
      uint8_t* stage = (uint8_t*)c->scratch[SCR_IN0].ptr;
      e = hipMemcpyAsync(stage, bytes + rec * off, rec * m, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) break;
      MZK_LAUNCH(c, "k_g1_decode", k_g1_decode, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, stage, m, compressed,
                 subgroup, A, off, (uint32_t*)s->bases + 24 * off, bad);
      e = hipGetLastError();
    }
    unsigned long long first = ~0ull;
    if (e == hipSuccess) e = hipMemcpyAsync(&first, bad, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      mzk_srs_destroy(s);
      set_error("mzk_srs_read_points: %s", hipGetErrorString(e));
      return MZK_ERR_HIP;
    }
    if (first != ~0ull) {
      mzk_srs_destroy(s);
      set_error("mzk_srs_read_points: point %llu: %s", first >> 8, reason_text((uint32_t)(first & 0xff)));
      return MZK_ERR_INVALID;
    }
    rc = srs_finish_bases(c, s);
    if (rc != MZK_OK) {
      mzk_srs_destroy(s);
      return rc;
    }
  }
  *out = s;
  return MZK_OK;
}

int mzk_srs_write_points(mzk_ctx* c, const mzk_srs* srs, size_t offset, size_t count, int format, uint8_t* bytes) {
  MZK_REQUIRE(c != nullptr && srs != nullptr && (bytes != nullptr || count == 0),
              "mzk_srs_write_points: null argument");
  MZK_REQUIRE(offset <= srs->n && count <= srs->n - offset, "mzk_srs_write_points: range beyond the bases");
  MZK_REQUIRE(format >= MZK_SERDE_PROCESSED && format <= MZK_SERDE_RAW_BYTES_UNCHECKED,
              "mzk_srs_write_points: unknown format %d", format);
  if (count == 0) return MZK_OK;
  const int compressed = format == MZK_SERDE_PROCESSED;
  const size_t rec = compressed ? 48 : 96;
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  MZK_TRY(c->scratch[SCR_IN0].reserve(SERDE_CHUNK * rec));
  const SerdeArgs A = serde_args();
  for (size_t off = 0; off < count; off += SERDE_CHUNK) {
    const size_t m = count - off < SERDE_CHUNK ? count - off : SERDE_CHUNK;
    uint8_t* stage = (uint8_t*)c->scratch[SCR_IN0].ptr;
    MZK_LAUNCH(c, "k_g1_encode", k_g1_encode, dim3((unsigned)((m + 63) / 64)), dim3(64), 0,
               (const uint32_t*)srs->bases + 24 * (offset + off), m, compressed, A, stage);
    MZK_HIP(hipGetLastError());
    MZK_HIP(hipMemcpyAsync(bytes + rec * off, stage, rec * m, hipMemcpyDeviceToHost, c->stream));
    MZK_HIP(hipStreamSynchronize(c->stream));
  }
  return MZK_OK;
}

}  // extern "C"
