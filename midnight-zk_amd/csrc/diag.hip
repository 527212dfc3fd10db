// Element-wise field / curve kernels for parity tests, and the micro-benchmarks
// that measure the chip's 32x32+64 multiply-add rate (the roof for MSM and NTT).
#include "ctx.hpp"
#include "g1.hpp"
#include "g1_28.hpp"

namespace mzk {

template <class P>
__global__ void k_field_op(int op, const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                           uint32_t* __restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int N = P::N;
  Field<P> x = load<P>(a + N * i);
  Field<P> y = load<P>(b + N * i);
  Field<P> r;
  switch (op) {
    case 0: r = mul(x, y); break;
    case 1: r = add(x, y); break;
    case 2: r = sub(x, y); break;
    case 3: r = sqr(x); break;
    case 4: r = from_mont(x); break;
    default: r = mul_cios(x, y); break;
  }
  store<P>(out + N * i, r);
}

__global__ void k_g1_op(int op, const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                        uint32_t* __restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  g1_aff_t p = load_affine(a + 24 * i);
  g1_aff_t q = load_affine(b + 24 * i);
  g1_xyzz_t r;
  if (op == 0) {
    r = g1_add_affine(g1_xyzz_t::from_affine(p), q);
  } else if (op == 1) {
    r = g1_add_affine(g1_xyzz_t::from_affine(p), p);
  } else {
    g1_xyzz_t p2 = g1_dbl(g1_xyzz_t::from_affine(p));
    g1_xyzz_t q2 = g1_dbl(g1_xyzz_t::from_affine(q));
    r = g1_add(p2, q2);
  }
  fp_t X, Y, Z;
  g1_to_jacobian(r, X, Y, Z);
  store<FpP>(out + 36 * i, X);
  store<FpP>(out + 36 * i + 12, Y);
  store<FpP>(out + 36 * i + 24, Z);
}

// 8 independent multiply-add chains per thread.
__global__ void k_bench_mad(int iters, const uint32_t* __restrict__ seed, uint64_t* __restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t x = seed[0] ^ t;
  uint32_t y = seed[1] + t;
  uint64_t acc[8];
#pragma unroll
  for (int k = 0; k < 8; k++) acc[k] = (uint64_t)seed[2 + k] + t;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = (uint64_t)((uint32_t)acc[k] ^ x) * y + (acc[k] >> 32);
  }
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) s ^= acc[k];
  if (s == 0x123456789abcdefull) sink[t] = s;
}

// 8 independent dependent chains acc = lo(acc) * y + acc: one v_mad_u64_u32 per
// step and nothing else, so the time is the multiply-add issue rate.
__global__ void k_bench_mad_pure(int iters, const uint32_t* __restrict__ seed, uint64_t* __restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t y = seed[1] | 1u;
  uint64_t acc[8];
#pragma unroll
  for (int k = 0; k < 8; k++) acc[k] = (uint64_t)seed[2 + k] + t;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = (uint64_t)(uint32_t)acc[k] * y + acc[k];
  }
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) s ^= acc[k];
  if (s == 0x123456789abcdefull) sink[t] = s;
}

template <class P, bool FIPS>
__global__ void k_bench_mul(int iters, const uint32_t* __restrict__ src, uint32_t* __restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  Field<P> x = load<P>(src);
  Field<P> y = load<P>(src + 16);
  x.v[0] ^= t;
  for (int it = 0; it < iters; it++) x = FIPS ? mul_fips(x, y) : mul_cios(x, y);
  if (x.v[0] == 0x12345678u && x.v[1] == t) store<P>(sink, x);
}

// radix-2^28 products, many independent chains
__global__ void k_bench_mul28(int iters, const uint32_t* __restrict__ src, uint32_t* __restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  f28 x = f28_from_std(src), y = f28_from_std(src + 12);
  x.l[0] ^= t & 0xff;
  for (int it = 0; it < iters; it++) x = f28_mul(x, y);
  if (x.l[0] == 0x1234567u && x.l[1] == t) sink[t] = x.l[2];
}

// mixed additions into one accumulator per thread: the 32-bit form (g1.hpp) and the
// radix-2^28 form (g1_28.hpp); the added points vary per iteration
__global__ void k_bench_madd(int iters, int form28, const uint32_t* __restrict__ src, uint32_t* __restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t xs[12], ys[12];
  for (int i = 0; i < 12; i++) {
    xs[i] = src[i];
    ys[i] = src[12 + i];
  }
  xs[0] ^= t & 0xff;
  if (form28) {
    g1x28 acc = g1x28_from_affine(f28_from_std(src + 4), f28_from_std(src + 12), false);
    for (int it = 0; it < iters; it++) {
      xs[1] += it;
      acc = g1x28_madd(acc, f28_from_std(xs), f28_from_std(ys), (it & 1) != 0);
    }
    if (acc.X.l[0] == 0x1234567u && acc.X.l[1] == t) sink[t] = acc.Y.l[0];
  } else {
    g1_aff_t q0;
    q0.x = load<FpP>(src + 4);
    q0.y = load<FpP>(src + 12);
    g1_xyzz_t acc = g1_xyzz_t::from_affine(q0);
    for (int it = 0; it < iters; it++) {
      xs[1] += it;
      g1_aff_t q;
      for (int i = 0; i < 12; i++) {
        q.x.v[i] = xs[i];
        q.y.v[i] = ys[i];
      }
      if (it & 1) q = g1_neg(q);
      acc = g1_add_affine(acc, q);
    }
    if (acc.X.v[0] == 0x12345678u && acc.X.v[1] == t) sink[t] = acc.Y.v[0];
  }
}

// dependent 32-bit multiply-lows (v_mul_lo_u32 issue rate)
__global__ void k_bench_mullo(int iters, const uint32_t* __restrict__ src, uint32_t* __restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[8];
  for (int k = 0; k < 8; k++) a[k] = src[k] ^ t;
  const uint32_t m = src[9] | 1u;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = a[k] * m;
  }
  uint32_t s = 0;
  for (int k = 0; k < 8; k++) s ^= a[k];
  if (s == 0x12345678u) sink[t] = s;
}

}  // namespace mzk

using namespace mzk;

extern "C" {

int mzk_diag_field_op(mzk_ctx* c, int field, int op, const void* a, const void* b, void* out, size_t n) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_REQUIRE(field == 0 || field == 1, "field must be 0 (fr) or 1 (fp)");
  MZK_REQUIRE(op >= 0 && op <= 5, "op out of range");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a && b && out, "null pointer");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  if (field == 0)
    hipLaunchKernelGGL(k_field_op<FrP>, dim3(grid_for(n, 128)), dim3(128), 0, c->stream, op,
                       (const uint32_t*)a, (const uint32_t*)b, (uint32_t*)out, n);
  else
    hipLaunchKernelGGL(k_field_op<FpP>, dim3(grid_for(n, 128)), dim3(128), 0, c->stream, op,
                       (const uint32_t*)a, (const uint32_t*)b, (uint32_t*)out, n);
  MZK_HIP(hipGetLastError());
  MZK_HIP(hipStreamSynchronize(c->stream));
  return MZK_OK;
}

int mzk_diag_g1_op(mzk_ctx* c, int op, const void* a, const void* b, void* out, size_t n) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_REQUIRE(op >= 0 && op <= 2, "op out of range");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(a && b && out, "null pointer");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_g1_op, dim3(grid_for(n, 64)), dim3(64), 0, c->stream, op, (const uint32_t*)a,
                     (const uint32_t*)b, (uint32_t*)out, n);
  MZK_HIP(hipGetLastError());
  MZK_HIP(hipStreamSynchronize(c->stream));
  return MZK_OK;
}

int mzk_diag_bench(mzk_ctx* c, int which, int blocks, int threads, int iters, double* ms, double* ops) {
  MZK_REQUIRE(c != nullptr && ms && ops, "null argument");
  MZK_REQUIRE(which >= 0 && which <= 10, "which out of range");
  MZK_REQUIRE(blocks > 0 && threads > 0 && threads <= 1024 && iters > 0, "bad launch shape");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  size_t sink_bytes = (size_t)blocks * threads * 8;
  MZK_TRY(c->scratch[SCR_IN0].reserve(sink_bytes + 4096));
  uint32_t* buf = (uint32_t*)c->scratch[SCR_IN0].ptr;
  // src: two fp-sized operands of small fixed values (any values exercise the same code)
  uint32_t host[32];
  for (int i = 0; i < 32; i++) host[i] = 0x01020304u * (i + 1) & 0x0fffffffu;
  MZK_HIP(hipMemcpyAsync(buf, host, sizeof host, hipMemcpyHostToDevice, c->stream));
  uint32_t* sink = buf + 64;
  hipEvent_t e0, e1;
  MZK_HIP(hipEventCreate(&e0));
  MZK_HIP(hipEventCreate(&e1));
  MZK_HIP(hipEventRecord(e0, c->stream));
  double per_thread = 0;
  switch (which) {
    case 0:
      hipLaunchKernelGGL(k_bench_mad, dim3(blocks), dim3(threads), 0, c->stream, iters, buf, (uint64_t*)sink);
      per_thread = 8.0 * iters;
      break;
    case 1:
    case 2:
      hipLaunchKernelGGL((k_bench_mul<FpP, false>), dim3(blocks), dim3(threads), 0, c->stream, iters, buf, sink);
      per_thread = iters;
      break;
    case 5:
      hipLaunchKernelGGL((k_bench_mul<FpP, true>), dim3(blocks), dim3(threads), 0, c->stream, iters, buf, sink);
      per_thread = iters;
      break;
    case 6:
      hipLaunchKernelGGL((k_bench_mul<FrP, true>), dim3(blocks), dim3(threads), 0, c->stream, iters, buf, sink);
      per_thread = iters;
      break;
    case 7:
      hipLaunchKernelGGL(k_bench_mul28, dim3(blocks), dim3(threads), 0, c->stream, iters, buf, sink);
      per_thread = iters;
      break;
    case 8:
    case 9:
      hipLaunchKernelGGL(k_bench_madd, dim3(blocks), dim3(threads), 0, c->stream, iters, which == 9 ? 1 : 0, buf,
                         sink);
      per_thread = iters;
      break;
    case 10:
      hipLaunchKernelGGL(k_bench_mullo, dim3(blocks), dim3(threads), 0, c->stream, iters, buf, sink);
      per_thread = 8.0 * iters;
      break;
    case 4:
      hipLaunchKernelGGL(k_bench_mad_pure, dim3(blocks), dim3(threads), 0, c->stream, iters, buf, (uint64_t*)sink);
      per_thread = 8.0 * iters;
      break;
    default:
      hipLaunchKernelGGL((k_bench_mul<FrP, false>), dim3(blocks), dim3(threads), 0, c->stream, iters, buf, sink);
      per_thread = iters;
      break;
  }
  MZK_HIP(hipGetLastError());
  MZK_HIP(hipEventRecord(e1, c->stream));
  MZK_HIP(hipEventSynchronize(e1));
  float t = 0;
  MZK_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t;
  *ops = per_thread * (double)blocks * threads;
  return MZK_OK;
}

}  // extern "C"
