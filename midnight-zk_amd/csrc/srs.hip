// Resident SRS bases (proofs/src/poly/kzg/params.rs:22-37) and their device-side
// generation g_i = [tau^i] G1 (ParamsKZG::unsafe_setup, params.rs:154-172).
//
// Generation: tau^i by a 2^j power table, a fixed-base comb over 8-bit digits of
// the canonical scalar (32 mixed additions per point against a 32 x 256 table of
// d * 2^(8w) G built once on the host), then batch normalisation to affine with
// one Fermat inversion per 64 points (Montgomery's trick).
#include <vector>

#include "ctx.hpp"
#include "fp28.hpp"
#include "g1.hpp"
#include "g1_28.hpp"
#include "host_ec.hpp"
#include "msm.hpp"
#include "srs.hpp"
#include "tune.hpp"

namespace mzk {

struct FrPows {
  uint32_t v[32][8];
};

// out[i] = tau^(start + i) as Montgomery, from pows[j] = tau^(2^j)
__global__ void k_tau_powers(uint32_t* __restrict__ out, size_t n, FrPows p) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t r = fr_t::one();
  for (int j = 0; j < 32; j++) {
    if ((i >> j) & 1) {
      fr_t q;
#pragma unroll
      for (int t = 0; t < 8; t++) q.v[t] = p.v[j][t];
      r = mul(r, q);
    }
  }
  store<FrP>(out + 8 * i, r);
}

// XYZZ [s_i] G by the comb table (32 windows of 8 bits).
__global__ void __launch_bounds__(64) k_fixed_base(const uint32_t* __restrict__ scalars, size_t n,
                                                   const uint32_t* __restrict__ comb, uint32_t* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t s = from_mont(load<FrP>(scalars + 8 * i));
  g1_xyzz_t acc = g1_xyzz_t::inf();
  for (int w = 0; w < 32; w++) {
    uint32_t d = (s.v[w >> 2] >> (8 * (w & 3))) & 0xffu;
    if (d) acc = g1_add_affine(acc, load_affine(comb + 24 * (size_t)(w * 256 + d)));
  }
  store_xyzz(out + 48 * i, acc);
}

// a^(p-2) on the device (one dependent chain per chunk of points): square and multiply
// from the top bit of the constant exponent.  The exponent bits are uniform across the
// wave (scalar branches) and no table of powers is kept, so nothing spills to scratch (a
// 16-entry table of powers did, 784 bytes per lane).
__device__ fp_t fp_inv_dev(const fp_t& a) {
  constexpr uint32_t E[12] = {0xffffaaa9u, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu, 0xf6b0f624u, 0x6730d2a0u,
                              0xf38512bfu, 0x64774b84u, 0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
  fp_t r = a;   // bit 380, the top bit of p - 2
  for (int i = 379; i >= 0; i--) {
    r = sqr(r);
    if ((E[i >> 5] >> (i & 31)) & 1u) r = mul(r, a);
  }
  return r;
}

// XYZZ -> affine for n points, `chunk` per thread: with u = ZZ * ZZZ,
// 1/ZZ = ZZZ / u and 1/ZZZ = ZZ / u; one inversion (~570 dependent products) per chunk.
__global__ void __launch_bounds__(64) k_batch_normalize(const uint32_t* __restrict__ in, size_t n, uint32_t chunk,
                                                        uint32_t* __restrict__ out, uint32_t* __restrict__ prefix) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t i0 = t * chunk;
  if (i0 >= n) return;
  size_t i1 = i0 + chunk < n ? i0 + chunk : n;
  fp_t acc = fp_t::one();
  for (size_t i = i0; i < i1; i++) {
    store<FpP>(prefix + 12 * i, acc);
    const uint32_t* p = in + 48 * i;
    fp_t zz = load<FpP>(p + 24);
    if (!zz.is_zero()) acc = mul(acc, mul(zz, load<FpP>(p + 36)));
  }
  fp_t inv = fp_inv_dev(acc);
  for (size_t i = i1; i-- > i0;) {
    const uint32_t* p = in + 48 * i;
    fp_t zz = load<FpP>(p + 24);
    uint32_t* o = out + 24 * i;
    if (zz.is_zero()) {
      store<FpP>(o, fp_t::zero());
      store<FpP>(o + 12, fp_t::zero());
      continue;
    }
    fp_t zzz = load<FpP>(p + 36);
    fp_t uinv = mul(inv, load<FpP>(prefix + 12 * i));
    inv = mul(inv, mul(zz, zzz));
    store<FpP>(o, mul(load<FpP>(p), mul(zzz, uinv)));
    store<FpP>(o + 12, mul(load<FpP>(p + 12), mul(zz, uinv)));
  }
}

// ------------------------------------------------------------------- host tables
static const uint64_t kG1x[6] = {0x5cb38790fd530c16ull, 0x7817fc679976fff5ull, 0x154f95c7143ba1c1ull,
                                 0xf0ae6acdf3d0e747ull, 0xedce6ecc21dbf440ull, 0x120177419e0bfb75ull};
static const uint64_t kG1y[6] = {0xbaac93d50ce72271ull, 0x8c22631a7918fd8eull, 0xdd595f13570725ceull,
                                 0x51ac582950405194ull, 0x0e1c8c3fad0059c0ull, 0x0bbc3efc5008a26aull};

// comb[w][d] = d * 2^(8w) G (affine, blst layout), d in [0, 256), identity at d = 0.
static std::vector<uint64_t> comb_table() {
  using namespace host;
  std::vector<Xyzz> pts(32 * 256);
  Affine g;
  memcpy(g.x.l, kG1x, 48);
  memcpy(g.y.l, kG1y, 48);
  Xyzz base = xyzz_from_affine(g);
  for (int w = 0; w < 32; w++) {
    Xyzz acc = Xyzz::inf();
    pts[w * 256] = acc;
    for (int d = 1; d < 256; d++) {
      acc = xyzz_add(acc, base);
      pts[w * 256 + d] = acc;
    }
    for (int k = 0; k < 8; k++) base = xyzz_dbl(base);
  }
  // batch-normalise: u = ZZ * ZZZ
  std::vector<Fp> pre(pts.size());
  Fp acc = fp_one();
  for (size_t i = 0; i < pts.size(); i++) {
    pre[i] = acc;
    if (!pts[i].is_inf()) acc = fp_mul(acc, fp_mul(pts[i].ZZ, pts[i].ZZZ));
  }
  Fp inv = fp_inv(acc);
  std::vector<uint64_t> out(pts.size() * 12, 0);
  for (size_t i = pts.size(); i-- > 0;) {
    if (pts[i].is_inf()) continue;
    Fp uinv = fp_mul(inv, pre[i]);
    inv = fp_mul(inv, fp_mul(pts[i].ZZ, pts[i].ZZZ));
    Fp x = fp_mul(pts[i].X, fp_mul(pts[i].ZZZ, uinv));
    Fp y = fp_mul(pts[i].Y, fp_mul(pts[i].ZZ, uinv));
    memcpy(&out[12 * i], x.l, 48);
    memcpy(&out[12 * i + 6], y.l, 48);
  }
  return out;
}

int batch_normalize(mzk_ctx* c, const uint32_t* d_xyzz, size_t n, uint32_t* d_affine, uint32_t* prefix) {
  if (!prefix) {
    MZK_TRY(c->scratch[SCR_MSM_D].reserve(n * 48 + 64));
    prefix = (uint32_t*)c->scratch[SCR_MSM_D].ptr;
  }
  // chunks of 16 while that still leaves fewer than ~2^18 threads (a 2^20 normalisation on
  // chunks of 64 kept three quarters of the SIMDs idle), 64 for the largest batches (the
  // shifted windows of an SRS: one inversion per 64 points)
  const uint32_t chunk = n >= ((size_t)1 << 22) ? 64 : 16;
  size_t threads = (n + chunk - 1) / chunk;
  MZK_LAUNCH(c, "k_batch_normalize", k_batch_normalize, dim3(grid_for(threads, 64)), dim3(64), 0, d_xyzz, n, chunk,
             d_affine, prefix);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

// out[i] = 2^c * in[i] (affine in, blst-layout XYZZ out): one doubling from affine, c - 1
// more, in the radix-2^28 form of the accumulation (fp28.hpp: 73 against 54 G products/s
// for the 32-bit form).
__global__ void __launch_bounds__(64) k_shift_window(const uint32_t* __restrict__ in, size_t n, int c,
                                                     uint32_t* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4* q = reinterpret_cast<const uint4*>(in + 24 * i);
  uint32_t w[24];
  uint32_t any = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const uint4 v = q[k];
    w[4 * k] = v.x; w[4 * k + 1] = v.y; w[4 * k + 2] = v.z; w[4 * k + 3] = v.w;
    any |= v.x | v.y | v.z | v.w;
  }
  g1x28 p = g1x28_inf();
  if (any != 0) {   // the affine identity doubles to the identity
    const g1x28 a = g1x28_from_affine(f28_from_std(w), f28_from_std(w + 12), false);
    p = g1x28_dbl_affine(a.X, a.Y);
    for (int k = 1; k < c; k++) p = g1x28_dbl(p);
  }
  uint32_t o[48];
  g1x28_to_std(p, o);
  uint4* qo = reinterpret_cast<uint4*>(out + 48 * i);
#pragma unroll
  for (int k = 0; k < 12; k++) qo[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
}

// out[i] = 2^c * in[i], blst-layout XYZZ in and out (the next window from the previous
// one, with no normalisation in between); coordinates enter the radix-2^28 form below 2p.
__global__ void __launch_bounds__(64) k_shift_window_xyzz(const uint32_t* __restrict__ in, size_t n, int c,
                                                          uint32_t* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4* q = reinterpret_cast<const uint4*>(in + 48 * i);
  uint32_t w[48];
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const uint4 v = q[k];
    w[4 * k] = v.x; w[4 * k + 1] = v.y; w[4 * k + 2] = v.z; w[4 * k + 3] = v.w;
  }
  uint32_t zz_any = 0;
#pragma unroll
  for (int k = 24; k < 36; k++) zz_any |= w[k];
  g1x28 p = g1x28_inf();
  if (zz_any != 0) {
    const f28 one = f28_one();
    p.X = f28_mul(f28_from_std(w), one);
    p.Y = f28_mul(f28_from_std(w + 12), one);
    p.ZZ = f28_mul(f28_from_std(w + 24), one);
    p.ZZZ = f28_mul(f28_from_std(w + 36), one);
    for (int k = 0; k < c; k++) p = g1x28_dbl(p);
  }
  uint32_t o[48];
  g1x28_to_std(p, o);
  uint4* qo = reinterpret_cast<uint4*>(out + 48 * i);
#pragma unroll
  for (int k = 0; k < 12; k++) qo[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
}

// Affine points (blst layout) -> 128-byte records for the bucket accumulation: x and y
// in the radix-2^28 form, below 2p and normalised (words 0..13, 14..27), word 28 = 1
// for the identity, words 29..31 zero.  One record is one 128-byte line.
__global__ void __launch_bounds__(64) k_to_records(const uint32_t* __restrict__ aff, size_t m,
                                                   uint32_t* __restrict__ rec) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t* a = aff + 24 * i;
  uint32_t any = 0;
  for (int k = 0; k < 24; k++) any |= a[k];
  uint32_t w[MZK_SRS_REC_WORDS];
#pragma unroll
  for (int k = 0; k < MZK_SRS_REC_WORDS; k++) w[k] = 0;
  if (any) {
    // below 2p and normalised (a product by one each): the accumulation opens a bucket
    // with these coordinates as they are
    const f28 one = f28_one();
    const f28 x = f28_mul(f28_from_std(a), one), y = f28_mul(f28_from_std(a + 12), one);
#pragma unroll
    for (int k = 0; k < 14; k++) {
      w[k] = x.l[k];
      w[14 + k] = y.l[k];
    }
  } else {
    w[28] = 1;
  }
  uint4* q = reinterpret_cast<uint4*>(rec + MZK_SRS_REC_WORDS * i);
#pragma unroll
  for (int k = 0; k < MZK_SRS_REC_WORDS / 4; k++) q[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

// Window-shifted copies tab[w n + i] = 2^(c w) g_i for w < W, so that all windows of a
// commitment can share one bucket set.  Built window by window in the blst layout (n
// XYZZ temporaries), then stored as accumulation records.
int srs_build_shifted(mzk_ctx* c, mzk_srs* s, int cbits) {
  const int W = 255 / cbits + 1;
  const size_t n = s->n;
  const size_t bytes = (size_t)W * n * 96;
  void* std_tab = nullptr;
  hipError_t e = hipMalloc(&std_tab, bytes);
  if (e == hipSuccess) e = hipMalloc(&s->shifted, (size_t)W * n * MZK_SRS_REC_WORDS * 4);
  if (e != hipSuccess) {
    if (std_tab) (void)hipFree(std_tab);
    s->shifted = nullptr;
    set_error("srs: hipMalloc for shifted bases failed: %s", hipGetErrorString(e));
    return MZK_ERR_NOMEM;
  }
  // windows 1..W-1 as XYZZ, each doubled c times from the one before with no normalisation
  // in between, then normalised a group of windows per call (one inversion per 64 points
  // once a group holds 2^22 points); temporaries of at most ~4 GiB (XYZZ + prefixes)
  const size_t group = std::max<size_t>(1, std::min<size_t>(W - 1, ((size_t)4 << 30) / (n * 240)));
  void* xyz_mem = nullptr;
  if (W > 1) {
    e = hipMalloc(&xyz_mem, group * n * 240);
    if (e != hipSuccess) {
      (void)hipFree(std_tab);
      (void)hipFree(s->shifted);
      s->shifted = nullptr;
      set_error("srs: hipMalloc for shifted-window temporaries failed: %s", hipGetErrorString(e));
      return MZK_ERR_NOMEM;
    }
  }
  int rc = MZK_OK;
  do {
    uint32_t* xyz = (uint32_t*)xyz_mem;
    uint32_t* prefix = xyz + 48 * group * n;
    uint32_t* tab = (uint32_t*)std_tab;
    if ((e = hipMemcpyAsync(tab, s->bases, n * 96, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) break;
    for (int w0 = 1; w0 < W && rc == MZK_OK; w0 += (int)group) {
      const int g = std::min<int>((int)group, W - w0);
      for (int k = 0; k < g; k++) {
        uint32_t* dst = xyz + 48 * (size_t)k * n;
        if (w0 == 1 && k == 0)
          MZK_LAUNCH(c, "k_shift_window", k_shift_window, dim3(grid_for(n, 64)), dim3(64), 0, tab, n, cbits, dst);
        else   // from the previous window: slot k - 1, or the last slot of the previous group
          MZK_LAUNCH(c, "k_shift_window", k_shift_window_xyzz, dim3(grid_for(n, 64)), dim3(64), 0,
                     xyz + 48 * (size_t)(k > 0 ? k - 1 : group - 1) * n, n, cbits, dst);
      }
      if ((e = hipGetLastError()) != hipSuccess) break;
      rc = batch_normalize(c, xyz, (size_t)g * n, tab + 24 * (size_t)w0 * n, prefix);
    }
    if (rc != MZK_OK || e != hipSuccess) break;
    MZK_LAUNCH(c, "k_to_records", k_to_records, dim3(grid_for((size_t)W * n, 64)), dim3(64), 0, tab, (size_t)W * n,
               (uint32_t*)s->shifted);
    if ((e = hipGetLastError()) != hipSuccess) break;
    e = hipStreamSynchronize(c->stream);
  } while (0);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(std_tab);
  if (xyz_mem) (void)hipFree(xyz_mem);
  if (rc == MZK_OK && e != hipSuccess) {
    set_error("srs: shifted bases: %s", hipGetErrorString(e));
    rc = MZK_ERR_HIP;
  }
  if (rc != MZK_OK) {
    (void)hipFree(s->shifted);
    s->shifted = nullptr;
    return rc;
  }
  s->shifted_c = cbits;
  return MZK_OK;
}

}  // namespace mzk

using namespace mzk;

// Commitments of at least this many terms use the window-shifted bases (one shared
// bucket set); smaller ones use per-window buckets on the plain bases.
static constexpr size_t SHARED_MIN_TERMS = 1u << 14;
static constexpr size_t SHARED_MIN_BATCH_AVG = 1u << 15;
static constexpr int SHARED_C = 16;
// Shifted tables cost (255/16 + 1) x 96 B per base; built for SRS of at least this size.
static constexpr size_t SHIFTED_MIN_SRS = 1u << 14;

static int srs_finish(mzk_ctx* c, mzk_srs* s) {
  if (s->n < SHIFTED_MIN_SRS || s->n * 16 >= (1ull << 31)) return MZK_OK;
  return srs_build_shifted(c, s, SHARED_C);
}

static int srs_alloc(mzk_ctx* c, size_t n, mzk_srs** out) {
  mzk_srs* s = new mzk_srs();
  s->ctx = c;
  s->device = c->device;
  s->n = n;
  hipError_t e = hipMalloc(&s->bases, n * 96);
  if (e != hipSuccess) {
    delete s;
    set_error("srs: hipMalloc(%zu) failed: %s", n * 96, hipGetErrorString(e));
    return MZK_ERR_NOMEM;
  }
  *out = s;
  return MZK_OK;
}

namespace mzk {
int srs_alloc_bases(mzk_ctx* c, size_t n, mzk_srs** out) { return srs_alloc(c, n, out); }
int srs_finish_bases(mzk_ctx* c, mzk_srs* s) { return srs_finish(c, s); }

int srs_commit_batch(mzk_ctx* c, const mzk_srs* s, const mzk_msm_job* jobs, size_t njobs, uint64_t* outs,
                     const std::function<int()>* on_enqueued) {
  MZK_REQUIRE(c != nullptr && s != nullptr && (jobs != nullptr || njobs == 0) && (outs != nullptr || njobs == 0),
              "mzk_msm_srs_batch_dev: null argument");
  if (njobs == 0) return MZK_OK;
  std::vector<MsmJob> v(njobs);
  size_t total = 0;
  for (size_t j = 0; j < njobs; j++) {
    MZK_REQUIRE(jobs[j].offset <= s->n && jobs[j].n <= s->n - jobs[j].offset,
                "mzk_msm_srs_batch_dev: job %zu (%zu bases from %zu) exceeds the SRS (%zu)", j, jobs[j].n,
                jobs[j].offset, s->n);
    MZK_REQUIRE(jobs[j].d_coeffs != nullptr || jobs[j].n == 0, "mzk_msm_srs_batch_dev: job %zu has no coefficients",
                j);
    v[j].d_coeffs = jobs[j].d_coeffs;
    v[j].n = jobs[j].n;
    v[j].ref_offset = (uint32_t)jobs[j].offset;
    total += jobs[j].n;
  }
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  // the shared set costs 2^15 buckets of reduction per job: worth it once jobs average
  // 2^15 terms; smaller jobs take per-window sets (small sets reduce in one workgroup each)
  if (s->shifted && total / njobs >= SHARED_MIN_BATCH_AVG) {
    MsmPlan p = msm_plan(total, 0, c->num_cus, s->shifted_c);
    return msm_run_jobs(c, p, v, (const uint32_t*)s->shifted, (uint32_t)s->n, outs, true, on_enqueued);
  }
  MsmPlan p = msm_plan(total / njobs + 1, 1, c->num_cus, 0);
  MsmPlan pt = msm_plan(total, 1, c->num_cus, p.c);   // K from the batch's total digits
  if (s->shifted)   // window 0 of the shifted tables: the bases as records
    return msm_run_jobs(c, pt, v, (const uint32_t*)s->shifted, 0, outs, true, on_enqueued);
  return msm_run_jobs(c, pt, v, (const uint32_t*)s->bases, 0, outs, false, on_enqueued);
}
}  // namespace mzk

extern "C" {

int mzk_srs_create(mzk_ctx* c, const uint64_t* bases, size_t n, mzk_srs** out) {
  MZK_REQUIRE(c != nullptr && out != nullptr && (bases != nullptr || n == 0), "mzk_srs_create: null argument");
  MZK_REQUIRE(n >= 1 && n < (1ull << 31), "mzk_srs_create: n out of range");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  mzk_srs* s = nullptr;
  MZK_TRY(srs_alloc(c, n, &s));
  hipError_t e = hipMemcpyAsync(s->bases, bases, n * 96, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    mzk_srs_destroy(s);
    set_error("srs upload: %s", hipGetErrorString(e));
    return MZK_ERR_HIP;
  }
  int rc = srs_finish(c, s);
  if (rc != MZK_OK) {
    mzk_srs_destroy(s);
    return rc;
  }
  *out = s;
  return MZK_OK;
}

// Affine [s_i]G for i < n into a fresh SRS: the scalars are the powers of tau when
// d_scalars is null, else the n device scalars given (Montgomery).
static int srs_fixed_base(mzk_ctx* c, const uint64_t* tau, const void* d_scalars, size_t n, mzk_srs** out) {
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  mzk_srs* s = nullptr;
  MZK_TRY(srs_alloc(c, n, &s));
  std::vector<uint64_t> comb = comb_table();
  int rc = MZK_OK;
  do {
    if ((rc = c->scratch[SCR_IN0].reserve(comb.size() * 8 + n * 32)) != MZK_OK) break;
    if ((rc = c->scratch[SCR_MSM_C].reserve(n * 192)) != MZK_OK) break;
    uint32_t* d_comb = (uint32_t*)c->scratch[SCR_IN0].ptr;
    uint32_t* d_pows = d_comb + comb.size() * 2;
    uint32_t* d_xyzz = (uint32_t*)c->scratch[SCR_MSM_C].ptr;
    if (hipMemcpyAsync(d_comb, comb.data(), comb.size() * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
      rc = MZK_ERR_HIP;
      set_error("srs: comb upload failed");
      break;
    }
    const uint32_t* d_sc = (const uint32_t*)d_scalars;
    if (d_sc == nullptr) {
      FrPows p;
      host::Fr t;
      memcpy(t.l, tau, 32);
      for (int j = 0; j < 32; j++) {
        memcpy(p.v[j], t.l, 32);
        t = host::fr_mul(t, t);
      }
      MZK_LAUNCH(c, "k_tau_powers", k_tau_powers, dim3(grid_for(n, 256)), dim3(256), 0, d_pows, n, p);
      d_sc = d_pows;
    }
    MZK_LAUNCH(c, "k_fixed_base", k_fixed_base, dim3(grid_for(n, 64)), dim3(64), 0, d_sc, n, d_comb, d_xyzz);
    if (hipGetLastError() != hipSuccess) {
      rc = MZK_ERR_HIP;
      set_error("srs: kernel launch failed");
      break;
    }
    if ((rc = batch_normalize(c, d_xyzz, n, (uint32_t*)s->bases)) != MZK_OK) break;
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = MZK_ERR_HIP;
      set_error("srs: generation failed");
      break;
    }
    rc = srs_finish(c, s);
  } while (0);
  if (rc != MZK_OK) {
    mzk_srs_destroy(s);
    return rc;
  }
  *out = s;
  return MZK_OK;
}

int mzk_srs_generate(mzk_ctx* c, const uint64_t tau[4], size_t n, mzk_srs** out) {
  MZK_REQUIRE(c != nullptr && out != nullptr && tau != nullptr, "mzk_srs_generate: null argument");
  MZK_REQUIRE(n >= 1 && n < (1ull << 31), "mzk_srs_generate: n out of range");
  return srs_fixed_base(c, tau, nullptr, n, out);
}

int mzk_srs_from_scalars_dev(mzk_ctx* c, const void* d_scalars, size_t n, mzk_srs** out) {
  MZK_REQUIRE(c != nullptr && out != nullptr && d_scalars != nullptr, "mzk_srs_from_scalars_dev: null argument");
  MZK_REQUIRE(n >= 1 && n < (1ull << 31), "mzk_srs_from_scalars_dev: n out of range");
  return srs_fixed_base(c, nullptr, d_scalars, n, out);
}

size_t mzk_srs_len(const mzk_srs* s) { return s ? s->n : 0; }

const void* mzk_srs_bases_dev(const mzk_srs* s) { return s ? s->bases : nullptr; }

int mzk_srs_download(mzk_srs* s, size_t offset, size_t count, uint64_t* out) {
  MZK_REQUIRE(s != nullptr && (out != nullptr || count == 0), "mzk_srs_download: null argument");
  MZK_REQUIRE(offset <= s->n && count <= s->n - offset, "mzk_srs_download: range out of bounds");
  std::lock_guard<std::mutex> g(s->ctx->mu);
  MZK_HIP(hipSetDevice(s->ctx->device));
  MZK_HIP(hipMemcpyAsync(out, (const char*)s->bases + offset * 96, count * 96, hipMemcpyDeviceToHost,
                         s->ctx->stream));
  MZK_HIP(hipStreamSynchronize(s->ctx->stream));
  return MZK_OK;
}

// Any context on the device may have queued commits reading the bases, and the creating
// context may already be destroyed: wait for the whole device, touching no context.
void mzk_srs_destroy(mzk_srs* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  if (s->bases) (void)hipFree(s->bases);
  if (s->shifted) (void)hipFree(s->shifted);
  delete s;
}

int mzk_msm_srs_dev(mzk_ctx* c, const mzk_srs* s, size_t offset, const void* d_coeffs, size_t n, uint64_t out[18]) {
  MZK_REQUIRE(c != nullptr && s != nullptr && out != nullptr, "mzk_msm_srs_dev: null argument");
  MZK_REQUIRE(offset <= s->n && n <= s->n - offset, "mzk_msm_srs_dev: %zu bases from %zu exceed the SRS (%zu)", n,
              offset, s->n);
  if (n == 0) {
    memset(out, 0, 18 * 8);
    return MZK_OK;
  }
  MZK_REQUIRE(d_coeffs != nullptr, "mzk_msm_srs_dev: null coefficients");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  const MsmTune t = tune_msm_get(n);
  if (s->shifted && (t.path >= 0 ? t.path == 1 : n >= SHARED_MIN_TERMS)) {
    MsmPlan p = msm_plan(n, 0, c->num_cus, s->shifted_c);
    return msm_run(c, p, (const uint32_t*)d_coeffs, n, (const uint32_t*)s->shifted, (uint32_t)s->n, (uint32_t)offset, true,
                   out);
  }
  MsmPlan p = msm_plan(n, 1, c->num_cus, 0);
  if (s->shifted)   // window 0 of the shifted tables is the bases themselves, as records
    return msm_run(c, p, (const uint32_t*)d_coeffs, n, (const uint32_t*)s->shifted, 0, (uint32_t)offset, true, out);
  return msm_run(c, p, (const uint32_t*)d_coeffs, n, (const uint32_t*)s->bases + 24 * offset, 0, 0, false, out);
}

int mzk_msm_srs_batch_dev(mzk_ctx* c, const mzk_srs* s, const mzk_msm_job* jobs, size_t njobs, uint64_t* outs) {
  return mzk::srs_commit_batch(c, s, jobs, njobs, outs, nullptr);
}

}  // extern "C"
