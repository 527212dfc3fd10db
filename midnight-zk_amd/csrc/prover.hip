// Native prover for the synthetic PLONK circuit (five advice columns a..e, fixed
// q_m..q_c, gate a b q_m + a q_l + b q_r - c q_o + q_c, every advice column in the
// permutation, quotient degree 4), for a group of proofs of one size in lockstep.
//
// It follows create_proof (proofs/src/plonk/prover.rs:432-474 -> compute_trace :66-299
// -> finalise_proof :302-423) and multi_open (proofs/src/poly/kzg/mod.rs:110-289) step
// for step, with the same deliberate differences as the Python form of this prover
// (midnight-zk_amd/plonk.py): seeded random draws and a descriptor string in the VK hash
// (the caller passes the hash).  Every phase issues
// one batched call for the whole group (commitments through mzk_msm_srs_batch_dev,
// transforms through mzk_ntt_batch_dev, evaluations through mzk_fr_eval_batch), so a
// group of proofs shares each pipeline; everything else is the library's vector
// kernels.  Host work is the transcript (BLAKE2b), challenges and point encodings.
#include <algorithm>
#include <chrono>
#include <functional>
#include <mutex>
#include <vector>

#include "blake2b.hpp"
#include "ctx.hpp"
#include "field.hpp"
#include "host_ec.hpp"
#include "srs.hpp"

namespace mzk {
namespace prover {

using host::Fr;

constexpr int NADV = 5;
constexpr int NFIX = 5;
constexpr int BF = 6;          // blinding factors (circuit.rs:2203-2246 for this circuit)
constexpr int QDEG = 4;        // quotient limbs
constexpr int CHUNK = 3;       // permutation columns per set
constexpr int NSETS = 2;
constexpr uint64_t RNG_STREAM = 0x524E47;
constexpr int NDRAWS = NADV * (BF + 1) + NSETS * BF + (QDEG - 1);   // 50

static Fr fr_of(const uint64_t v[4]) {
  Fr r;
  memcpy(r.l, v, 32);
  return r;
}

static Fr fr_canon(const Fr& a) {   // Montgomery -> canonical limbs
  Fr one = host::fr_zero();
  one.l[0] = 1;
  return host::fr_mul(a, one);
}

static Fr fr_neg(const Fr& a) { return host::fr_sub(host::fr_zero(), a); }

// Fq::from_uniform_bytes (fq.rs:712-736): 512-bit little-endian integer mod r
static Fr fr_from_uniform(const uint8_t b[64]) {
  Fr lo, hi;
  memcpy(lo.l, b, 32);
  memcpy(hi.l, b + 32, 32);
  const Fr r2 = host::fr_const(host::kFrR2);
  const Fr r3 = host::fr_mul(r2, r2);   // R^3
  return host::fr_add(host::fr_mul(lo, r2), host::fr_mul(hi, r3));
}

struct Transcript {
  host::Blake2b h;
  std::vector<uint8_t> proof;
  Transcript() : h((const uint8_t*)"Domain separator for transcript", 31) {}
  void absorb(const uint8_t* d, size_t len) {
    const uint8_t pre = 1;
    h.update(&pre, 1);
    h.update(d, len);
  }
  void common_fr(const Fr& a) {
    const Fr c = fr_canon(a);
    absorb((const uint8_t*)c.l, 32);
  }
  void write_fr(const Fr& a) {
    const Fr c = fr_canon(a);
    absorb((const uint8_t*)c.l, 32);
    proof.insert(proof.end(), (const uint8_t*)c.l, (const uint8_t*)c.l + 32);
  }
  void write_point(const uint8_t c[48]) {
    absorb(c, 48);
    proof.insert(proof.end(), c, c + 48);
  }
  Fr squeeze() {
    const uint8_t pre = 0;
    h.update(&pre, 1);
    uint8_t d[64];
    h.digest(d);
    return fr_from_uniform(d);
  }
};

// G1Affine::to_compressed (ZCash encoding) of `count` Jacobian results (18 words each),
// with one field inversion for all of them: Montgomery's trick over their non-zero Z
// (an inversion costs the host ~45 us, as much as a few hundred products, and each
// phase's commitments are compressed before its challenge can be drawn)
static void compress_many(const uint64_t* jac, size_t count, uint8_t* out) {
  std::vector<host::Fp> pre(count);
  host::Fp run = host::fp_one();
  for (size_t i = 0; i < count; i++) {
    host::Fp z;
    memcpy(z.l, jac + 18 * i + 12, 48);
    pre[i] = run;
    if (!host::fp_is_zero(z)) run = host::fp_mul(run, z);
  }
  host::Fp inv = host::fp_inv(run);   // 1 / (product of the non-zero Z)
  host::Fp plain_one = host::fp_zero();
  plain_one.l[0] = 1;   // a product by 1 leaves Montgomery form
  // (p - 1) / 2
  static const uint64_t HALF[6] = {0xdcff7fffffffd555ull, 0x0f55ffff58a9ffffull, 0xb39869507b587b12ull,
                                   0xb23ba5c279c2895full, 0x258dd3db21a5d66bull, 0x0d0088f51cbff34dull};
  for (size_t i = count; i-- > 0;) {
    uint8_t* o = out + 48 * i;
    memset(o, 0, 48);
    host::Fp X, Y, Z;
    memcpy(X.l, jac + 18 * i, 48);
    memcpy(Y.l, jac + 18 * i + 6, 48);
    memcpy(Z.l, jac + 18 * i + 12, 48);
    if (host::fp_is_zero(Z)) {
      o[0] = 0xc0;   // the identity
      continue;
    }
    const host::Fp zi = host::fp_mul(inv, pre[i]);   // 1 / Z
    inv = host::fp_mul(inv, Z);                      // 1 / (product of the earlier non-zero Z)
    const host::Fp zi2 = host::fp_mul(zi, zi);
    const host::Fp x = host::fp_mul(host::fp_mul(X, zi2), plain_one);
    const host::Fp y = host::fp_mul(host::fp_mul(Y, host::fp_mul(zi2, zi)), plain_one);
    for (int l = 0; l < 6; l++)
      for (int b = 0; b < 8; b++) o[47 - (8 * l + b)] = (uint8_t)(x.l[l] >> (8 * b));
    o[0] |= 0x80;
    bool larger = false;
    for (int l = 5; l >= 0; l--) {
      if (y.l[l] != HALF[l]) {
        larger = y.l[l] > HALF[l];
        break;
      }
    }
    if (larger) o[0] |= 0x20;
  }
}

static Fr fr_pow_u64(const Fr& a, uint64_t e) { return host::fr_pow(a, e); }

// The field constants of a proof at 2^k rows, which depend on k alone: computed once per
// k and process (five inversions and their powers cost the host ~80 us per proof, spent
// where the device waits for the next phase's launches).
struct KConsts {
  Fr omega, omega_inv, delta, omega_last;
  Fr te[4];   // 1 / (X^n - 1) on the four cosets zeta w_ext^i of the extended domain
};
static const KConsts& k_consts(uint32_t k, uint32_t blinding_rows) {
  static std::mutex mu;
  static std::vector<std::pair<uint64_t, KConsts*>> cache;   // (k, blinding rows) -> constants, kept
  const uint64_t key = ((uint64_t)k << 32) | blinding_rows;
  std::lock_guard<std::mutex> g(mu);
  for (auto& e : cache)
    if (e.first == key) return *e.second;
  KConsts* kc = new KConsts();
  const Fr one = host::fr_one();
  const uint64_t n = (uint64_t)1 << k;
  kc->omega = host::fr_root_of_unity(k);
  kc->omega_inv = host::fr_inv(kc->omega);
  kc->delta = fr_pow_u64(host::fr_from_u64(7), 1ull << 32);   // fq.rs:501-509
  kc->omega_last = fr_pow_u64(kc->omega_inv, blinding_rows + 1);
  const Fr wext = host::fr_root_of_unity(k + 2), zeta = host::fr_const(host::kFrZeta);
  Fr cur = zeta;
  for (int i = 0; i < 4; i++) {
    kc->te[i] = host::fr_inv(host::fr_sub(fr_pow_u64(cur, n), one));
    cur = host::fr_mul(cur, wext);
  }
  cache.push_back({key, kc});
  return *kc;
}

}  // namespace prover
}  // namespace mzk

using namespace mzk;
using namespace mzk::prover;

#define PV_TRY(expr)                  \
  do {                                \
    int rc_ = (expr);                 \
    if (rc_ != MZK_OK) return rc_;    \
  } while (0)

namespace {

// z[0] = 1, the blinding rows from the draws, and z[u] (the set's last usable value,
// before any correction) copied to *last (permutation/prover.rs:186-201).  Elements
// move as two 16-byte words.
struct ElemArg {
  uint32_t v[8];
};
__device__ __forceinline__ void copy_elem(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src) {
  const uint4* s4 = reinterpret_cast<const uint4*>(src);
  uint4* d4 = reinterpret_cast<uint4*>(dst);
  d4[0] = s4[0];
  d4[1] = s4[1];
}
__global__ void k_perm_z_finish(uint32_t* __restrict__ z, const uint32_t* __restrict__ blind, uint32_t bf,
                                uint64_t n, uint64_t u, ElemArg one, uint32_t* __restrict__ last) {
  const uint32_t t = threadIdx.x;
  if (t == 0) {
    copy_elem(last, z + 8 * u);
    uint4* d4 = reinterpret_cast<uint4*>(z);
    d4[0] = make_uint4(one.v[0], one.v[1], one.v[2], one.v[3]);
    d4[1] = make_uint4(one.v[4], one.v[5], one.v[6], one.v[7]);
  }
  if (t < bf) copy_elem(z + 8 * (n - bf + t), blind + 8 * t);
}

// The advice columns with their blinding rows, one launch for all columns of a proof:
// row i < u of column j from the caller's column, rows u..n-1 from the draws
// (draws[j (bf + 1) + i - u]).  Ten copies enqueued one by one left the device idle
// ~0.15 ms at the start of every proof.
struct AdviceRows {
  const uint32_t* src[NADV];
  uint32_t* dst[NADV];
};
__global__ void k_advice_rows(AdviceRows a, const uint32_t* __restrict__ draws, uint64_t n, uint64_t u,
                              uint32_t bf) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = blockIdx.y;
  const uint32_t* s = i < u ? a.src[j] + 8 * i : draws + 8 * ((uint64_t)j * (bf + 1) + (i - u));
  copy_elem(a.dst[j] + 8 * i, s);
}

// The quotient's limbs (plonk/prover.rs:743-785): limb i holds h's coefficients
// i (n - 1) .. i (n - 1) + n - 2, then blinding value i (0 for the last limb); the first
// coefficient of limb i > 0 less blinding value i - 1.  One launch for all limbs of a
// proof, in place of eleven copies, fills and one-element subtractions.
struct QuotientLimbs {
  uint32_t* dst[QDEG];
};
__global__ void k_quotient_limbs(const uint32_t* __restrict__ h, QuotientLimbs q, const uint32_t* __restrict__ blind,
                                 uint64_t n) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t i = blockIdx.y;
  fr_t v;
  if (r + 1 < n) {
    v = load<FrP>(h + 8 * ((uint64_t)i * (n - 1) + r));
    if (i > 0 && r == 0) v = sub(v, load<FrP>(blind + 8 * (i - 1)));
  } else {
    v = i + 1 < QDEG ? load<FrP>(blind + 8 * i) : fr_t::zero();
  }
  store<FrP>(q.dst[i] + 8 * r, v);
}

// z of set s > 0 continues from the earlier sets: z[i] *= last[0] ... last[s - 1], the
// factor read on the device (no host round trip for the last values).
__global__ void k_perm_z_continue(uint32_t* __restrict__ z, uint64_t n, const uint32_t* __restrict__ last,
                                  uint32_t s) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t f = load<FrP>(last);
  for (uint32_t t = 1; t < s; t++) f = mul(f, load<FrP>(last + 8 * t));
  store<FrP>(z + 8 * i, mul(load<FrP>(z + 8 * i), f));
}

struct ProofBufs {
  uint32_t* adv[NADV];
  uint32_t* polys[NADV + NSETS];   // advice then z, coefficient form
  uint32_t* z[NSETS];              // Lagrange form
  uint32_t* cos[NADV + NSETS];
  uint32_t* hext;
  uint32_t* h;
  uint32_t* limbs[QDEG];
  uint32_t* lin;
  uint32_t* q0;
  uint32_t* k1;   // Kate chain temporaries
  uint32_t* k2;
  uint32_t* fa;
  uint32_t* fb;
  uint32_t* fc;
  uint32_t* f;
  uint32_t* fin;
  uint32_t* pi;
  uint32_t* draws;
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// on_enqueued: see msm_run_jobs (runs with the context lock held, before the host waits)
int msm_batch(mzk_ctx* c, const mzk_srs* srs, const std::vector<std::pair<const void*, size_t>>& jobs,
              std::vector<uint64_t>& outs, const std::function<int()>* on_enqueued = nullptr) {
  std::vector<mzk_msm_job> j(jobs.size());
  for (size_t i = 0; i < jobs.size(); i++) {
    j[i].d_coeffs = jobs[i].first;
    j[i].n = jobs[i].second;
    j[i].offset = 0;
  }
  outs.assign(18 * jobs.size(), 0);
  return mzk::srs_commit_batch(c, srs, j.data(), j.size(), outs.data(), on_enqueued);
}

int eval_batch(mzk_ctx* c, const std::vector<std::pair<const void*, size_t>>& polys, const std::vector<Fr>& xs,
               std::vector<Fr>& out) {
  const size_t m = polys.size();
  std::vector<const void*> ptr(m);
  std::vector<size_t> ns(m);
  std::vector<uint64_t> x(4 * m), o(4 * m);
  for (size_t i = 0; i < m; i++) {
    ptr[i] = polys[i].first;
    ns[i] = polys[i].second;
    memcpy(&x[4 * i], xs[i].l, 32);
  }
  PV_TRY(mzk_fr_eval_batch(c, ptr.data(), ns.data(), x.data(), m, o.data()));
  out.resize(m);
  for (size_t i = 0; i < m; i++) memcpy(out[i].l, &o[4 * i], 32);
  return MZK_OK;
}

int lincomb(mzk_ctx* c, const std::vector<std::pair<const void*, size_t>>& polys, const std::vector<Fr>& s,
            void* out, size_t out_len) {
  std::vector<const void*> ptr;
  std::vector<size_t> lens;
  std::vector<uint64_t> sc;
  for (size_t i = 0; i < polys.size(); i++) {
    ptr.push_back(polys[i].first);
    lens.push_back(polys[i].second);
    sc.insert(sc.end(), s[i].l, s[i].l + 4);
  }
  return mzk_fr_lincomb_dev(c, ptr.data(), lens.data(), sc.data(), ptr.size(), out, out_len);
}

int vscal(mzk_ctx* c, int op, const void* a, const Fr& s, void* out, size_t n) {
  return mzk_fr_vec_scalar(c, op, a, s.l, out, n);
}

// construct_intermediate_sets (kzg/utils.rs:40-133) for (label, point) queries
void intermediate_sets(const std::vector<std::pair<int, Fr>>& q, std::vector<int>& labels, std::vector<int>& set_of,
                       std::vector<std::vector<Fr>>& point_sets) {
  std::vector<Fr> points;
  std::vector<std::vector<int>> lpts;
  for (auto& e : q) {
    int pi = -1;
    for (size_t i = 0; i < points.size(); i++)
      if (host::fr_eq(points[i], e.second)) pi = (int)i;
    if (pi < 0) {
      pi = (int)points.size();
      points.push_back(e.second);
    }
    auto it = std::find(labels.begin(), labels.end(), e.first);
    if (it == labels.end()) {
      labels.push_back(e.first);
      lpts.push_back({pi});
    } else {
      lpts[it - labels.begin()].push_back(pi);
    }
  }
  std::vector<std::vector<int>> keys;
  for (auto& p : lpts) {
    std::vector<int> k = p;
    std::sort(k.begin(), k.end());
    auto it = std::find(keys.begin(), keys.end(), k);
    if (it == keys.end()) {
      set_of.push_back((int)keys.size());
      keys.push_back(k);
    } else {
      set_of.push_back((int)(it - keys.begin()));
    }
  }
  for (auto& k : keys) {
    std::vector<Fr> ps;
    for (int i : k) ps.push_back(points[i]);
    point_sets.push_back(ps);
  }
}

// Cross-stream ordering for one prover call: ev[0] is recorded on the main stream when
// the advice columns are in place and ev[1] on the auxiliary stream after their
// transforms; ev[2] / ev[3] the same for the permutation products z.  On every exit the
// auxiliary stream is drained (it writes into the call's arena) and the events released.
struct AuxStream {
  mzk_ctx* aux = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool used = false;
  ~AuxStream() {
    if (used) (void)hipStreamSynchronize(aux->stream);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// Columns to transform on the auxiliary stream: (source, (coefficients out, coset out)).
typedef std::vector<std::pair<const void*, std::pair<uint32_t*, uint32_t*>>> TransformCols;

// The transforms of columns cols[] of every proof on the auxiliary stream once `ready`
// (already recorded on the context stream) has passed: to coefficients (src -> polys), then
// to the extended coset (polys -> cos); `done` is recorded after them.
static int aux_enqueue(AuxStream& ax, hipEvent_t ready, hipEvent_t done, const TransformCols& cols, uint32_t k) {
  {
    std::lock_guard<std::mutex> g(ax.aux->mu);
    MZK_HIP(hipStreamWaitEvent(ax.aux->stream, ready, 0));
  }
  ax.used = true;
  std::vector<mzk_ntt_job> jobs(cols.size());
  for (size_t i = 0; i < cols.size(); i++) {
    jobs[i].d_in = cols[i].first;
    jobs[i].d_out = cols[i].second.first;
    jobs[i].n_real = (uint64_t)1 << k;
    jobs[i].log_n = k;
    jobs[i].flags = MZK_NTT_INVERSE;
  }
  MZK_TRY(mzk_ntt_batch_dev(ax.aux, jobs.data(), jobs.size()));
  for (size_t i = 0; i < cols.size(); i++) {
    jobs[i].d_in = cols[i].second.first;
    jobs[i].d_out = cols[i].second.second;
    jobs[i].log_n = k + 2;
    jobs[i].flags = MZK_NTT_COSET;
  }
  MZK_TRY(mzk_ntt_batch_dev(ax.aux, jobs.data(), jobs.size()));
  std::lock_guard<std::mutex> g(ax.aux->mu);
  MZK_HIP(hipEventRecord(done, ax.aux->stream));
  return MZK_OK;
}

// Each column to coefficients (n) and to the extended coset (4n) on the auxiliary stream,
// after the work now on the context stream; `done` is recorded after them.
static int aux_transforms(mzk_ctx* c, AuxStream& ax, hipEvent_t ready, hipEvent_t done, const TransformCols& cols,
                          uint32_t k) {
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipSetDevice(c->device));
    MZK_HIP(hipEventRecord(ready, c->stream));
  }
  return aux_enqueue(ax, ready, done, cols, k);
}

}  // namespace

extern "C" {

int mzk_plonk_prove(mzk_ctx* c, const mzk_plonk_key* key, size_t np, const void* const* advice,
                    const uint64_t* rng_seeds, uint8_t* proofs, double* phase_ms) {
  MZK_REQUIRE(c && key && (np == 0 || (advice && rng_seeds && proofs)), "mzk_plonk_prove: null argument");
  if (np == 0) return MZK_OK;
  // the arena and the per-phase buffers stay in use across the library calls below, each
  // of which takes the context lock on its own: concurrent provers on one context wait here
  std::lock_guard<std::mutex> whole(c->prove_mu);
  const uint32_t k = key->k;
  MZK_REQUIRE(k >= 4 && k <= 24, "mzk_plonk_prove: k = %u out of range", k);
  const size_t n = (size_t)1 << k, N = 4 * n, u = n - BF - 1;
  MZK_REQUIRE(key->g && key->g_lagrange && key->g->n >= n && key->g_lagrange->n >= n,
              "mzk_plonk_prove: SRS shorter than 2^%u", k);
  MZK_REQUIRE(!key->g_lagrange_delta || key->g_lagrange_delta->n == n,
              "mzk_plonk_prove: delta bases must hold exactly 2^%u points", k);
  const void* need[] = {key->l0, key->l_last, key->l_active, key->coset_x, key->omega_powers};
  for (const void* p : need) MZK_REQUIRE(p, "mzk_plonk_prove: key vector missing");
  for (int j = 0; j < NFIX; j++) MZK_REQUIRE(key->fixed_polys[j] && key->fixed_cosets[j], "mzk_plonk_prove: fixed %d", j);
  for (int j = 0; j < NADV; j++)
    MZK_REQUIRE(key->sigma_values[j] && key->sigma_polys[j] && key->sigma_cosets[j], "mzk_plonk_prove: sigma %d", j);
  for (size_t p = 0; p < np * NADV; p++) MZK_REQUIRE(advice[p], "mzk_plonk_prove: advice column %zu missing", p);
  double t0 = now_ms();
  double* ph = phase_ms;
  auto mark = [&](int i) {
    if (!ph) return;
    mzk_ctx_synchronize(c);
    ph[i] = now_ms() - t0;
  };

  // ---- device arena: per proof 64 n scalars plus the draws
  const size_t per = (size_t)(NADV + (NADV + NSETS) + NSETS + 4 * (NADV + NSETS) + 4 + 4 + QDEG + 10) * n + 64;
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipSetDevice(c->device));
    // plus, for the group: numerators and denominators of every (proof, set), and the
    // sets' last values
    PV_TRY(c->scratch[SCR_PROVER].reserve(per * 32 * np + (2 * n + 1) * 32 * np * NSETS + 256));
  }
  uint32_t* arena = (uint32_t*)c->scratch[SCR_PROVER].ptr;
  std::vector<ProofBufs> B(np);
  for (size_t p = 0; p < np; p++) {
    uint32_t* q = arena + 8 * per * p;
    auto take = [&](size_t elems) {
      uint32_t* r = q;
      q += 8 * elems;
      return r;
    };
    ProofBufs& b = B[p];
    for (int j = 0; j < NADV; j++) b.adv[j] = take(n);
    for (int j = 0; j < NADV + NSETS; j++) b.polys[j] = take(n);
    for (int s = 0; s < NSETS; s++) b.z[s] = take(n);
    for (int j = 0; j < NADV + NSETS; j++) b.cos[j] = take(N);
    b.hext = take(N);
    b.h = take(N);
    for (int i = 0; i < QDEG; i++) b.limbs[i] = take(n);
    b.lin = take(n);
    b.q0 = take(n);
    b.k1 = take(n);
    b.k2 = take(n);
    b.fa = take(n);
    b.fb = take(n);
    b.fc = take(n);
    b.f = take(n);
    b.fin = take(n);
    b.pi = take(n);
    b.draws = take(64);
  }
  hipStream_t st = c->stream;
  std::vector<Transcript> T(np);
  const Fr one = host::fr_one();
  const KConsts& kc = k_consts(k, BF);
  const Fr omega = kc.omega, delta = kc.delta;

  // ---- VK hash, advice with blinding rows, advice commitments (prover.rs:101-108, 541-631)
  for (size_t p = 0; p < np; p++) {
    T[p].common_fr(fr_of(key->transcript_repr));
    PV_TRY(mzk_fr_from_seed_dev(c, rng_seeds[p], RNG_STREAM, 0, NDRAWS, B[p].draws));
    AdviceRows ar;
    for (int j = 0; j < NADV; j++) {
      ar.src[j] = (const uint32_t*)advice[p * NADV + j];
      ar.dst[j] = B[p].adv[j];
    }
    std::lock_guard<std::mutex> g(c->mu);
    hipLaunchKernelGGL(k_advice_rows, dim3((unsigned)((n + 255) / 256), NADV), dim3(256), 0, st, ar, B[p].draws,
                       (uint64_t)n, (uint64_t)u, (uint32_t)BF);
    MZK_HIP(hipGetLastError());
  }
  // the advice columns' transforms (to coefficients, to the extended coset) depend on the
  // advice alone: they run on the auxiliary stream from the end of the advice commitments'
  // kernels on, beside the host's tail of those commitments, the transcript and the
  // permutation products (small, latency-bound kernels that leave most SIMDs idle); the
  // numerator waits for them.  Started with the commitments instead, they competed with
  // the accumulation for the same VALU slots and saved nothing.
  AuxStream ax;
  TransformCols adv_cols;
  {
    if (!c->aux) PV_TRY(mzk_ctx_create(c->device, &c->aux));
    ax.aux = c->aux;
    MZK_HIP(hipSetDevice(c->device));
    for (hipEvent_t& e : ax.ev) MZK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (size_t p = 0; p < np; p++)
      for (int j = 0; j < NADV; j++) adv_cols.push_back({B[p].adv[j], {B[p].polys[j], B[p].cos[j]}});
  }
  std::vector<uint64_t> jac;
  {
    std::vector<std::pair<const void*, size_t>> jobs;
    for (size_t p = 0; p < np; p++)
      for (int j = 0; j < NADV; j++) jobs.push_back({B[p].adv[j], n});
    bool queued = false;
    const std::function<int()> after_commit_kernels = [&]() -> int {
      MZK_HIP(hipEventRecord(ax.ev[0], c->stream));   // the context lock is held here
      MZK_TRY(aux_enqueue(ax, ax.ev[0], ax.ev[1], adv_cols, k));
      queued = true;
      return MZK_OK;
    };
    PV_TRY(msm_batch(c, key->g_lagrange, jobs, jac, &after_commit_kernels));
    if (!queued) PV_TRY(aux_transforms(c, ax, ax.ev[0], ax.ev[1], adv_cols, k));
    std::vector<uint8_t> cb(48 * jobs.size());
    compress_many(jac.data(), jobs.size(), cb.data());
    for (size_t p = 0; p < np; p++)
      for (int j = 0; j < NADV; j++) T[p].write_point(&cb[48 * (p * NADV + j)]);
  }
  mark(0);
  std::vector<Fr> beta(np), gamma(np), y(np);
  for (size_t p = 0; p < np; p++) {
    T[p].squeeze();   // theta (no lookups)
    beta[p] = T[p].squeeze();
    gamma[p] = T[p].squeeze();
  }

  // ---- permutation grand products (permutation/prover.rs:81-255): the products of every
  // (proof, set) in one pass, one batch inversion and one quotient for all of them, the
  // running products, then one read-back of the sets' last values for the corrections
  {
    const size_t J = np * NSETS;
    uint32_t* num_all = arena + 8 * per * np;
    uint32_t* den_all = num_all + 8 * J * n;
    uint32_t* lasts = den_all + 8 * J * n;
    std::vector<mzk_perm_set> sets(J);
    for (size_t p = 0; p < np; p++)
      for (int s = 0; s < NSETS; s++) {
        mzk_perm_set& e = sets[p * NSETS + s];
        memset(&e, 0, sizeof e);
        const int j0 = s * CHUNK, j1 = std::min(NADV, j0 + CHUNK);
        e.ncols = (uint32_t)(j1 - j0);
        for (int j = j0; j < j1; j++) {
          e.values[j - j0] = B[p].adv[j];
          e.sigma[j - j0] = key->sigma_values[j];
          const Fr bd = host::fr_mul(beta[p], fr_pow_u64(delta, (uint64_t)j));
          memcpy(e.delta_beta[j - j0], bd.l, 32);
        }
        memcpy(e.beta, beta[p].l, 32);
        memcpy(e.gamma, gamma[p].l, 32);
        e.num = num_all + 8 * (p * NSETS + s) * n;
        e.den = den_all + 8 * (p * NSETS + s) * n;
      }
    PV_TRY(mzk_plonk_perm_products_dev(c, key->omega_powers, n, sets.data(), J));
    PV_TRY(mzk_fr_batch_invert(c, den_all, J * n));
    PV_TRY(mzk_fr_vec_op(c, 2, num_all, den_all, num_all, J * n));
    // z_0 = 1, z_i = z_(i-1) mv_(i-1): every set's running products in one batch
    std::vector<mzk_prefix_job> pj(J);
    for (size_t js = 0; js < J; js++) {
      pj[js].d_a = num_all + 8 * js * n;
      pj[js].d_out = B[js / NSETS].z[js % NSETS] + 8;
      pj[js].n = n - 1;
    }
    PV_TRY(mzk_fr_prefix_product_batch(c, pj.data(), J));
    for (size_t p = 0; p < np; p++)
      for (int s = 0; s < NSETS; s++) {
        const size_t js = p * NSETS + s;
        uint32_t* z = B[p].z[s];
        // blinding rows from the draws
        const uint32_t* blind = B[p].draws + 8 * (size_t)(NADV * (BF + 1) + s * BF);
        std::lock_guard<std::mutex> g(c->mu);
        ElemArg one_w;
        memcpy(one_w.v, one.l, 32);
        hipLaunchKernelGGL(k_perm_z_finish, dim3(1), dim3(64), 0, st, z, blind, (uint32_t)BF, (uint64_t)n,
                           (uint64_t)u, one_w, lasts + 8 * js);
        MZK_HIP(hipGetLastError());
      }
    // set s > 0 continues from the earlier sets' last values: scale by their product
    for (size_t p = 0; p < np; p++)
      for (int s = 1; s < NSETS; s++) {
        std::lock_guard<std::mutex> g(c->mu);
        hipLaunchKernelGGL(k_perm_z_continue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B[p].z[s],
                           (uint64_t)n, lasts + 8 * p * NSETS, (uint32_t)s);
        MZK_HIP(hipGetLastError());
      }
  }
  {   // z's transforms overlap its commitment (few non-zero deltas: mostly the MSM's sort
      // and reduction phases)
    TransformCols cols;
    for (size_t p = 0; p < np; p++)
      for (int s = 0; s < NSETS; s++) cols.push_back({B[p].z[s], {B[p].polys[NADV + s], B[p].cos[NADV + s]}});
    PV_TRY(aux_transforms(c, ax, ax.ev[2], ax.ev[3], cols, k));
  }
  {
    // z has long constant runs (rows outside every copy cycle): committed as z.to_delta()
    // against the suffix-sum bases (permutation/prover.rs:236-245), the zero deltas give the
    // MSM no window digits.  The deltas go where the numerators were (free by now).
    std::vector<std::pair<const void*, size_t>> jobs;
    const bool delta = key->g_lagrange_delta != nullptr;
    uint32_t* deltas = arena + 8 * per * np;
    for (size_t p = 0; p < np; p++)
      for (int s = 0; s < NSETS; s++) {
        if (delta) {
          uint32_t* d = deltas + 8 * (p * NSETS + s) * n;
          PV_TRY(mzk_fr_to_delta_dev(c, B[p].z[s], d, n, 1));
          jobs.push_back({d, n});
        } else {
          jobs.push_back({B[p].z[s], n});
        }
      }
    PV_TRY(msm_batch(c, delta ? key->g_lagrange_delta : key->g_lagrange, jobs, jac));
    std::vector<uint8_t> cb(48 * jobs.size());
    compress_many(jac.data(), jobs.size(), cb.data());
    for (size_t p = 0; p < np; p++)
      for (int s = 0; s < NSETS; s++) T[p].write_point(&cb[48 * (p * NSETS + s)]);
  }
  mark(1);
  for (size_t p = 0; p < np; p++) {
    T[p].squeeze();   // trash challenge (no trashcans)
    y[p] = T[p].squeeze();
  }

  // ---- to coefficients, to the extended coset, numerator / (X^n - 1), back (prover.rs:634-772):
  // the columns' transforms ran on the auxiliary stream; the numerator waits for them
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipStreamWaitEvent(st, ax.ev[1], 0));
    MZK_HIP(hipStreamWaitEvent(st, ax.ev[3], 0));
  }
  mark(2);
  mark(3);
  const Fr* te = kc.te;
  for (size_t p = 0; p < np; p++) {
    mzk_plonk_numerator A;
    memset(&A, 0, sizeof A);
    A.n_ext = N;
    A.rot_next = 4;
    A.rot_last = (BF + 1) * 4;
    A.ncols = NADV;
    A.nsets = NSETS;
    A.chunk = CHUNK;
    A.t_count = 4;
    const ProofBufs& b = B[p];
    A.a = b.cos[0];
    A.b = b.cos[1];
    A.c = b.cos[2];
    A.qm = key->fixed_cosets[0];
    A.ql = key->fixed_cosets[1];
    A.qr = key->fixed_cosets[2];
    A.qo = key->fixed_cosets[3];
    A.qc = key->fixed_cosets[4];
    for (int j = 0; j < NADV; j++) {
      A.col[j] = b.cos[j];
      A.sigma[j] = key->sigma_cosets[j];
      const Fr bd = host::fr_mul(beta[p], fr_pow_u64(delta, (uint64_t)j));
      memcpy(A.delta_beta[j], bd.l, 32);
    }
    for (int s = 0; s < NSETS; s++) A.z[s] = b.cos[NADV + s];
    A.l0 = key->l0;
    A.l_last = key->l_last;
    A.l_active = key->l_active;
    A.coset_x = key->coset_x;
    memcpy(A.y, y[p].l, 32);
    memcpy(A.beta, beta[p].l, 32);
    memcpy(A.gamma, gamma[p].l, 32);
    for (int i = 0; i < 4; i++) memcpy(A.t_evals[i], te[i].l, 32);
    PV_TRY(mzk_plonk_numerator_dev(c, &A, b.hext));
  }
  mark(4);
  {
    std::vector<mzk_ntt_job> jobs;
    for (size_t p = 0; p < np; p++) {
      mzk_ntt_job t;
      t.d_in = B[p].hext;
      t.d_out = B[p].h;
      t.n_real = N;
      t.log_n = k + 2;
      t.flags = MZK_NTT_INVERSE | MZK_NTT_COSET;
      jobs.push_back(t);
    }
    PV_TRY(mzk_ntt_batch_dev(c, jobs.data(), jobs.size()));
  }
  // limbs of n - 1 coefficients, each with a blinding tail; limb i > 0 minus the previous tail
  for (size_t p = 0; p < np; p++) {
    ProofBufs& b = B[p];
    const uint32_t* blind = b.draws + 8 * (size_t)(NADV * (BF + 1) + NSETS * BF);
    QuotientLimbs q;
    for (int i = 0; i < QDEG; i++) q.dst[i] = b.limbs[i];
    std::lock_guard<std::mutex> g(c->mu);
    hipLaunchKernelGGL(k_quotient_limbs, dim3((unsigned)((n + 255) / 256), QDEG), dim3(256), 0, st, b.h, q, blind,
                       (uint64_t)n);
    MZK_HIP(hipGetLastError());
  }
  {
    std::vector<std::pair<const void*, size_t>> jobs;
    for (size_t p = 0; p < np; p++)
      for (int i = 0; i < QDEG; i++) jobs.push_back({B[p].limbs[i], n});
    PV_TRY(msm_batch(c, key->g, jobs, jac));
    std::vector<uint8_t> cb(48 * jobs.size());
    compress_many(jac.data(), jobs.size(), cb.data());
    for (size_t p = 0; p < np; p++)
      for (int i = 0; i < QDEG; i++) T[p].write_point(&cb[48 * (p * QDEG + i)]);
  }
  mark(5);

  // ---- evaluations (prover.rs:798-869, permutation/prover.rs:266-340)
  std::vector<Fr> x(np), xn1(np), xl(np);
  const Fr omega_last = kc.omega_last;
  {
    std::vector<std::pair<const void*, size_t>> polys;
    std::vector<Fr> pts;
    for (size_t p = 0; p < np; p++) {
      x[p] = T[p].squeeze();
      xn1[p] = host::fr_mul(x[p], omega);
      xl[p] = host::fr_mul(x[p], omega_last);
      for (int j = 0; j < NADV; j++) polys.push_back({B[p].polys[j], n}), pts.push_back(x[p]);
      for (int j = 0; j < NFIX; j++) polys.push_back({key->fixed_polys[j], n}), pts.push_back(x[p]);
      for (int j = 0; j < NADV; j++) polys.push_back({key->sigma_polys[j], n}), pts.push_back(x[p]);
      const void* z0 = B[p].polys[NADV];
      const void* z1 = B[p].polys[NADV + 1];
      polys.push_back({z0, n}), pts.push_back(x[p]);
      polys.push_back({z0, n}), pts.push_back(xn1[p]);
      polys.push_back({z0, n}), pts.push_back(xl[p]);
      polys.push_back({z1, n}), pts.push_back(x[p]);
      polys.push_back({z1, n}), pts.push_back(xn1[p]);
    }
    std::vector<Fr> ev;
    PV_TRY(eval_batch(c, polys, pts, ev));
    const size_t per_p = ev.size() / np;
    for (size_t p = 0; p < np; p++)
      for (size_t i = 0; i < per_p; i++) T[p].write_fr(ev[p * per_p + i]);
  }
  mark(6);

  // ---- linearization and multi-open (linearization/prover.rs:42-86, kzg/mod.rs:110-289)
  // query labels: advice 0..4, perm 10 + s, fixed 20 + j, sigma 30 + j, lin 40
  std::vector<std::vector<Fr>> point_sets;
  std::vector<int> labels, set_of;
  std::vector<std::vector<std::pair<int, Fr>>> Qs(np);
  for (size_t p = 0; p < np; p++) {
    auto& q = Qs[p];
    for (int j = 0; j < NADV; j++) q.push_back({j, x[p]});
    q.push_back({10, x[p]});
    q.push_back({10, xn1[p]});
    q.push_back({11, x[p]});
    q.push_back({11, xn1[p]});
    q.push_back({10, xl[p]});
    for (int j = 0; j < NFIX; j++) q.push_back({20 + j, x[p]});
    for (int j = 0; j < NADV; j++) q.push_back({30 + j, x[p]});
    q.push_back({40, x[p]});
  }
  std::vector<Fr> x3(np), x4(np);
  for (size_t p = 0; p < np; p++) {
    ProofBufs& b = B[p];
    const Fr split = fr_pow_u64(x[p], n - 1);
    const Fr xn = host::fr_mul(split, x[p]);
    Fr sp = host::fr_sub(xn, one);
    std::vector<std::pair<const void*, size_t>> lp;
    std::vector<Fr> lc;
    for (int i = 0; i < QDEG; i++) {
      lp.push_back({b.limbs[i], n});
      lc.push_back(fr_neg(sp));
      sp = host::fr_mul(sp, split);
    }
    PV_TRY(lincomb(c, lp, lc, b.lin, n));
  }
  std::vector<Fr> x1(np), x2(np);
  for (size_t p = 0; p < np; p++) {
    x1[p] = T[p].squeeze();
    x2[p] = T[p].squeeze();
  }
  // q polynomials per proof, then every proof's point-set divisions stage by stage: stage i
  // divides each set with more than i points once, all sets of all proofs in one batch
  // (the reference divides the sets in parallel, kzg/mod.rs:226-236)
  struct Chain {
    const void* cur;
    size_t len;
    std::vector<Fr> pts;
    uint32_t* out;
    uint32_t* tmp[2];
  };
  std::vector<Chain> chains;
  for (size_t p = 0; p < np; p++) {
    ProofBufs& b = B[p];
    labels.clear();
    set_of.clear();
    point_sets.clear();
    intermediate_sets(Qs[p], labels, set_of, point_sets);
    auto poly_of = [&](int label) -> std::pair<const void*, size_t> {
      if (label < NADV) return {b.polys[label], n};
      if (label == 10) return {b.polys[NADV], n};
      if (label == 11) return {b.polys[NADV + 1], n};
      if (label >= 20 && label < 20 + NFIX) return {key->fixed_polys[label - 20], n};
      if (label >= 30 && label < 30 + NADV) return {key->sigma_polys[label - 30], n};
      return {b.lin, n};
    };
    std::vector<std::vector<std::pair<const void*, size_t>>> members(point_sets.size());
    for (size_t i = 0; i < labels.size(); i++) members[set_of[i]].push_back(poly_of(labels[i]));
    std::vector<int> ord(point_sets.size());
    for (size_t i = 0; i < ord.size(); i++) ord[i] = (int)i;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int bb) { return point_sets[a].size() < point_sets[bb].size(); });
    MZK_REQUIRE(ord.size() == 3, "mzk_plonk_prove: unexpected query structure");
    // q polynomials: sum of the set's members with powers of x1 (into q0, k1, k2 by sorted slot)
    uint32_t* qbuf[3] = {b.q0, b.k1, b.k2};
    for (int slot = 0; slot < 3; slot++) {
      auto& m = members[ord[slot]];
      std::vector<Fr> pw;
      Fr t = one;
      for (size_t i = 0; i < m.size(); i++) {
        pw.push_back(t);
        t = host::fr_mul(t, x1[p]);
      }
      PV_TRY(lincomb(c, m, pw, qbuf[slot], n));
    }
    // intermediate quotients ping-pong through fin, pi and lin (free from here until the
    // final phase; lin only fed the q polynomials): a set of L points takes min(L - 1, 2)
    uint32_t* spare[3] = {b.fin, b.pi, b.lin};
    int used = 0;
    uint32_t* fout[3] = {b.fa, b.fb, b.fc};
    for (int slot = 0; slot < 3; slot++) {
      Chain ch;
      ch.cur = qbuf[slot];
      ch.len = n;
      ch.pts = point_sets[ord[slot]];
      ch.out = fout[slot];
      const int need = (int)std::min<size_t>(ch.pts.size() > 0 ? ch.pts.size() - 1 : 0, 2);
      MZK_REQUIRE(used + need <= 3, "mzk_plonk_prove: point sets need more intermediate buffers");
      ch.tmp[0] = need > 0 ? spare[used] : nullptr;
      ch.tmp[1] = need > 1 ? spare[used + 1] : nullptr;
      used += need;
      chains.push_back(ch);
    }
  }
  for (size_t stage = 0;; stage++) {
    std::vector<mzk_kate_job> kj;
    for (Chain& ch : chains) {
      if (stage >= ch.pts.size()) continue;
      uint32_t* dst = stage + 1 == ch.pts.size() ? ch.out : ch.tmp[stage & 1];
      mzk_kate_job j;
      j.d_a = ch.cur;
      j.n = ch.len;
      memcpy(j.b, ch.pts[stage].l, 32);
      j.d_q = dst;
      kj.push_back(j);
      ch.cur = dst;
      ch.len -= 1;
    }
    if (kj.empty()) break;
    PV_TRY(mzk_fr_kate_div_batch(c, kj.data(), kj.size()));
  }
  for (size_t p = 0; p < np; p++) {
    ProofBufs& b = B[p];
    // f = sum_slot x2^slot q_slot / prod (X - point)
    std::vector<std::pair<const void*, size_t>> fparts;
    for (int slot = 0; slot < 3; slot++) fparts.push_back({chains[3 * p + slot].out, chains[3 * p + slot].len});
    std::vector<Fr> pw = {one, x2[p], host::fr_mul(x2[p], x2[p])};
    PV_TRY(lincomb(c, fparts, pw, b.f, n - 1));
  }
  {
    std::vector<std::pair<const void*, size_t>> jobs;
    for (size_t p = 0; p < np; p++) jobs.push_back({B[p].f, n - 1});
    PV_TRY(msm_batch(c, key->g, jobs, jac));
    std::vector<uint8_t> cb(48 * jobs.size());
    compress_many(jac.data(), jobs.size(), cb.data());
    for (size_t p = 0; p < np; p++) T[p].write_point(&cb[48 * p]);
  }
  {
    std::vector<std::pair<const void*, size_t>> polys;
    std::vector<Fr> pts;
    for (size_t p = 0; p < np; p++) {
      x3[p] = T[p].squeeze();
      polys.push_back({B[p].q0, n}), pts.push_back(x3[p]);
      polys.push_back({B[p].k1, n}), pts.push_back(x3[p]);
      polys.push_back({B[p].k2, n}), pts.push_back(x3[p]);
    }
    std::vector<Fr> ev;
    PV_TRY(eval_batch(c, polys, pts, ev));
    for (size_t p = 0; p < np; p++)
      for (int i = 0; i < 3; i++) T[p].write_fr(ev[3 * p + i]);
  }
  {
    std::vector<std::pair<const void*, size_t>> polys;
    std::vector<Fr> pts;
    for (size_t p = 0; p < np; p++) {
      ProofBufs& b = B[p];
      x4[p] = T[p].squeeze();
      Fr t = one;
      std::vector<Fr> pw;
      for (int i = 0; i < 4; i++) {
        pw.push_back(t);
        t = host::fr_mul(t, x4[p]);
      }
      std::vector<std::pair<const void*, size_t>> fp = {{b.q0, n}, {b.k1, n}, {b.k2, n}, {b.f, n - 1}};
      PV_TRY(lincomb(c, fp, pw, b.fin, n));
      polys.push_back({b.fin, n}), pts.push_back(x3[p]);
    }
    std::vector<Fr> v;
    PV_TRY(eval_batch(c, polys, pts, v));
    std::vector<mzk_kate_job> kj(np);
    for (size_t p = 0; p < np; p++) {
      ProofBufs& b = B[p];
      PV_TRY(vscal(c, 1, b.fin, fr_neg(v[p]), b.fin, 1));   // final(X) - v
      kj[p].d_a = b.fin;
      kj[p].n = n;
      memcpy(kj[p].b, x3[p].l, 32);
      kj[p].d_q = b.pi;
    }
    PV_TRY(mzk_fr_kate_div_batch(c, kj.data(), np));
    std::vector<std::pair<const void*, size_t>> jobs;
    for (size_t p = 0; p < np; p++) jobs.push_back({B[p].pi, n - 1});
    PV_TRY(msm_batch(c, key->g, jobs, jac));
    std::vector<uint8_t> cb(48 * jobs.size());
    compress_many(jac.data(), jobs.size(), cb.data());
    for (size_t p = 0; p < np; p++) T[p].write_point(&cb[48 * p]);
  }
  mark(7);
  for (size_t p = 0; p < np; p++) {
    MZK_REQUIRE(T[p].proof.size() == MZK_PLONK_PROOF_BYTES, "mzk_plonk_prove: proof of %zu bytes",
                T[p].proof.size());
    memcpy(proofs + MZK_PLONK_PROOF_BYTES * p, T[p].proof.data(), MZK_PLONK_PROOF_BYTES);
  }
  return MZK_OK;
}

}  // extern "C"
