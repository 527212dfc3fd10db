// Autotuner for the MSM and NTT plans.
//
// The MSM planner's choices -- the resident-SRS path (one bucket set per window, or
// one shared set over the window-shifted bases), the window bits c of the per-window
// path, and how many accumulation threads per CU the sorted digits are cut into (the
// bucket partitioning: K references per thread) -- and the NTT's four-step split
// (sub-DFT sizes 2^l1 x 2^l2 staged in LDS) and workgroup target (sub-DFTs per
// workgroup) come from built-in cost models unless the table here says otherwise.
// mzk_tune_msm / mzk_tune_ntt time each candidate on seeded inputs on the caller's
// device, check that every candidate returns the same result (plans change the order
// of additions, never the group element or the DFT values), and keep the fastest per
// size.  The table is text-exportable so a tuned table can be shipped with a
// deployment and imported at start-up.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "ntt.hpp"
#include "srs.hpp"
#include "tune.hpp"

namespace mzk {

static constexpr int TUNE_LOGS = 32;

struct TuneTable {
  std::mutex mu;
  MsmTune msm[TUNE_LOGS];
  NttTune ntt[TUNE_LOGS];
};

static TuneTable& table() {
  static TuneTable t;
  return t;
}

static thread_local bool forced_msm_on = false, forced_ntt_on = false;
static thread_local MsmTune forced_msm;
static thread_local NttTune forced_ntt;

static int log2_floor(size_t n) {
  int l = 0;
  while (l + 1 < TUNE_LOGS && ((size_t)1 << (l + 1)) <= n) l++;
  return l;
}

MsmTune tune_msm_get(size_t n) {
  if (forced_msm_on) return forced_msm;
  TuneTable& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  return t.msm[log2_floor(n)];
}

NttTune tune_ntt_get(uint32_t log_n) {
  if (forced_ntt_on) return forced_ntt;
  if (log_n >= (uint32_t)TUNE_LOGS) return NttTune();
  TuneTable& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  return t.ntt[log_n];
}

void tune_force_msm(const MsmTune* v) {
  forced_msm_on = v != nullptr;
  if (v) forced_msm = *v;
}

void tune_force_ntt(const NttTune* v) {
  forced_ntt_on = v != nullptr;
  if (v) forced_ntt = *v;
}

struct ForceMsm {
  explicit ForceMsm(const MsmTune& t) { tune_force_msm(&t); }
  ~ForceMsm() { tune_force_msm(nullptr); }
};
struct ForceNtt {
  explicit ForceNtt(const NttTune& t) { tune_force_ntt(&t); }
  ~ForceNtt() { tune_force_ntt(nullptr); }
};

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// affine form of a Jacobian result, for comparing candidates (equal points, any Z)
static std::vector<uint64_t> affine_of(const uint64_t jac[18]) {
  std::vector<uint64_t> a(12);
  mzk_g1_jacobian_to_affine(jac, a.data());
  return a;
}

}  // namespace mzk

using namespace mzk;

extern "C" {

int mzk_tune_msm(mzk_ctx* c, const mzk_srs* s, uint32_t log_lo, uint32_t log_hi, uint32_t reps) {
  MZK_REQUIRE(c != nullptr && s != nullptr, "mzk_tune_msm: null argument");
  MZK_REQUIRE(log_lo <= log_hi && log_hi < 31 && ((size_t)1 << log_hi) <= s->n,
              "mzk_tune_msm: sizes 2^%u..2^%u outside the SRS (%zu bases)", log_lo, log_hi, s->n);
  MZK_REQUIRE(reps >= 1 && reps <= 100, "mzk_tune_msm: reps %u outside [1, 100]", reps);
  const size_t nmax = (size_t)1 << log_hi;
  void* d = nullptr;
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipSetDevice(c->device));
    MZK_HIP(hipMalloc(&d, nmax * 32));
  }
  int rc = mzk_fr_from_seed_dev(c, 0x7475AE, 11, 0, nmax, d);
  const bool has_shared = s->shifted != nullptr;
  for (uint32_t L = log_lo; rc == MZK_OK && L <= log_hi; L++) {
    const size_t n = (size_t)1 << L;
    std::vector<MsmTune> cand;
    for (int cb : {4, 8, 16})
      for (int tpc : {256, 512, 1024}) {
        MsmTune t;
        t.path = 0;
        t.c = (int8_t)cb;
        t.tpc_window = (uint16_t)tpc;
        cand.push_back(t);
      }
    if (has_shared)
      for (int tpc : {256, 512, 1024}) {
        MsmTune t;
        t.path = 1;
        t.tpc_shared = (uint16_t)tpc;
        cand.push_back(t);
      }
    std::vector<uint64_t> ref;
    for (const MsmTune& t : cand) {   // warm-up, and the result check
      ForceMsm f(t);
      uint64_t out[18];
      if ((rc = mzk_msm_srs_dev(c, s, 0, d, n, out)) != MZK_OK) break;
      const std::vector<uint64_t> a = affine_of(out);
      if (ref.empty()) {
        ref = a;
      } else if (a != ref) {
        set_error("mzk_tune_msm: plan (path %d, c %d) gave a different point at n = 2^%u", t.path, t.c, L);
        rc = MZK_ERR_INVALID;
        break;
      }
    }
    if (rc != MZK_OK) break;
    // rounds over all candidates, best of each candidate's rounds: clock changes during
    // the tuning then weigh on every candidate alike instead of on the last ones timed
    std::vector<double> tmin(cand.size(), 1e300);
    for (uint32_t r = 0; r < reps && rc == MZK_OK; r++)
      for (size_t i = 0; i < cand.size() && rc == MZK_OK; i++) {
        ForceMsm f(cand[i]);
        uint64_t out[18];
        const double t0 = now_ms();
        rc = mzk_msm_srs_dev(c, s, 0, d, n, out);
        tmin[i] = std::min(tmin[i], now_ms() - t0);
      }
    if (rc != MZK_OK) break;
    const MsmTune pick = cand[std::min_element(tmin.begin(), tmin.end()) - tmin.begin()];
    TuneTable& tt = table();
    std::lock_guard<std::mutex> g(tt.mu);
    MsmTune& e = tt.msm[L];
    e.path = has_shared ? pick.path : -1;
    if (pick.path == 0) {
      e.c = pick.c;
      e.tpc_window = pick.tpc_window;
    } else {
      e.tpc_shared = pick.tpc_shared;
    }
  }
  {
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d);
  }
  return rc;
}

int mzk_tune_ntt(mzk_ctx* c, uint32_t log_lo, uint32_t log_hi, uint32_t reps) {
  MZK_REQUIRE(c != nullptr, "mzk_tune_ntt: null context");
  MZK_REQUIRE(log_lo <= log_hi && log_hi <= 22, "mzk_tune_ntt: sizes 2^%u..2^%u outside [1, 2^22]", log_lo, log_hi);
  MZK_REQUIRE(reps >= 1 && reps <= 100, "mzk_tune_ntt: reps %u outside [1, 100]", reps);
  const uint32_t lo = std::max<uint32_t>(log_lo, 12);   // smaller sizes are one LDS pass: nothing to choose
  if (lo > log_hi) return MZK_OK;
  const size_t nmax = (size_t)1 << log_hi;
  void *din = nullptr, *dout = nullptr;
  {
    std::lock_guard<std::mutex> g(c->mu);
    MZK_HIP(hipSetDevice(c->device));
    MZK_HIP(hipMalloc(&din, nmax * 32));
    if (hipMalloc(&dout, nmax * 32) != hipSuccess) {
      (void)hipFree(din);
      set_error("mzk_tune_ntt: hipMalloc failed");
      return MZK_ERR_NOMEM;
    }
  }
  int rc = mzk_fr_from_seed_dev(c, 0x4E5454, 12, 0, nmax, din);
  std::vector<uint8_t> ref, got;
  for (uint32_t L = lo; rc == MZK_OK && L <= log_hi; L++) {
    const size_t n = (size_t)1 << L;
    std::vector<NttTune> cand;
    for (uint32_t l1 = 1; l1 <= 11; l1++) {
      if (L - l1 > 11 || L <= l1 || (int)l1 - (int)(L / 2) > 2 || (int)(L / 2) - (int)l1 > 2) continue;
      for (int gr : {256, 1024, 4096}) {
        NttTune t;
        t.l1 = (uint8_t)l1;
        t.groups = (uint16_t)gr;
        cand.push_back(t);
      }
    }
    mzk_ntt_job job;
    memset(&job, 0, sizeof job);
    job.d_in = din;
    job.d_out = dout;
    job.n_real = n;
    job.log_n = L;
    job.flags = 0;
    ref.clear();
    for (const NttTune& t : cand) {   // warm-up, and the result check
      ForceNtt f(t);
      if ((rc = mzk_ntt_dev(c, &job)) != MZK_OK) break;
      got.resize(n * 32);
      {
        std::lock_guard<std::mutex> g(c->mu);
        if (hipMemcpyAsync(got.data(), dout, n * 32, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
          set_error("mzk_tune_ntt: result copy failed");
          rc = MZK_ERR_HIP;
          break;
        }
      }
      if (ref.empty()) {
        ref = got;
      } else if (got != ref) {
        set_error("mzk_tune_ntt: split 2^%u x 2^%u gave different values at 2^%u", t.l1, L - t.l1, L);
        rc = MZK_ERR_INVALID;
        break;
      }
    }
    if (rc != MZK_OK) break;
    std::vector<double> tmin(cand.size(), 1e300);   // interleaved rounds, as for the MSM
    for (uint32_t r = 0; r < reps && rc == MZK_OK; r++)
      for (size_t i = 0; i < cand.size() && rc == MZK_OK; i++) {
        ForceNtt f(cand[i]);
        const double t0 = now_ms();
        rc = mzk_ntt_dev(c, &job);
        if (rc == MZK_OK) rc = mzk_ctx_synchronize(c);
        tmin[i] = std::min(tmin[i], now_ms() - t0);
      }
    if (rc != MZK_OK) break;
    const NttTune pick = cand[std::min_element(tmin.begin(), tmin.end()) - tmin.begin()];
    TuneTable& tt = table();
    std::lock_guard<std::mutex> g(tt.mu);
    tt.ntt[L] = pick;
  }
  {
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(din);
    (void)hipFree(dout);
  }
  return rc;
}

int mzk_tune_export(char* buf, size_t len, size_t* needed) {
  std::string s;
  {
    TuneTable& t = table();
    std::lock_guard<std::mutex> g(t.mu);
    char line[128];
    for (int L = 0; L < TUNE_LOGS; L++) {
      const MsmTune& m = t.msm[L];
      if (m.path >= 0 || m.c || m.tpc_window || m.tpc_shared) {
        snprintf(line, sizeof line, "msm %d %d %d %u %u\n", L, m.path, m.c, m.tpc_window, m.tpc_shared);
        s += line;
      }
      const NttTune& e = t.ntt[L];
      if (e.l1 || e.groups) {
        snprintf(line, sizeof line, "ntt %d %u %u\n", L, e.l1, e.groups);
        s += line;
      }
    }
  }
  if (needed) *needed = s.size() + 1;
  MZK_REQUIRE(buf != nullptr || len == 0, "mzk_tune_export: null buffer");
  if (len == 0) return MZK_OK;
  MZK_REQUIRE(len > s.size(), "mzk_tune_export: %zu bytes needed", s.size() + 1);
  memcpy(buf, s.c_str(), s.size() + 1);
  return MZK_OK;
}

int mzk_tune_import(const char* text) {
  MZK_REQUIRE(text != nullptr, "mzk_tune_import: null text");
  MsmTune msm[TUNE_LOGS];
  NttTune ntt[TUNE_LOGS];
  {
    TuneTable& t = table();
    std::lock_guard<std::mutex> g(t.mu);
    std::copy(t.msm, t.msm + TUNE_LOGS, msm);
    std::copy(t.ntt, t.ntt + TUNE_LOGS, ntt);
  }
  int lineno = 0;
  const char* p = text;
  while (*p) {
    const char* e = strchr(p, '\n');
    const size_t n = e ? (size_t)(e - p) : strlen(p);
    std::string line(p, n);
    p += n + (e ? 1 : 0);
    lineno++;
    if (line.empty() || line[0] == '#') continue;
    int L, a, b;
    unsigned u, v;
    char extra;
    if (sscanf(line.c_str(), "msm %d %d %d %u %u %c", &L, &a, &b, &u, &v, &extra) == 5) {
      MZK_REQUIRE(L >= 0 && L < TUNE_LOGS && a >= -1 && a <= 1 && (b == 0 || b == 4 || b == 8 || b == 16) &&
                      u <= 4096 && v <= 4096,
                  "mzk_tune_import: line %d: bad msm entry", lineno);
      msm[L].path = (int8_t)a;
      msm[L].c = (int8_t)b;
      msm[L].tpc_window = (uint16_t)u;
      msm[L].tpc_shared = (uint16_t)v;
    } else if (sscanf(line.c_str(), "ntt %d %u %u %c", &L, &u, &v, &extra) == 3) {
      MZK_REQUIRE(L >= 0 && L < TUNE_LOGS && u <= 11 && v <= 65535, "mzk_tune_import: line %d: bad ntt entry",
                  lineno);
      ntt[L].l1 = (uint8_t)u;
      ntt[L].groups = (uint16_t)v;
    } else {
      MZK_REQUIRE(false, "mzk_tune_import: line %d: cannot parse '%s'", lineno, line.c_str());
    }
  }
  TuneTable& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  std::copy(msm, msm + TUNE_LOGS, t.msm);
  std::copy(ntt, ntt + TUNE_LOGS, t.ntt);
  return MZK_OK;
}

void mzk_tune_reset(void) {
  TuneTable& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  for (int L = 0; L < TUNE_LOGS; L++) {
    t.msm[L] = MsmTune();
    t.ntt[L] = NttTune();
  }
}

}  // extern "C"
