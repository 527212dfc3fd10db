// The NTT entry point that allocates nothing: mzk_ntt_dev_noalloc.
//
// It runs the same passes as mzk_ntt_dev (same kernels, same plans, hence the same
// values), but the scratch those passes need -- the pass-to-pass buffer of 2^log_n
// elements and the twiddle tables -- lives in `scratch`, a file-scope device array of
// fixed size, instead of context scratch grown with hipMalloc and cached plans.  It is for
// callers that cannot allocate: code under HIP graph capture, tight serving loops.
//
// `scratch` names a host-side stand-in in host code, not a device address.  Its device
// address is looked up once per device, with that device current, when the first
// context on the device is created (noalloc_setup, from mzk_ctx_create), and only that
// address is passed to kernels.  Each device holds its own copy of the array until the
// process ends; every call on a device shares it, so calls that overlap on one device
// race (callers order them, e.g. on one stream).  Nothing in it needs to start at zero:
// each call rewrites the tables it reads, and the first pass writes every buffer element
// the later passes read.
#include <mutex>

#include "ctx.hpp"
#include "host_ec.hpp"
#include "ntt.hpp"

namespace mzk {

static constexpr uint32_t NOALLOC_MAX_LOG = MZK_NTT_NOALLOC_MAX_LOG;
static constexpr size_t NOALLOC_TMP_BYTES = (size_t)32 << NOALLOC_MAX_LOG;   // 2^20 elements: 32 MiB
// twiddle records of any plan up to 2^20: at most 1280 sub-DFT + 2048 + 512 twist records
static constexpr size_t NOALLOC_TABLE_RECORDS = 4096;
static constexpr size_t NOALLOC_BYTES = NOALLOC_TMP_BYTES + NOALLOC_TABLE_RECORDS * NTT_TW29 * 4;
static constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

static std::mutex g_setup_mu;
static void* g_scratch_dev[NOALLOC_MAX_DEVICES];   // device address of `scratch`, per device

// Called with `device` current.  Looks up the device address of `scratch` once per device.
int noalloc_setup(int device) {
  MZK_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, "noalloc setup: device %d above %d", device,
              NOALLOC_MAX_DEVICES - 1);
  std::lock_guard<std::mutex> g(g_setup_mu);
  if (g_scratch_dev[device]) return MZK_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return MZK_ERR_HIP;
  }
  g_scratch_dev[device] = p;
  return MZK_OK;
}

static void* scratch_dev(int device) {
  std::lock_guard<std::mutex> g(g_setup_mu);
  return device >= 0 && device < NOALLOC_MAX_DEVICES ? g_scratch_dev[device] : nullptr;
}

}  // namespace mzk

using namespace mzk;

extern "C" {

int mzk_ntt_dev_noalloc(mzk_ctx* c, const mzk_ntt_job* job) {
  MZK_REQUIRE(c != nullptr && job != nullptr, "mzk_ntt_dev_noalloc: null argument");
  MZK_TRY(ntt_check_job(*job, 0));
  MZK_REQUIRE(job->log_n <= NOALLOC_MAX_LOG,
              "mzk_ntt_dev_noalloc: 2^%u elements above the maximum 2^%u of the allocation-free path "
              "(use mzk_ntt_dev)", job->log_n, NOALLOC_MAX_LOG);
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  uint8_t* base = (uint8_t*)scratch_dev(c->device);
  MZK_REQUIRE(base != nullptr, "mzk_ntt_dev_noalloc: device %d was not set up", c->device);
  uint32_t* tmp = (uint32_t*)base;
  uint32_t* tables = (uint32_t*)(base + NOALLOC_TMP_BYTES);
  const int inverse = (job->flags & MZK_NTT_INVERSE) != 0, coset = (job->flags & MZK_NTT_COSET) != 0;
  const uint64_t N = 1ull << job->log_n;
  host::Fr w = host::fr_root_of_unity(job->log_n);
  if (inverse) w = host::fr_inv(w);
  NttPlan plan;   // tables in `scratch`: plan.mem stays null, nothing to free
  MZK_TRY(plan.init_at(c->stream, job->log_n, w, tables, NOALLOC_TABLE_RECORDS));
  PassArgs pa[3];
  uint64_t groups[3];
  int np;
  if (inverse)
    np = ntt_build_passes(plan, (const uint32_t*)job->d_in, (uint32_t*)job->d_out, tmp, N, 0, coset ? 2 : 1,
                          host::fr_inv(host::fr_from_u64(N)), N, pa, groups);
  else
    np = ntt_build_passes(plan, (const uint32_t*)job->d_in, (uint32_t*)job->d_out, tmp, job->n_real, coset, 0,
                          host::fr_one(), N, pa, groups);
  for (int i = 0; i < np; i++) MZK_TRY(ntt_launch_pass_raw(c->stream, pa[i], groups[i]));
  return MZK_OK;
}

}  // extern "C"
