// Context, error reporting and the seeded input generator.
#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <chrono>

#include "ctx.hpp"
#include "field.hpp"

namespace mzk {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int Scratch::reserve(size_t bytes) {
  if (bytes <= cap) return MZK_OK;
  if (ptr) {
    MZK_HIP(hipFree(ptr));
    ptr = nullptr;
    cap = 0;
  }
  size_t want = bytes + bytes / 8;
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    return MZK_ERR_NOMEM;
  }
  cap = want;
  return MZK_OK;
}

void Scratch::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
}

int Pinned::reserve(size_t bytes) {
  if (bytes <= cap) return MZK_OK;
  if (ptr) {
    MZK_HIP(hipHostFree(ptr));
    ptr = nullptr;
    cap = 0;
  }
  const size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 16);
  const hipError_t e = hipHostMalloc(&ptr, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error("hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    return MZK_ERR_NOMEM;
  }
  cap = want;
  return MZK_OK;
}

void Pinned::release() {
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  cap = 0;
}

double host_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void KernelProfiler::begin(hipStream_t st, const char* name) {
  Pending p;
  p.name = name;
  if (hipEventCreate(&p.start) != hipSuccess || hipEventCreate(&p.stop) != hipSuccess) return;
  (void)hipEventRecord(p.start, st);
  pending.push_back(p);
}

void KernelProfiler::end(hipStream_t st) {
  if (!pending.empty()) (void)hipEventRecord(pending.back().stop, st);
}

int KernelProfiler::collect() {
  for (auto& p : pending) {
    MZK_HIP(hipEventSynchronize(p.stop));
    float ms = 0;
    MZK_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    auto& t = totals[p.name];
    t.first += ms;
    t.second += 1;
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  pending.clear();
  return MZK_OK;
}

// ---------------------------------------------------------------- launch graphs
// Destroys replaced graphs once nothing is queued on the stream (they may still be
// running until then); with too many waiting, one wait frees them all.
static int reap_retired(mzk_ctx* c, bool force) {
  GraphCache& G = c->graphs;
  if (G.retired.empty()) return MZK_OK;
  if (force || G.retired.size() > GraphCache::MAX_RETIRED) {
    MZK_HIP(hipStreamSynchronize(c->stream));
  } else if (hipStreamQuery(c->stream) != hipSuccess) {
    return MZK_OK;   // still busy (hipErrorNotReady): try again on a later call
  }
  for (hipGraphExec_t ex : G.retired) (void)hipGraphExecDestroy(ex);
  G.retired.clear();
  return MZK_OK;
}

int run_graph(mzk_ctx* c, const std::string& key, const std::function<int()>& launch) {
  GraphCache& G = c->graphs;
  if (!G.on || c->prof.on) return launch();
  MZK_TRY(reap_retired(c, false));
  const uint64_t now = ++G.tick;
  auto it = G.map.find(key);
  if (it == G.map.end()) {
    // a shape's first call runs eagerly: capture and instantiation (~16 ms for a 96-job
    // batch) pay only when the shape repeats, and one-off shapes do not evict cached ones
    if (G.seen.insert(key).second) {
      if (G.seen.size() > GraphCache::MAX_SEEN) {
        G.seen.clear();
        G.seen.insert(key);
      }
      return launch();
    }
    if (G.map.size() >= GraphCache::MAX_GRAPHS) {
      auto old = G.map.begin();
      for (auto e = G.map.begin(); e != G.map.end(); ++e)
        if (e->second.last_use < old->second.last_use) old = e;
      // every cached graph ran recently: the working set exceeds the cache, keep it
      if (now - old->second.last_use < GraphCache::STALE_TICKS) return launch();
      G.retired.push_back(old->second.exec);
      G.map.erase(old);
    }
    MZK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != MZK_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc != MZK_OK) return rc;
      set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
      return MZK_ERR_HIP;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      set_error("hipGraphInstantiate: %s", hipGetErrorString(ei));
      return MZK_ERR_HIP;
    }
    it = G.map.emplace(key, GraphCache::Entry{ex, 0}).first;
    G.captures++;
  } else {
    G.replays++;
  }
  it->second.last_use = now;
  MZK_HIP(hipGraphLaunch(it->second.exec, c->stream));
  return MZK_OK;
}

// Caller has synchronised the stream.
void free_graphs(mzk_ctx* c) {
  for (auto& kv : c->graphs.map) (void)hipGraphExecDestroy(kv.second.exec);
  for (hipGraphExec_t ex : c->graphs.retired) (void)hipGraphExecDestroy(ex);
  c->graphs.map.clear();
  c->graphs.retired.clear();
  c->graphs.seen.clear();
}

// ---------------------------------------------------------------- seeded stream
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint64_t splitmix64_host(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Fq::from_uniform_bytes (fq.rs:712-736): a0 * R2 + a1 * R3 as Montgomery products.
__global__ void k_fr_from_seed(uint64_t key, uint64_t offset, size_t n, uint32_t* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr uint32_t R2[8] = {0xf3f29c6du, 0xc999e990u, 0x87925c23u, 0x2b6cedcbu,
                              0x7254398fu, 0x05d31496u, 0x9f59ff11u, 0x0748d9d9u};
  constexpr uint32_t R3[8] = {0x439b73afu, 0xc62c1807u, 0x8cf06990u, 0x1b3e0d18u,
                              0xc7b5f418u, 0x73d13c71u, 0xc8db33e9u, 0x6e2a5bb9u};
  uint64_t idx = offset + i;
  fr_t a0, a1, r2, r3;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    uint64_t w0 = splitmix64(key + idx * 8 + j);
    uint64_t w1 = splitmix64(key + idx * 8 + 4 + j);
    a0.v[2 * j] = (uint32_t)w0;
    a0.v[2 * j + 1] = (uint32_t)(w0 >> 32);
    a1.v[2 * j] = (uint32_t)w1;
    a1.v[2 * j + 1] = (uint32_t)(w1 >> 32);
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    r2.v[j] = R2[j];
    r3.v[j] = R3[j];
  }
  fr_t r = add(mul(a0, r2), mul(a1, r3));
  store<FrP>(out + 8 * i, r);
}

}  // namespace mzk

using namespace mzk;

extern "C" {

const char* mzk_last_error(void) { return g_err; }

int mzk_version(void) { return 100; }

int mzk_ctx_create(int device, mzk_ctx** out) {
  MZK_REQUIRE(out != nullptr, "mzk_ctx_create: null out");
  int ndev = 0;
  MZK_HIP(hipGetDeviceCount(&ndev));
  MZK_REQUIRE(device >= 0 && device < ndev, "mzk_ctx_create: device %d out of range (%d devices)",
              device, ndev);
  MZK_HIP(hipSetDevice(device));
  MZK_TRY(noalloc_setup(device));
  mzk_ctx* c = new mzk_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_error("hipStreamCreateWithFlags: %s", hipGetErrorString(e));
    return MZK_ERR_HIP;
  }
  hipDeviceProp_t prop;
  MZK_HIP(hipGetDeviceProperties(&prop, device));
  c->num_cus = prop.multiProcessorCount;
  c->lds_per_block = prop.sharedMemPerBlock;
  *out = c;
  return MZK_OK;
}

void mzk_ctx_destroy(mzk_ctx* c) {
  if (!c) return;
  if (c->aux) mzk_ctx_destroy(c->aux);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_ntt_plans(c);
  free_graphs(c);
  for (auto& s : c->scratch) s.release();
  c->pinned.release();
  (void)hipStreamDestroy(c->stream);
  delete c;
}

void* mzk_ctx_stream(mzk_ctx* c) { return c ? (void*)c->stream : nullptr; }

int mzk_ctx_synchronize(mzk_ctx* c) {
  MZK_REQUIRE(c != nullptr, "null context");
  MZK_HIP(hipSetDevice(c->device));
  MZK_HIP(hipStreamSynchronize(c->stream));
  return MZK_OK;
}

int mzk_ctx_mem_info(mzk_ctx* c, size_t* free_bytes, size_t* total_bytes) {
  MZK_REQUIRE(c != nullptr && free_bytes != nullptr && total_bytes != nullptr, "mzk_ctx_mem_info: null argument");
  MZK_HIP(hipSetDevice(c->device));
  MZK_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return MZK_OK;
}

int mzk_prof_enable(mzk_ctx* c, int on) {
  MZK_REQUIRE(c != nullptr, "null context");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  MZK_TRY(c->prof.collect());
  c->prof.totals.clear();
  c->prof.on = on != 0;
  return MZK_OK;
}

int mzk_ctx_set_graphs(mzk_ctx* c, int on) {
  MZK_REQUIRE(c != nullptr, "null context");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  c->graphs.on = on != 0;
  if (!c->graphs.on) {
    MZK_HIP(hipStreamSynchronize(c->stream));
    free_graphs(c);
  }
  return MZK_OK;
}

int mzk_ctx_graph_stats(mzk_ctx* c, uint64_t out[3]) {
  MZK_REQUIRE(c != nullptr && out != nullptr, "mzk_ctx_graph_stats: null argument");
  std::lock_guard<std::mutex> g(c->mu);
  out[0] = c->graphs.captures;
  out[1] = c->graphs.replays;
  out[2] = c->graphs.map.size();
  return MZK_OK;
}

int mzk_prof_read(mzk_ctx* c, char* buf, size_t len) {
  MZK_REQUIRE(c != nullptr && buf != nullptr && len > 0, "mzk_prof_read: null argument");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  MZK_TRY(c->prof.collect());
  std::string out;
  char line[256];
  for (auto& kv : c->prof.totals) {
    snprintf(line, sizeof line, "%s %.6f %ld\n", kv.first.c_str(), kv.second.first, kv.second.second);
    out += line;
  }
  c->prof.totals.clear();
  MZK_REQUIRE(out.size() < len, "mzk_prof_read: buffer too small (%zu bytes needed)", out.size() + 1);
  memcpy(buf, out.c_str(), out.size() + 1);
  return MZK_OK;
}

int mzk_fr_from_seed_dev(mzk_ctx* c, uint64_t seed, uint64_t stream, uint64_t offset, size_t n,
                         void* d_out) {
  MZK_REQUIRE(c != nullptr, "null context");
  if (n == 0) return MZK_OK;
  MZK_REQUIRE(d_out != nullptr, "mzk_fr_from_seed_dev: null output");
  std::lock_guard<std::mutex> g(c->mu);
  MZK_HIP(hipSetDevice(c->device));
  uint64_t key = splitmix64_host(seed ^ (stream * 0xD1B54A32D192ED03ull));
  hipLaunchKernelGGL(k_fr_from_seed, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, key, offset, n,
                     (uint32_t*)d_out);
  MZK_HIP(hipGetLastError());
  return MZK_OK;
}

}  // extern "C"
