// The context object behind the opaque mzk_ctx handle.
#pragma once
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "common.hpp"

namespace mzk {

// A growable device allocation reused across calls on one context.
struct Scratch {
  void* ptr = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes);
  void release();
};

// Pinned host memory reused across calls (hipHostMalloc): copies from and to it run
// asynchronously on the stream instead of through the runtime's pageable staging.  A user
// must wait for its copies before the next call may reuse the buffer.
struct Pinned {
  void* ptr = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes);
  void release();
};

enum ScratchSlot {
  SCR_NTT = 0,      // NTT ping-pong buffer
  SCR_MSM_A,        // MSM digits / sort keys
  SCR_MSM_B,        // MSM sorted point references
  SCR_MSM_C,        // MSM buckets
  SCR_MSM_D,        // MSM small reductions
  SCR_IN0,          // staging for host-pointer entry points
  SCR_IN1,
  SCR_PROVER,       // the native prover's per-proof vectors
  SCR_COUNT
};

uint64_t splitmix64_host(uint64_t x);

struct NttPlan;

// Per-kernel timing with HIP events on the context stream (off unless enabled).
struct KernelProfiler {
  bool on = false;
  struct Pending {
    const char* name;
    hipEvent_t start, stop;
  };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<double, long>> totals;   // name -> (ms, launches)
  void begin(hipStream_t st, const char* name);
  void end(hipStream_t st);
  int collect();   // waits for pending events and folds them into totals
  void host(const char* name, double ms) {   // host-side phase (waits, serial tails)
    if (!on) return;
    auto& t = totals[name];
    t.first += ms;
    t.second += 1;
  }
};

// Wall-clock milliseconds since an arbitrary origin (host-phase timing).
double host_ms();

// Executable graphs of launch sequences, keyed by everything the sequence's launches
// depend on (kernel, grid, block, LDS bytes and every argument, serialised): replaying
// a graph whose key matches is the same work as launching the sequence again.
// A shape is captured on its second call.  When the cache is full, the least recently
// used graph is replaced only if it has not run during the last STALE_TICKS calls;
// otherwise the new shape runs eagerly.  A working set that cycles through more shapes
// than the cache holds (up to STALE_TICKS of them) then keeps MAX_GRAPHS of them cached
// instead of recapturing on every call, and a workload that moves on to new shapes gets
// graphs for them after STALE_TICKS calls.  Replaced graphs are destroyed once the
// stream has gone idle, never with a wait.
struct GraphCache {
  static constexpr size_t MAX_GRAPHS = 48;
  static constexpr size_t MAX_SEEN = 4096;   // keys run once so far (cleared when full)
  static constexpr uint64_t STALE_TICKS = 8 * MAX_GRAPHS;
  static constexpr size_t MAX_RETIRED = 64;  // above this, one wait frees them all
  std::unordered_set<std::string> seen;
  struct Entry {
    hipGraphExec_t exec;
    uint64_t last_use;
  };
  std::map<std::string, Entry> map;
  std::vector<hipGraphExec_t> retired;
  uint64_t tick = 0;                         // run_graph calls so far
  uint64_t captures = 0, replays = 0;
  bool on = true;
};

// Byte string of plain values and pointers, for graph keys.
struct GraphKey {
  std::string s;
  template <class T>
  GraphKey& add(const T& v) {
    s.append(reinterpret_cast<const char*>(&v), sizeof v);
    return *this;
  }
};

}  // namespace mzk
struct mzk_ctx;
namespace mzk {
// Frees the context's cached NTT plans (ntt.hip).
void free_ntt_plans(mzk_ctx* c);
// Runs `launch` (kernel launches on the context stream and nothing else: no waits, no
// copies, no allocations) through the context's graph cache: a cached graph under `key`
// is replayed, otherwise the launches are captured into one, instantiated and launched.
// With graphs off or the kernel profiler on, `launch` simply runs.  Caller holds c->mu.
int run_graph(mzk_ctx* c, const std::string& key, const std::function<int()>& launch);
void free_graphs(mzk_ctx* c);
// Per-device set-up of the allocation-free NTT's device array (noalloc.hip); called with
// the device current.
int noalloc_setup(int device);

}  // namespace mzk

// Launches a kernel on the context stream, bracketed by profiler events when on.
#define MZK_LAUNCH(ctx, name, kernel, grid, block, lds, ...)                        \
  do {                                                                             \
    if ((ctx)->prof.on) (ctx)->prof.begin((ctx)->stream, name);                     \
    hipLaunchKernelGGL(kernel, grid, block, lds, (ctx)->stream, __VA_ARGS__);       \
    if ((ctx)->prof.on) (ctx)->prof.end((ctx)->stream);                             \
  } while (0)

struct mzk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::mutex prove_mu;   // held for a whole mzk_plonk_prove call (its arena spans many locked calls)
  int num_cus = 0;
  size_t lds_per_block = 0;
  mzk::Scratch scratch[mzk::SCR_COUNT];
  mzk::Pinned pinned;   // MSM descriptors in, channel sums out (the MSM waits before returning)
  mzk::KernelProfiler prof;
  uint64_t last_msm[6] = {0, 0, 0, 0, 0, 0};   // c, W, sets, K, M, n of the last MSM
  std::map<std::string, mzk::NttPlan*> ntt_plans;   // (log_n, split, root) -> resident twiddle tables
  mzk::GraphCache graphs;
  // a second context (own stream and scratch) on the same device for work the prover can
  // overlap with this context's stream; created on first use under prove_mu
  mzk_ctx* aux = nullptr;
};
