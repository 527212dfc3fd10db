// The context object behind the opaque mzk_ctx handle.
#pragma once
#include <mutex>

#include "common.hpp"

namespace mzk {

// A growable device allocation reused across calls on one context.
struct Scratch {
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
  SCR_COUNT
};

uint64_t splitmix64_host(uint64_t x);

}  // namespace mzk

struct mzk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  int num_cus = 0;
  size_t lds_per_block = 0;
  mzk::Scratch scratch[mzk::SCR_COUNT];
  void* host_staging = nullptr;
};
