// Shared host-side plumbing for the HIP library: status codes, the thread-local
// error message behind mzk_last_error(), and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>

#include "../../include/mzk_gpu.h"

namespace mzk {

void set_error(const char* fmt, ...);

#define MZK_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::mzk::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                       \
      return MZK_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define MZK_REQUIRE(cond, ...)        \
  do {                                \
    if (!(cond)) {                    \
      ::mzk::set_error(__VA_ARGS__);  \
      return MZK_ERR_INVALID;         \
    }                                 \
  } while (0)

#define MZK_TRY(expr)         \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != MZK_OK) return rc_; \
  } while (0)

// Blocks needed to give every one of n items its own thread (callers bound n so
// that this stays below 2^31).
inline unsigned grid_for(size_t n, unsigned block) {
  size_t g = (n + block - 1) / block;
  return (unsigned)(g == 0 ? 1 : g);
}

}  // namespace mzk
