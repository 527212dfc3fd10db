// NTT plans (twiddle tables resident on the device) and the two-pass driver.
#pragma once
#include "ctx.hpp"
#include "host_ec.hpp"

namespace mzk {

// Tables for one transform size and root: sub-DFT twiddles for the two passes and
// the two-level twist table omega_N^(lo + 2^11 hi).
struct NttPlan {
  uint32_t log_n = 0, l1 = 0, l2 = 0;
  void* mem = nullptr;
  // tables as radix-2^29 records (values below r, 12 words each)
  uint32_t* tw1 = nullptr;       // omega_{N1}^i, i < N1/2   (pass 2)
  uint32_t* tw2 = nullptr;       // omega_{N2}^i, i < N2/2   (pass 1)
  uint32_t* twist_lo = nullptr;  // omega_N^i, i < 2^min(11, log_n)
  uint32_t* twist_hi = nullptr;  // omega_N^(i 2^11)
  NttPlan() = default;
  NttPlan(const NttPlan&) = delete;
  NttPlan& operator=(const NttPlan&) = delete;
  ~NttPlan();
  int init(hipStream_t st, uint32_t log_n, const host::Fr& omega);
};

// DFT of src (zero beyond n_real; optionally zeta-coset distributed on input) into
// dst; out_mode 0: plain, 1: times scale, 2: times scale * [1, zeta^2, zeta][i mod 3].
// src may equal dst.  Enqueued on the context stream; caller holds the context lock.
int ntt_run(mzk_ctx* c, const NttPlan& plan, const uint32_t* src, uint32_t* dst, uint64_t n_real, int in_coset,
            int out_mode, const host::Fr& scale);

}  // namespace mzk
