// NTT plans (twiddle tables resident on the device) and the pass driver.
#pragma once
#include "ctx.hpp"
#include "host_ec.hpp"

namespace mzk {

// Sub-DFTs staged in LDS hold at most 2^NTT_MAX_LOG_L elements; a transform takes one
// pass up to that size, two passes (Bailey four-step) up to its square and three above.
static constexpr int NTT_MAX_LOG_L = 11;
static constexpr int NTT_MAX_LOG_N = 32;   // the field's two-adicity S (fq.rs:43-48)
static constexpr int NTT_TW29 = 12;        // words per radix-2^29 table record (9 used; 16-byte aligned)

// Tables for one transform size and root, in radix-2^29 records (values below r):
//   tw[p]    = omega_{L_p}^i, i < max(L_p / 2, 1), for the sub-DFT size L_p of pass p;
//   twist_lo = omega_N^i, i < 2^11, and twist_hi = omega_N^(i 2^11), i < N / 2^11
//              (the four-step twists, two or three passes only).
struct NttPlan {
  uint32_t log_n = 0;
  uint32_t np = 0;                      // passes: 1, 2 or 3
  uint32_t lp[3] = {0, 0, 0};           // log2 sub-DFT size of each pass, in pass order
  uint32_t* tw[3] = {nullptr, nullptr, nullptr};
  uint32_t* twist_lo = nullptr;
  uint32_t* twist_hi = nullptr;
  void* mem = nullptr;                  // owned table memory (null when the caller provides it)
  NttPlan() = default;
  NttPlan(const NttPlan&) = delete;
  NttPlan& operator=(const NttPlan&) = delete;
  ~NttPlan();
  // The pass split of a 2^log_n transform (the tuner's split for two passes, when set).
  static void split(uint32_t log_n, uint32_t* np, uint32_t lp[3]);
  // Table records the plan of this split needs.
  static size_t records(uint32_t log_n, uint32_t np, const uint32_t lp[3]);
  // Allocates the tables and fills them on stream st.
  int init(hipStream_t st, uint32_t log_n, const host::Fr& omega);
  // Fills the tables into caller memory of cap_records records on stream st, allocating
  // nothing; the memory must stay valid while transforms of this plan are queued.
  int init_at(hipStream_t st, uint32_t log_n, const host::Fr& omega, uint32_t* tables, size_t cap_records);
};

// One pass of a transform: sub-DFT b's element j is read at
//   (b & in_mask) in_lo + (b >> in_shift) in_hi + j in_j,
// its output k written at the same form with the out_ fields, and (when do_twist) scaled
// by omega_N^e, e = (b & tw_mask) ((b >> tw_shift) + (k << tw_kshift)).
struct FrWords {
  uint32_t v[8];
};
struct PassArgs {
  const uint32_t* src;
  uint32_t* dst;
  const uint32_t* tw;        // omega_L^i, i < L/2 (records)
  const uint32_t* twist_lo;
  const uint32_t* twist_hi;
  uint64_t in_mask, in_lo, in_hi, in_j;
  uint64_t out_mask, out_lo, out_hi, out_k;
  uint64_t tw_mask;
  uint64_t n_real;           // input indices >= n_real read as zero
  uint32_t in_shift, out_shift, tw_shift, tw_kshift;
  uint32_t log_l;            // sub-DFT size L = 2^log_l
  uint32_t C;                // sub-DFTs per workgroup
  int do_twist;
  int in_coset;              // input index i times [1, zeta, zeta^2][i mod 3]
  int out_mode;              // 0 none, 1 times scale, 2 times scale*[1, zeta^2, zeta][k mod 3]
  int pairs;                 // stages two at a time in registers (large launches)
  FrWords zeta, zeta2;
  FrWords post[3];           // scale * pattern, indexed by output index mod 3
};

// The passes of one transform of src (zero beyond n_real; zeta-coset distributed on input
// when in_coset) into dst, through tmp (N elements; unused for one pass); out_mode 0:
// plain, 1: times scale, 2: times scale * [1, zeta^2, zeta][i mod 3].  launch_elems is the
// size of the launch the passes join (it picks the stage pairing).  Returns the pass
// count; groups[p] = workgroups of pass p.
int ntt_build_passes(const NttPlan& plan, const uint32_t* src, uint32_t* dst, uint32_t* tmp, uint64_t n_real,
                     int in_coset, int out_mode, const host::Fr& scale, uint64_t launch_elems, PassArgs pa[3],
                     uint64_t groups[3]);
// Launches one pass on stream st without the context's profiler or graphs.
int ntt_launch_pass_raw(hipStream_t st, const PassArgs& a, uint64_t groups);

// Validates one transform job (sizes, flags, pointers); job index i names it in the error.
int ntt_check_job(const mzk_ntt_job& j, size_t i);

// DFT of src into dst as ntt_build_passes describes; src may equal dst.  Enqueued on the
// context stream (scratch tmp from the context); caller holds the context lock.
int ntt_run(mzk_ctx* c, const NttPlan& plan, const uint32_t* src, uint32_t* dst, uint64_t n_real, int in_coset,
            int out_mode, const host::Fr& scale);

}  // namespace mzk
