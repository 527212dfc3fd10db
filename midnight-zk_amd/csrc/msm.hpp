// MSM planning and the shared device pipeline (msm.hip), used by the per-call
// entry points and by the resident-SRS commit path.
#pragma once
#include <functional>
#include <vector>

#include "ctx.hpp"

namespace mzk {

struct MsmPlan {
  int c = 16;          // window bits
  int W = 16;          // windows = 255 / c + 1 (signed digits never carry past them)
  uint32_t B = 0;      // buckets per set = 2^(c-1) (digit magnitudes 1..2^(c-1))
  int per_window = 1;  // 1: one bucket set per window; 0: one set over window-shifted bases
  int sets = 16;
  uint32_t NB = 0;     // sets * B
  uint32_t K = 32;     // sorted references per accumulation thread
};

MsmPlan msm_plan(size_t n, int per_window, int num_cus, int c_override);

// One MSM of a batch: n device scalars; its point references start at ref_offset.
struct MsmJob {
  const void* d_coeffs = nullptr;
  size_t n = 0;
  uint32_t ref_offset = 0;
};

// Runs all jobs through one pipeline (one sort, one accumulation, one reduction):
// outs[18 j ..] receives job j's Jacobian result.  records: d_bases holds 128-byte
// accumulation records (a resident SRS's shifted tables) instead of blst affine points.
// on_enqueued (optional) runs once the device sequence is on the context stream and
// before the host waits for it, with the context lock held: a caller can order other
// streams after the MSM's kernels (an event recorded there fires when they end).
int msm_run_jobs(mzk_ctx* c, const MsmPlan& p, const std::vector<MsmJob>& jobs, const uint32_t* d_bases,
                 uint32_t ref_stride, uint64_t* outs, bool records = false,
                 const std::function<int()>* on_enqueued = nullptr);

// Runs the pipeline on device scalars (n x 8 u32 Montgomery) and device bases.
// Reference r of scalar i in window w is ref_offset + i (+ w * ref_stride when
// per_window == 0); out is the Jacobian result.
int msm_run(mzk_ctx* c, const MsmPlan& p, const uint32_t* d_scalars, size_t n, const uint32_t* d_bases,
            uint32_t ref_stride, uint32_t ref_offset, bool records, uint64_t out[18]);

}  // namespace mzk
