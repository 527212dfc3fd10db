// The resident-SRS object behind the opaque mzk_srs handle.
#pragma once
#include <functional>

#include "ctx.hpp"

#define MZK_SRS_REC_WORDS 32

struct mzk_srs {
  mzk_ctx* ctx = nullptr;    // the creating context (used by mzk_srs_download); never read at destroy
  int device = 0;            // the device holding the bases
  size_t n = 0;
  void* bases = nullptr;     // n x 24 u32 affine (blst_p1_affine layout)
  void* shifted = nullptr;   // optional window-shifted copies 2^(c w) g_i, as accumulation
                             // records of MZK_SRS_REC_WORDS words (x, y radix-2^28, identity flag)
  int shifted_c = 0;
};

namespace mzk {
// XYZZ (n x 48 u32) -> affine (n x 24 u32) on the context stream.
// XYZZ (blst layout) -> affine; prefix: n * 48 bytes of scratch, or null for the context's
int batch_normalize(mzk_ctx* c, const uint32_t* d_xyzz, size_t n, uint32_t* d_affine, uint32_t* prefix = nullptr);
// For the SRS builders outside srs.hip (called with the context lock held and the
// device set): a new SRS of n bases, and its shifted tables once the bases are written.
int srs_alloc_bases(mzk_ctx* c, size_t n, mzk_srs** out);
int srs_finish_bases(mzk_ctx* c, mzk_srs* s);
// mzk_msm_srs_batch_dev with the msm_run_jobs hook: on_enqueued runs once the batch's
// device sequence is enqueued, before the host waits for it.
int srs_commit_batch(mzk_ctx* c, const mzk_srs* s, const mzk_msm_job* jobs, size_t njobs, uint64_t* outs,
                     const std::function<int()>* on_enqueued);
}  // namespace mzk
