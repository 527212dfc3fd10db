// Plan table of the autotuner (tune.hip): per log2 of the problem size, the MSM path,
// window bits and accumulation threads, and the NTT four-step split and workgroup
// target.  Zero fields mean "the built-in model"; the planners in msm.hip, srs.hip and
// ntt.hip read the table through these getters.  The table is process-wide; the
// tuner's candidate runs override it on the calling thread only.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace mzk {

struct MsmTune {
  int8_t path = -1;          // resident-SRS single MSM: -1 model, 0 per-window buckets, 1 shared buckets
  int8_t c = 0;              // per-window window bits (4, 8 or 16)
  uint16_t tpc_window = 0;   // accumulation threads per CU, per-window path
  uint16_t tpc_shared = 0;   // accumulation threads per CU, shared-bucket path
};

struct NttTune {
  uint8_t l1 = 0;            // log2 of the second pass's sub-DFT size (rows); l2 = log_n - l1
  uint16_t groups = 0;       // workgroups each pass aims for (sub-DFTs per workgroup follow)
};

MsmTune tune_msm_get(size_t n);
NttTune tune_ntt_get(uint32_t log_n);
void tune_force_msm(const MsmTune* t);   // null: back to the table
void tune_force_ntt(const NttTune* t);

}  // namespace mzk
