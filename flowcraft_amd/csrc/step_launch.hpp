// Launcher of the fused step kernel, one explicit instantiation per (Q, storage) in
// step_inst.hip.  Keeping the KStep instantiations out of fc_lbm.hip lets the build
// compile them as separate translation units in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_kernels.hpp"

namespace fc {

// Enqueues one KStep over `tiles` workgroups on `stream`: cpt cells per thread (1 or 2),
// the periodic-seam variant when `periodic`, cached population stores when `resident`
// (one cell per thread only).
template <int Q, class S, class Lay>
hipError_t LaunchStep(const StepArgs& a, unsigned tiles, int cpt, bool periodic, bool resident, hipStream_t stream);

}  // namespace fc
