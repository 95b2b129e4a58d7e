// kelp_batch.cuh -- types shared by the batched-cloud kernels
// (kelp_batch.hip, one unit per field dtype and vertical coordinate) and
// their host side (kelp_batch_host.hip).
#pragma once

#include "kelp_internal.h"

namespace kelp {

// one cloud's step, written by batch_bracket_kernel for batch_step_kernel
struct BatchStep {
    int32_t active, slot_a, slot_b, slot_t;
    double alpha0;
    double alphas[4];
    uint64_t step_i;
};

struct SlotRecs {
    const void* rec[kMaxSlots];
    const float4* scale[kMaxSlots];  // row scales of float16 slots
    uint32_t row_len;                // records per row (the mesh's nodes)
};

// Launch batch_step_kernel<order, MAXV, T, SIGMA> over the particles (the
// mesh's element width) in workgroups of `block` threads.
template <typename T, bool SIGMA>
void launch_batch(hipStream_t stream, int order, int block, const DMesh& m, const SlotRecs& sr, const StepCfg& c,
                  const PState& ps, const kelp_cloud* clouds, const BatchStep* steps, int64_t n_clouds);

}  // namespace kelp
