// Differentiation form of the 18-DOF solver: a forward kernel that integrates with per-step
// controls, records the (x, y) trajectory and stores the carried state every K steps, and a
// reverse kernel that walks the segments backwards (recompute the segment's states from its
// checkpoint into a per-vehicle buffer, then run the step adjoint of d18_adjoint.h step by step).
//
// One vehicle per lane as in the integration kernels; every lane of a wave runs the same step
// counts, so the segment loops are wave-uniform.
#include <hip/hip_runtime.h>

#include "d18_adjoint_lanes.h"

// Waves per SIMD of the reverse kernel (its step adjoint needs far more registers than the forward).
#ifndef LFDM_ADJ_BWD_WAVES
#define LFDM_ADJ_BWD_WAVES 1
#endif

namespace lfdm {

using namespace dev;

namespace {

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void adj_forward(AdjLaunch L) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = uniform(slot >> 6);
    if (wave * kWave >= L.S) return;
    adj_forward_lane(L, slot, L.sets[uniform(L.wave_set[wave])]);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LFDM_ADJ_BWD_WAVES, LFDM_ADJ_BWD_WAVES))) void
adj_backward(AdjLaunch L) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = uniform(slot >> 6);
    if (wave * kWave >= L.S) return;
    adj_backward_lane(L, slot, L.sets[uniform(L.wave_set[wave])]);
}

}  // namespace

hipError_t launch_adj_forward(const AdjLaunch& L, hipStream_t stream) {
    if (L.S <= 0 || L.nsteps <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_forward, (L.S + 255) / 256, 256, 0, stream, L);
    return hipGetLastError();
}

hipError_t launch_adj_backward(const AdjLaunch& L, hipStream_t stream) {
    if (L.S <= 0 || L.nsteps <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_backward, (L.S + 255) / 256, 256, 0, stream, L);
    return hipGetLastError();
}

}  // namespace lfdm
