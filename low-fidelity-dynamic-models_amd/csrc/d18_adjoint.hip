// Differentiation form of the 18-DOF solver.
//  - adj_forward: integrates with per-step controls, records the (x, y) trajectory and stores the
//    carried state every K steps.
//  - adj_recompute: one lane per (segment of a group, vehicle slot); reruns the segment from its
//    checkpoint into the tape.  Segments are independent, so a group of G segments gives G times
//    the wavefronts of the fleet and the recompute runs at full occupancy even for fleets that fill
//    only one wavefront per SIMD.  The tape is blocked by wavefront (one contiguous 19 KB run per
//    wavefront and step) and the per-vehicle constants sit in LDS.
//  - step_jacobian: one lane per slot; the step's Jacobian at the current record, one step adjoint
//    sweep per row.
//  - adj_reverse: one lane per slot; walks the group's segments backwards through the step adjoint
//    of d18_adjoint.h.  The step adjoint needs the whole register file of a SIMD, so the per-vehicle
//    constants and their adjoints live in LDS (23 + 21 doubles per lane; both structs have an odd
//    number of doubles, so a wavefront's 64-bit LDS accesses at that stride hit distinct banks).
//
// One vehicle per lane as in the integration kernels; every lane of a wave runs the same step
// counts, so the segment loops are wave-uniform.
#include <hip/hip_runtime.h>

#include "d18_adjoint_lanes.h"

namespace lfdm {

using namespace dev;

namespace {

constexpr int kThreads = 256;

// Waves per SIMD of the segment recompute (three: its 164 VGPRs fit without spills, and the group's
// (segment, vehicle) lanes fill them).
#ifndef LFDM_ADJ_RECOMPUTE_WAVES
#define LFDM_ADJ_RECOMPUTE_WAVES 3
#endif

static_assert(sizeof(K18) % 16 == 8 && sizeof(K18b) % 16 == 8, "odd number of doubles per lane in LDS");

template <int W>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(W, W))) void adj_forward(AdjLaunch L) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = uniform(slot >> 6);
    if (wave * kWave >= L.S) return;
    adj_forward_lane(L, slot, pset_at(L.sets, uniform(L.wave_set[wave])));
}

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(LFDM_ADJ_RECOMPUTE_WAVES,
                                                                            LFDM_ADJ_RECOMPUTE_WAVES))) void
adj_recompute(AdjLaunch L) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    // S is a multiple of the wavefront size, so q and the slot's wave are wave-uniform
    const int q = uniform((int)(id / (size_t)L.S));
    const int slot = (int)(id - (size_t)q * L.S);
    if (q >= L.seg_hi - L.seg_lo) return;
    __shared__ K18 ks[kThreads];
    adj_recompute_lane(L, q, slot, pset_at(L.sets, uniform(L.wave_set[slot >> 6])), ks[threadIdx.x]);
}

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) void adj_reverse(AdjLaunch L) {
    __shared__ K18 ks[kThreads];
    __shared__ K18b kbs[kThreads];
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = uniform(slot >> 6);
    if (wave * kWave >= L.S) return;
    adj_reverse_lane(L, slot, pset_at(L.sets, uniform(L.wave_set[wave])), ks[threadIdx.x], kbs[threadIdx.x]);
}

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) void step_jacobian(AdjLaunch L) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = uniform(slot >> 6);
    if (wave * kWave >= L.S) return;
    jac_lane(L, slot, pset_at(L.sets, uniform(L.wave_set[wave])));
}

}  // namespace

hipError_t launch_step_jacobian(const AdjLaunch& L, hipStream_t stream) {
    if (L.S <= 0) return hipSuccess;
    hipLaunchKernelGGL(step_jacobian, (L.S + kThreads - 1) / kThreads, kThreads, 0, stream, L);
    return hipGetLastError();
}

hipError_t launch_adj_forward(const AdjLaunch& L, int waves_per_simd, hipStream_t stream) {
    if (L.S <= 0 || L.nsteps <= 0) return hipSuccess;
    if (waves_per_simd == 1)
        hipLaunchKernelGGL(adj_forward<1>, (L.S + kThreads - 1) / kThreads, kThreads, 0, stream, L);
    else
        hipLaunchKernelGGL(adj_forward<2>, (L.S + kThreads - 1) / kThreads, kThreads, 0, stream, L);
    return hipGetLastError();
}

hipError_t launch_adj_recompute(const AdjLaunch& L, hipStream_t stream) {
    const size_t lanes = (size_t)(L.seg_hi - L.seg_lo) * (size_t)L.S;
    if (L.S <= 0 || lanes == 0) return hipSuccess;
    hipLaunchKernelGGL(adj_recompute, (unsigned)((lanes + kThreads - 1) / kThreads), kThreads, 0, stream, L);
    return hipGetLastError();
}

hipError_t launch_adj_reverse(const AdjLaunch& L, hipStream_t stream) {
    if (L.S <= 0 || L.seg_hi <= L.seg_lo) return hipSuccess;
    hipLaunchKernelGGL(adj_reverse, (L.S + kThreads - 1) / kThreads, kThreads, 0, stream, L);
    return hipGetLastError();
}

}  // namespace lfdm
