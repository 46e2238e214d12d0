// SolveStep(t, functor) of the reference GPU solvers (dof18_halfImplicit_gpu.cuh:233-290 and the
// templated Integrate kernels :454-640; the 11/24-DOF twins alike): the functor
// `__device__ void operator()(double t, DriverInput* controls)` supplies the controls of every vehicle at
// every step of one kernel-sim-time chunk, called on the vehicle's own lane with the step's time, so it may
// read blockIdx / threadIdx (slot = blockIdx.x * blockDim.x + threadIdx.x; with one Construct group the slot
// is the vehicle index, as in the reference) or per-vehicle device data.  A functor may also take the
// vehicle index itself: `operator()(double t, DriverInput* controls, unsigned vehicle)`, which stays the
// vehicle's index whatever the slot layout (several Construct groups, horizon layout).
//
// Needs a HIP compiler (hipcc): the library's integration kernel is instantiated here on the functor's type
// and launched on the handle's stream with the launch record the library prepares
// (lfdm_functor_chunk_begin / _done); nothing is allocated and nothing synchronizes.  Compile with
// -ffp-contract=off, as the library is, so the step keeps the reference's operation order, and with
// -mllvm -disable-machine-licm, as the library is, so the step's constants are not held in vector
// registers across the step loop (without it the kernel is correct but spills more).
#pragma once
#include <hip/hip_runtime.h>

#include "../low-fidelity-dynamic-models_amd/csrc/integrate.h"
#include "lfdm_amd.hpp"

namespace lfdm_amd {
namespace detail {

template <class Solver>
struct ModelOf;
template <>
struct ModelOf<d18SolverHalfImplicitGPU> {
    using type = lfdm::kern::M18;
};
template <>
struct ModelOf<d11SolverHalfImplicitGPU> {
    using type = lfdm::kern::M11;
};
template <>
struct ModelOf<d24SolverHalfImplicitGPU> {
    using type = lfdm::kern::M24;
};

// Calls the functor with the vehicle index where it takes one (overload resolution in device code, where a
// __device__ call operator is visible; the int argument prefers this form).
template <class Func>
__device__ auto call_functor(Func& f, double t, DriverInput* in, unsigned vehicle, int)
    -> decltype(f(t, in, vehicle), void()) {
    f(t, in, vehicle);
}
template <class Func>
__device__ void call_functor(Func& f, double t, DriverInput* in, unsigned, long) {
    f(t, in);
}

// The kernel's control source around the user's functor.
template <class Func>
struct FunctorControls {
    Func func;
    __device__ void eval(double t, int vehicle, double& s, double& th, double& b) {
        DriverInput in;
        in.m_time = t;
        call_functor(func, t, &in, (unsigned)vehicle, 0);
        s = in.m_steering;
        th = in.m_throttle;
        b = in.m_braking;
    }
};

template <class M, bool NR, int W, bool STRICT, class F>
hipError_t launch_functor_at(const lfdm::Launch& L, int block, int grid, hipStream_t st, const F& f) {
    hipLaunchKernelGGL((lfdm::kern::integrate<M, NR, lfdm::kCtrlFunctor, W, STRICT, F>), dim3(grid), dim3(block),
                       lfdm::kern::k_lds_bytes<M>(block), st, L, f);
    return hipGetLastError();
}

// The library's choice of kernel (lfdm::kern::launch_model): the strict step at the model's default
// occupancy; the fast step at the occupancy the launch names (L.waves: the wavefronts per SIMD of a
// fleet that fills fewer than the default), else at the model's default.
template <class M, bool NR, class F>
hipError_t launch_functor_tire(const lfdm::Launch& L, int block, int grid, hipStream_t st, const F& f) {
    if (L.strict) return launch_functor_at<M, NR, M::kWaves, true>(L, block, grid, st, f);
    if constexpr (M::kWaves > 3) {
        if (L.waves == 3) return launch_functor_at<M, NR, 3, false>(L, block, grid, st, f);
    }
    if constexpr (M::kWaves > 2) {
        if (L.waves == 2) return launch_functor_at<M, NR, 2, false>(L, block, grid, st, f);
    }
    if constexpr (M::kWaves > 1) {
        if (L.waves == 1) return launch_functor_at<M, NR, 1, false>(L, block, grid, st, f);
    }
    return launch_functor_at<M, NR, M::kWaves, false>(L, block, grid, st, f);
}

}  // namespace detail

// One chunk with controls from `func`; returns t + kernel_sim_time like the reference.  The solver must
// not have CSV output set (the reference exits in that case; here it throws).
template <class Solver, class Func>
double SolveStep(Solver& solver, double t, Func func) {
    using M = typename detail::ModelOf<Solver>::type;
    lfdm_solver* h = solver.handle();
    int n = 0;
    check(lfdm_chunk_steps(h, t, &n));
    lfdm::Launch L;
    int block = 0, grid = 0;
    void* stream = nullptr;
    check(lfdm_functor_chunk_begin(h, t, n, &L, (int)sizeof(L), &block, &grid, &stream));
    if (n > 0 && L.S > 0) {
        const detail::FunctorControls<Func> f{func};
        const hipStream_t st = static_cast<hipStream_t>(stream);
        const hipError_t e = L.nr ? detail::launch_functor_tire<M, true>(L, block, grid, st, f)
                                  : detail::launch_functor_tire<M, false>(L, block, grid, st, f);
        if (e != hipSuccess) throw Error(LFDM_E_HIP, std::string("functor integration launch: ") + hipGetErrorString(e));
    }
    check(lfdm_functor_chunk_done(h, n));
    double ksim = 0.0;
    check(lfdm_get_kernel_sim_time(h, &ksim));
    return t + ksim;
}

}  // namespace lfdm_amd
