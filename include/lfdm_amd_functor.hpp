// SolveStep(t, functor) of the reference GPU solvers (dof18_halfImplicit_gpu.cuh:233-290 and the
// templated Integrate kernels :454-640): the functor `__device__ void operator()(double t,
// DriverInput* controls)` supplies the controls of every step of one kernel-sim-time chunk.
//
// Needs a HIP compiler (hipcc).  The functor is evaluated on the device at each step's time (t
// accumulated from the chunk start by `t += step`, as the integration itself does) into a
// [steps][3] control table that the integration kernel then reads for the whole fleet.
#pragma once
#include <hip/hip_runtime.h>

#include "lfdm_amd.hpp"

namespace lfdm_amd {
namespace detail {

template <class Func>
__global__ void eval_functor_controls(Func func, double t0, double step, int nsteps, double* ctrl) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double t = t0;
    for (int k = 0; k < nsteps; ++k) {
        DriverInput in;
        in.m_time = t;
        func(t, &in);
        ctrl[3 * k + 0] = in.m_steering;
        ctrl[3 * k + 1] = in.m_throttle;
        ctrl[3 * k + 2] = in.m_braking;
        t = t + step;
    }
}

// Device table reused across calls of one process (grown on demand).
struct FunctorTable {
    double* ptr = nullptr;
    int cap = 0;
    ~FunctorTable() {
        if (ptr) (void)hipFree(ptr);
    }
};

}  // namespace detail

// One chunk with controls from `func`; returns t + kernel_sim_time like the reference.  The solver
// must not have output enabled (the reference exits in that case; here it throws).
template <class Solver, class Func>
double SolveStep(Solver& solver, double t, Func func) {
    static detail::FunctorTable table;
    lfdm_solver* h = solver.handle();
    check(lfdm_synchronize(h));  // the previous chunk may still read the table
    int n = 0;
    check(lfdm_chunk_steps(h, t, &n));
    if (n > table.cap) {
        if (table.ptr) (void)hipFree(table.ptr);
        table.ptr = nullptr;
        if (hipMalloc((void**)&table.ptr, (size_t)n * 3 * sizeof(double)) != hipSuccess)
            throw Error(LFDM_E_HIP, "cannot allocate the functor control table");
        table.cap = n;
    }
    hipLaunchKernelGGL(detail::eval_functor_controls<Func>, dim3(1), dim3(64), 0, 0, func, t, solver.GetStep(), n,
                       table.ptr);
    if (hipDeviceSynchronize() != hipSuccess) throw Error(LFDM_E_HIP, "functor control evaluation failed");
    check(lfdm_set_controls_shared_device(h, n, table.ptr));
    check(lfdm_set_time(h, t));
    check(lfdm_advance(h, n));
    double ksim = 0.0;
    check(lfdm_get_kernel_sim_time(h, &ksim));
    return t + ksim;
}

}  // namespace lfdm_amd
