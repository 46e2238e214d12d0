// The reference's GPU known-answer tests and functor demo, written against the C++ drop-in classes
// (include/lfdm_amd.hpp, include/lfdm_amd_functor.hpp) exactly as they are written against
// d11/d18/d24SolverHalfImplicitGPU: wheeled_vehicle_models/tests/{11,18,24}dof_gpu.cu and
// 18dof-gpu/demos/HMMWV/demo_hmmwv_controlsFunctor_18gpu.cu.
//
// Usage: dropin_test <data dir (tests/golden/data)>.  Prints one line per check and exits non-zero
// on the first failure.
#include <cmath>
#include <cstdio>
#include <string>

#include "lfdm_amd_functor.hpp"

using namespace lfdm_amd;

namespace {

int failures = 0;

void expect_near(const char* what, double expected, double got, double tol) {
    const bool ok = std::fabs(expected - got) <= tol;
    std::printf("%s %s: expected %.6g, got %.6g\n", ok ? "PASS" : "FAIL", what, expected, got);
    std::fflush(stdout);
    if (!ok) ++failures;
}

// Throttle from t = 0.5 s, a kernel-sim-time boundary (chunks of 0.1 s start at 0, 0.1, ... 0.4,
// 0.5 exactly), so the per-step functor and the per-chunk scalar SolveStep see the same controls.
struct MyFunctor {
    __device__ void operator()(double t, DriverInput* controls) {
        if (t < 0.5) {
            controls->m_steering = 0.0;
            controls->m_throttle = 0.0;
            controls->m_braking = 0.0;
        } else {
            controls->m_steering = 0.0;
            controls->m_throttle = 0.5;
            controls->m_braking = 0.0;
        }
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <data dir>\n", argv[0]);
        return 2;
    }
    const std::string data = argv[1];
    const std::string in = data + "/input/";
    try {
        {  // tests/18dof_gpu.cu:13-56 (50 vehicles, acc3, 10 s)
            const std::string d = data + "/18dof/HMMWV/";
            d18SolverHalfImplicitGPU solver(50);
            solver.Construct(d + "vehicle.json", d + "tmeasy.json", 50, in + "acc3.txt");
            solver.SetThreadsPerBlock(32);
            solver.SetTimeStep(1e-3);
            d18GPU::VehicleState veh_st;
            d18GPU::TMeasyState tirelf_st, tirerf_st, tirelr_st, tirerr_st;
            solver.Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, 50);
            solver.SetEndTime(10.0);
            solver.Solve();
            d18GPU::SimState sim_state = solver.GetSimState(49);
            expect_near("18-DOF acc3 x", 88.5335, sim_state._veh_state._x, 1.0);
        }
        {  // tests/18dof_gpu.cu:58-99 (two Construct groups, 22 s)
            const std::string d = data + "/18dof/HMMWV/";
            d18SolverHalfImplicitGPU solver(1000);
            solver.Construct(d + "vehicle.json", d + "tmeasy.json", 500, in + "acc3.txt");
            solver.Construct(d + "vehicle.json", d + "tmeasy.json", 500, in + "double_lane4.txt");
            solver.SetThreadsPerBlock(32);
            solver.SetTimeStep(1e-3);
            d18GPU::VehicleState veh_st;
            d18GPU::TMeasyState tirelf_st, tirerf_st, tirelr_st, tirerr_st;
            solver.Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, 1000);
            solver.SetEndTime(22.0);
            solver.Solve();
            expect_near("18-DOF fleet vehicle 499 x", 241.112, solver.GetSimState(499)._veh_state._x, 1.0);
            expect_near("18-DOF fleet vehicle 999 x", 106.548, solver.GetSimState(999)._veh_state._x, 1.0);
        }
        {  // tests/11dof_gpu.cu:13-42
            const std::string d = data + "/11dof/HMMWV/";
            d11SolverHalfImplicitGPU solver(50);
            solver.Construct(d + "vehicle.json", d + "tmeasy.json", 50, in + "acc3.txt");
            solver.SetThreadsPerBlock(32);
            solver.SetTimeStep(1e-3);
            d11GPU::VehicleState veh_st;
            d11GPU::TMeasyState tiref_st, tirer_st;
            solver.Initialize(veh_st, tiref_st, tirer_st, 50);
            solver.SetEndTime(10.0);
            solver.Solve();
            expect_near("11-DOF acc3 x", 70.6758, solver.GetSimState(49)._veh_state._x, 1.0);
        }
        {  // tests/24dof_gpu.cu:14-61
            const std::string d = data + "/24dof/HMMWV/";
            d24SolverHalfImplicitGPU solver(50);
            solver.Construct(d + "vehicle.json", d + "tmeasy.json", d + "suspension.json", 50, in + "acc3.txt");
            solver.SetThreadsPerBlock(32);
            solver.SetTimeStep(1e-3);
            d24GPU::VehicleState veh_st;
            d24GPU::TMeasyState tirelf_st, tirerf_st, tirelr_st, tirerr_st;
            d24GPU::SuspensionState suslf_st, susrf_st, suslr_st, susrr_st;
            solver.Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, suslf_st, susrf_st, suslr_st,
                              susrr_st, 50);
            // Initialize wrote the static tire loads into the caller's structs, as the reference does
            expect_near("24-DOF initialized front tire load > 0", 1.0, tirelf_st._fz > 0 ? 1.0 : 0.0, 0.0);
            solver.SetEndTime(10.0);
            solver.Solve();
            expect_near("24-DOF acc3 x", 89.1625, solver.GetSimState(49)._v_states._x, 1.0);
        }
        {  // demo_hmmwv_controlsFunctor_18gpu.cu: SolveStep(t, functor) every 0.1 s for 10 s
            const std::string d = data + "/18dof/HMMWV/";
            d18SolverHalfImplicitGPU solver(64);
            solver.Construct(d + "vehicle.json", d + "tmeasy.json", 64);
            solver.SetTimeStep(1e-3);
            solver.SetKernelSimTime(0.1);
            d18GPU::VehicleState veh_st;
            d18GPU::TMeasyState tirelf_st, tirerf_st, tirelr_st, tirerr_st;
            solver.Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, 64);
            double t = 0.;
            while (t < 10.0 - 1e-4) t = SolveStep(solver, t, MyFunctor());
            // same controls through SolveStep(t, s, th, b)
            d18SolverHalfImplicitGPU ref(64);
            ref.Construct(d + "vehicle.json", d + "tmeasy.json", 64);
            ref.SetTimeStep(1e-3);
            ref.SetKernelSimTime(0.1);
            ref.Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, 64);
            double t2 = 0.;
            while (t2 < 10.0 - 1e-4) t2 = ref.SolveStep(t2, 0.0, t2 < 0.5 ? 0.0 : 0.5, 0.0);
            const double a = solver.GetSimState(63)._veh_state._x, b = ref.GetSimState(63)._veh_state._x;
            expect_near("functor SolveStep equals scalar SolveStep (x)", b, a, 0.0);
            expect_near("functor SolveStep moved the vehicle", 1.0, a > 10.0 ? 1.0 : 0.0, 0.0);
        }
    } catch (const Error& e) {
        std::printf("FAIL exception %d: %s\n", e.code(), e.what());
        return 1;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures);
    return failures ? 1 : 0;
}
