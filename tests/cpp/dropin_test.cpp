// The reference's GPU known-answer tests and functor demo, written against the C++ drop-in classes
// (include/lfdm_amd.hpp, include/lfdm_amd_functor.hpp) exactly as they are written against
// d11/d18/d24SolverHalfImplicitGPU: wheeled_vehicle_models/tests/{11,18,24}dof_gpu.cu and
// 18dof-gpu/demos/HMMWV/demo_hmmwv_controlsFunctor_18gpu.cu.
//
// Usage: dropin_test <data dir (tests/golden/data)> [records file].  Prints one line per check and exits
// non-zero on a failure; writes the per-vehicle functor fleet's final records (200 x 75 doubles) to the
// records file when one is given.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

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

// Per-vehicle controls the way a reference user writes them: the functor reads the kernel's vehicle
// index from the thread index (dof18_halfImplicit_gpu.cuh:465).  Exact binary fractions and small integers
// only, so the oracle side (tests/test_cpp_dropin.py) forms the same doubles.
__device__ __host__ inline void per_vehicle_controls(double t, unsigned v, DriverInput* c) {
    c->m_steering = t < 1.0 ? 0.0 : 0.0625 * (double)(v % 9) - 0.25;
    c->m_throttle = t < 0.5 ? 0.0 : 0.25 + 0.0625 * (double)(v % 11);
    c->m_braking = (v % 13 == 0 && t > 6.0) ? 0.375 : 0.0;
}

struct PerVehicleFunctor {
    __device__ void operator()(double t, DriverInput* controls) {
        per_vehicle_controls(t, blockIdx.x * blockDim.x + threadIdx.x, controls);
    }
};

// The same controls from the vehicle index the library passes (independent of the slot layout).
struct PerVehicleIndexFunctor {
    __device__ void operator()(double t, DriverInput* controls, unsigned vehicle) {
        per_vehicle_controls(t, vehicle, controls);
    }
};

bool write_records(const std::string& path, const std::vector<std::vector<double>>& recs) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    for (const auto& r : recs) std::fwrite(r.data(), sizeof(double), r.size(), f);
    std::fclose(f);
    return true;
}

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
        {  // per-vehicle functor controls (200 vehicles, one Construct group, 10 s in 0.1 s chunks)
            const std::string d = data + "/18dof/HMMWV/";
            const unsigned n = 200;
            d18SolverHalfImplicitGPU solver(n);
            solver.Construct(d + "vehicle.json", d + "tmeasy.json", n);
            solver.SetTimeStep(1e-3);
            solver.SetKernelSimTime(0.1);
            d18GPU::VehicleState veh_st;
            d18GPU::TMeasyState tirelf_st, tirerf_st, tirelr_st, tirerr_st;
            solver.Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, n);
            double t = 0.;
            while (t < 10.0 - 1e-4) t = SolveStep(solver, t, PerVehicleFunctor());
            // the vehicle-index form on two Construct groups (slots of the second group are padded away
            // from the vehicle indices): the same controls per vehicle, so the same bits
            d18SolverHalfImplicitGPU two(n);
            two.Construct(d + "vehicle.json", d + "tmeasy.json", 70);
            two.Construct(d + "vehicle.json", d + "tmeasy.json", n - 70);
            two.SetTimeStep(1e-3);
            two.SetKernelSimTime(0.1);
            two.Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, n);
            double t2 = 0.;
            while (t2 < 10.0 - 1e-4) t2 = SolveStep(two, t2, PerVehicleIndexFunctor());
            std::vector<std::vector<double>> recs;
            double worst = 0.0;
            for (unsigned v = 0; v < n; ++v) {
                recs.push_back(solver.GetRecord(v));
                const auto r2 = two.GetRecord(v);
                for (size_t k = 0; k < r2.size(); ++k) worst = std::fmax(worst, std::fabs(r2[k] - recs.back()[k]));
            }
            expect_near("vehicle-index functor on two groups equals thread-index functor (max |diff|)", 0.0, worst,
                        0.0);
            const double x0 = recs[0][60], x5 = recs[5][60];
            expect_near("per-vehicle controls give per-vehicle trajectories", 1.0, x0 != x5 ? 1.0 : 0.0, 0.0);
            if (argc > 2 && !write_records(argv[2], recs)) {
                std::printf("FAIL cannot write %s\n", argv[2]);
                return 1;
            }
        }
        {  // per-vehicle parameters: m_sim_data[i]._veh_param of the reference, through GetParams / SetParams
            const std::string d = data + "/18dof/HMMWV/";
            d18SolverHalfImplicitGPU solver(64), plain(64);
            for (auto* s : {&solver, &plain}) {
                s->Construct(d + "vehicle.json", d + "tmeasy.json", 64, in + "acc3.txt");
                d18GPU::VehicleState veh_st;
                d18GPU::TMeasyState tirelf_st, tirerf_st, tirelr_st, tirerr_st;
                s->Initialize(veh_st, tirelf_st, tirerf_st, tirelr_st, tirerr_st, 64);
                s->SetEndTime(10.0);
            }
            d18GPU::VehicleParam veh;
            d18GPU::TMeasyParam tire;
            solver.GetParams(3, veh, tire);
            expect_near("GetParams reads the JSON's CG height", 0.8, veh._h > 0.1 ? 0.8 : 0.0, 0.0);
            solver.SetParams(1, veh, tire);  // unchanged values: vehicle 1 keeps its trajectory
            veh._h += 0.2;
            tire._cx *= 0.8;
            solver.SetParams(3, veh, tire);
            d18GPU::VehicleParam back;
            d18GPU::TMeasyParam back_t;
            solver.GetParams(3, back, back_t);
            expect_near("SetParams then GetParams (h)", veh._h, back._h, 0.0);
            expect_near("SetParams then GetParams (cx)", tire._cx, back_t._cx, 0.0);
            solver.Solve();
            plain.Solve();
            expect_near("unchanged parameters keep the trajectory (x)", plain.GetSimState(1)._veh_state._x,
                        solver.GetSimState(1)._veh_state._x, 0.0);
            expect_near("other vehicles keep their trajectory (x)", plain.GetSimState(2)._veh_state._x,
                        solver.GetSimState(2)._veh_state._x, 0.0);
            expect_near("the edited vehicle moves otherwise", 1.0,
                        plain.GetSimState(3)._veh_state._x != solver.GetSimState(3)._veh_state._x ? 1.0 : 0.0, 0.0);
        }
    } catch (const Error& e) {
        std::printf("FAIL exception %d: %s\n", e.code(), e.what());
        return 1;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures);
    return failures ? 1 : 0;
}
