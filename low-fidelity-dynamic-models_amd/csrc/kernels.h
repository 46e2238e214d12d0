// Launch records and host-side entry points of the HIP kernels.
//
// Records (one per vehicle; SoA on the device as [field][slot]); the full field lists are in
// integrate.h:
//   11-DOF  2 tires x 15 (TMeasy) / 11 (TMeasyNr) + 12 vehicle fields       = 42 / 34
//   18-DOF  4 tires x 15 / 11 + 15 vehicle fields                           = 75 / 59
//   24-DOF  4 tires x 23 / 19 + 4 struts x 24 + 24 vehicle fields           = 212 / 196
#pragma once
#include <hip/hip_runtime.h>

#include "lfdm_types.h"

namespace lfdm {

// Parameter set i of a launch.  Sets are never written while a kernel runs, so they are read
// through the constant address space: reads at a wave-uniform address then become scalar loads into
// SGPRs instead of vector loads into VGPR pairs (which also lowers the step's VGPR pressure).
__device__ __forceinline__ const PSet& pset_at(const PSet* sets, int i) {
    return *(const PSet*)((const __attribute__((address_space(4))) PSet*)sets + i);
}

// Most waves per SIMD each model's integration kernels are built for (each also exists at every lower
// occupancy; the maximum is the default on fleets of more than one wavefront per SIMD).
constexpr int kMaxWaves11 = 4, kMaxWaves18 = 3, kMaxWaves24 = 2;

struct Launch {
    const PSet* sets;       // parameter sets
    const int* wave_set;   // set index of every 64-slot wavefront
    const int* slot_user;  // user vehicle index of every slot, -1 for padding
    double* state;         // [record field][S]
    const double* ovr;     // [kOverrides][S]
    const DrvRow* drv;     // driver-table pool (kCtrlTable)
    const double* knots;   // [ns + 2 np][n_user] (kCtrlKnots)
    const double* steps;   // [nsteps][3][n_user] for this launch (kCtrlSteps)
    const int* out_steps;  // step counts after which an output row is written, ascending
    double* out;           // [n_out][channels][n_user]
    double t0, dt;
    double c_s, c_t, c_b;  // kCtrlConst
    double ds, dp;         // knot spacings (kCtrlKnots)
    int ns, np;
    int nsteps, n_out, S, n_user, mode, nr;
    int steps_shared;      // kCtrlSteps: 1 = steps[nsteps][3] shared by every vehicle
    // per-vehicle horizons (null: every vehicle runs all nsteps): a vehicle stops after
    // horizon[slot] steps counted from step 0 of the solve; step_base steps were done before this
    // launch; wave_hmax[wave] is the largest horizon of the wavefront
    const int* horizon;
    const int* wave_hmax;
    int step_base;
    int waves;  // waves per SIMD of the kernel: 1 up to the model's maximum, or 0 for its default
    int strict;  // 1: the step with the reference's own quotients and transcendentals (LFDM_ARITH_STRICT)
};

// Integration launches of the 11-, 18- and 24-DOF models (one kernel per tire model and control
// source inside each).
hipError_t launch_d11(const Launch& L, int block, hipStream_t stream);
hipError_t launch_d18(const Launch& L, int block, hipStream_t stream);
hipError_t launch_d24(const Launch& L, int block, hipStream_t stream);

// Differentiation form (18-DOF TMeasy).  Forward: from the record state, nsteps steps with per-step
// controls ctrl[nsteps][3][n_user]; writes xy[nsteps][2][n_user] (position after each step) and
// the carried state every K steps into ckpt[ceil(nsteps/K)][32][S].  Backward, one group of
// segments [seg_lo, seg_hi) at a time, last group first: the recompute launch reruns the group's
// segments from their checkpoints into tape[seg_hi - seg_lo][S / 64][K][38][64] (the carried state and
// tire outputs the step adjoint reads, every step, blocked by wavefront), the reverse launch walks them backwards.  From xy_bar (layout of xy) the
// last group's reverse writes ctrl_bar (layout of ctrl), state_bar[75][n_user] (adjoint of the
// initial record) and param_bar[7][n_user] (m, jz, jx, dfy0Pn, dfy0P2n, torque scale,
// maxBrakeTorque); between groups the running adjoints live in carry[52][S].
constexpr int kAdjCkptRow = 32;  // doubles per vehicle in a checkpoint
constexpr int kAdjTapeRow = 38;  // doubles per vehicle and step in the tape
constexpr int kAdjCarry = 52;    // doubles per vehicle in carry

struct AdjLaunch {
    const PSet* sets;
    const int* wave_set;
    const int* slot_user;
    const double* state;
    const double* ovr;
    const double* ctrl;
    double* xy;
    const double* xy_bar;
    double* ctrl_bar;
    double* state_bar;
    double* param_bar;
    double* ckpt;
    double* tape;
    double* carry;
    double dt;
    int nsteps, K, S, n_user;
    int seg_lo, seg_hi;  // backward: the group of this launch
    int fwd_tape_lo;     // forward: segments from this one on also go to the tape (block 0 onwards)
    double* jac;         // step Jacobian: [kJacRows][kJacCols][n_user]
};

// Step Jacobian (lfdm_step_jacobian): rows are the stepped packed states, columns the packed states
// then steering, throttle and braking, both in the reference's packY order of a vehicle with a torque
// converter (xe ye per tire, omega x 4, crankOmega, x y u v psi wz phi wx).
constexpr int kJacRows = 21;
constexpr int kJacCols = kJacRows + 3;

// waves_per_simd 1 or 2: register budget of the forward kernel (1 when the fleet fills at most one
// wavefront per SIMD, so the step runs without spills).
hipError_t launch_adj_forward(const AdjLaunch& L, int waves_per_simd, hipStream_t stream);
hipError_t launch_adj_recompute(const AdjLaunch& L, hipStream_t stream);
hipError_t launch_adj_reverse(const AdjLaunch& L, hipStream_t stream);
// Every vehicle's step Jacobian at its current record under the per-vehicle controls ctrl[3][n_user].
hipError_t launch_step_jacobian(const AdjLaunch& L, hipStream_t stream);

// Records `msg` as the calling thread's lfdm_last_error() text and returns `code`.
int set_last_error(int code, const char* msg);

// Gathers full records of `n` slots from SoA state into out[n][rec].
hipError_t launch_gather(const double* state, int S, int rec, const int* slots, int n, double* out,
                         hipStream_t stream);
// Scatters field-major rows src[F][n] (vehicle v in column v) into dst[F][S] at user_slot[v].
hipError_t launch_scatter_fields(double* dst, int S, int F, const int* user_slot, int n, const double* src,
                                 hipStream_t stream);
// Scatters records in[n][rec] into the SoA state at `slots`.
hipError_t launch_scatter(double* state, int S, int rec, const int* slots, int n, const double* in,
                          hipStream_t stream);

}  // namespace lfdm
