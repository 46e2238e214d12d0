// Launch records and host-side entry points of the HIP kernels.
//
// Canonical 18-DOF record (one per vehicle; SoA on the device as [field][slot]):
//   per tire LF, RF, LR, RR  TMeasy:   xe ye xedot yedot omega dOmega xt rStat fx fy fz vsx vsy My engTor (15)
//                            TMeasyNr: omega dOmega xt rStat fx fy fz vsx vsy My engTor             (11)
//   vehicle:                 x y u v psi wz phi wx udot vdot wxdot wzdot crankOmega dOmega_crank gear  (15)
#pragma once
#include <hip/hip_runtime.h>

#include "lfdm_types.h"

namespace lfdm {

struct Launch18 {
    const P18* sets;       // parameter sets
    const int* wave_set;   // set index of every 64-slot wavefront
    const int* slot_user;  // user vehicle index of every slot, -1 for padding
    double* state;         // [record field][S]
    const double* ovr;     // [kOverrides][S]
    const DrvRow* drv;     // driver-table pool (kCtrlTable)
    const double* knots;   // [ns + 2 np][n_user] (kCtrlKnots)
    const double* steps;   // [nsteps][3][n_user] for this launch (kCtrlSteps)
    const int* out_steps;  // step counts after which an output row is written, ascending
    double* out;           // [n_out][13][n_user]
    double t0, dt;
    double c_s, c_t, c_b;  // kCtrlConst
    double ds, dp;         // knot spacings (kCtrlKnots)
    int ns, np;
    int nsteps, n_out, S, n_user, mode, nr;
};

hipError_t launch_d18(const Launch18& L, int block, hipStream_t stream);

// Gathers full records of `n` slots from SoA state into out[n][rec].
hipError_t launch_gather(const double* state, int S, int rec, const int* slots, int n, double* out,
                         hipStream_t stream);
// Scatters records in[n][rec] into the SoA state at `slots`.
hipError_t launch_scatter(double* state, int S, int rec, const int* slots, int n, const double* in,
                          hipStream_t stream);

}  // namespace lfdm
