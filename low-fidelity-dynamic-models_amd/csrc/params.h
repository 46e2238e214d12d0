// Setup-time parameter handling: JSON files -> flat parameter vectors -> device records.
//
// Flat parameter vectors (the exchange format of the C ABI and of the fixtures), per model:
//   11-DOF [0..9]   a b h m jz muf mur maxSteer crankInertia maxBrakeTorque
//   18-DOF [0..19]  a b h m jz jx jxz cf cr muf mur hrcf hrcr krof kror brof bror maxSteer crankInertia
//                   maxBrakeTorque
//   24-DOF [0..15]  a b h m jz jx jy cf cr muf mur hrcf hrcr maxSteer crankInertia maxBrakeTorque
// then, for every model:
//   nonLinearSteer tcbool throttleMod driveType whichWheels   (0/1)
//   ngears, gearRatios[ngears]
//   six maps, each as n, x0, y0, x1, y1, ...: shift (x = downshift, y = upshift, rad/s), steer, torque
//   (rad/s -> N m), losses, capacity factor, torque ratio
//   tire: 35 TMeasy scalars  jw rr mu r0 pn pnmax cx cy kt dx dy rdyncoPn rdyncoP2n fzRdynco rdyncoCrit
//                            dfx0Pn dfx0P2n fxmPn fxmP2n fxsPn fxsP2n sxmPn sxmP2n sxsPn sxsP2n
//                            dfy0Pn dfy0P2n fymPn fymP2n fysPn fysP2n symPn symP2n sysPn sysP2n
//      or 40 TMeasyNr scalars jw rr mu r0 width rim_radius pn pnmax kt rdyncoPn rdyncoP2n fzRdynco rdyncoCrit
//                            dfx0Pn ... sysP2n (as above) vcoulomb frblend_begin frblend_end
//                            bearingCapacity li p_li p_use
//   24-DOF only: suspension ks bs
// Tire values are stored after tireInit (fzRdynco, rdyncoCrit recomputed) and, for the 11-DOF
// model, with the loaders' axle factor of 2 applied (its single-track tire stands for an axle).
#pragma once
#include <string>
#include <vector>

#include "lfdm_types.h"

namespace lfdm {

// Reads the reference's JSON files of one model (dof 11, 18 or 24; `sus` is used by 24 only).
bool params_from_json(int dof, const std::string& veh, const std::string& tire, const std::string& sus, int tire_nr,
                      std::vector<double>& flat, std::string& err);
bool pset_from_flat(int dof, const double* flat, int len, int tire_nr, PSet& out, std::string& err);

// Positions in a flat parameter vector of the values the per-vehicle parameters stand for (jx is -1 for
// the 11-DOF model, which has no roll inertia; the torque map's ordinates are flat[torque + 2 k + 1]).
struct FlatSlots {
    int m, jz, jx, max_brake, dfy0Pn, dfy0P2n, torque, torque_n;
};
bool flat_slots(int dof, const double* flat, int len, int tire_nr, FlatSlots& out, std::string& err);

// initializeTireSus (dof24.cpp:34-89) on one 24-DOF record, with the vehicle's own mass (which
// may differ from the set's through the per-vehicle parameters).
void init_tire_sus24(const PSet& p, double m, int tire_nr, double* rec);

// Driver table file: whitespace-separated "time steering throttle braking" rows, read until the
// first line that does not parse (utils.cpp:12-31).
bool load_driver_file(const std::string& path, std::vector<DrvRow>& rows, std::string& err);

}  // namespace lfdm
