// Setup-time parameter handling: JSON files -> flat parameter vectors -> device records.
//
// Flat 18-DOF vector (the exchange format of the C ABI and of the fixtures):
//   [0..19]  a b h m jz jx jxz cf cr muf mur hrcf hrcr krof kror brof bror maxSteer crankInertia maxBrakeTorque
//   [20..24] nonLinearSteer tcbool throttleMod driveType whichWheels   (0/1)
//   ngears, gearRatios[ngears]
//   six maps, each as n, x0, y0, x1, y1, ...: shift (x = downshift, y = upshift, rad/s), steer, torque
//   (rad/s -> N m), losses, capacity factor, torque ratio
//   tire: 35 TMeasy scalars  jw rr mu r0 pn pnmax cx cy kt dx dy rdyncoPn rdyncoP2n fzRdynco rdyncoCrit
//                            dfx0Pn dfx0P2n fxmPn fxmP2n fxsPn fxsP2n sxmPn sxmP2n sxsPn sxsP2n
//                            dfy0Pn dfy0P2n fymPn fymP2n fysPn fysP2n symPn symP2n sysPn sysP2n
//      or 40 TMeasyNr scalars jw rr mu r0 width rim_radius pn pnmax kt rdyncoPn rdyncoP2n fzRdynco rdyncoCrit
//                            dfx0Pn ... sysP2n (as above) vcoulomb frblend_begin frblend_end
//                            bearingCapacity li p_li p_use
// Tire values are stored after tireInit (fzRdynco, rdyncoCrit recomputed).
#pragma once
#include <string>
#include <vector>

#include "lfdm_types.h"

namespace lfdm {

bool params18_from_json(const std::string& veh, const std::string& tire, int tire_nr, std::vector<double>& flat,
                        std::string& err);
bool p18_from_flat(const double* flat, int len, int tire_nr, P18& out, std::string& err);

// Driver table file: whitespace-separated "time steering throttle braking" rows, read until the
// first line that does not parse (utils.cpp:12-31).
bool load_driver_file(const std::string& path, std::vector<DrvRow>& rows, std::string& err);

}  // namespace lfdm
