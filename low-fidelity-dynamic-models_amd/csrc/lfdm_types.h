// Plain-old-data layouts shared by the host library and the HIP kernels.
//
// A parameter set ("group") is what one Construct call of the reference produces: one vehicle
// JSON plus one tire JSON (plus, for 24-DOF, one suspension JSON) and one optional driver table.
// Sets live in device memory as fixed-size records and are read by whole wavefronts through
// scalar loads; every vehicle slot of a wavefront belongs to one set (groups are padded to 64
// slots on the device).
#pragma once
#include <cstdint>

namespace lfdm {

constexpr int kWave = 64;
constexpr int kMaxMap = 48;    // entries per piecewise-linear map
constexpr int kMaxGears = 16;  // gears per transmission
constexpr int kOverrides = 7;  // per-vehicle fields: m, jz, jx, dfy0Pn, dfy0P2n, torque_scale, maxBrakeTorque

struct MapXY {
    double x, y;
};

struct Map {
    int n;
    int pad;
    MapXY e[kMaxMap];
};

// TMeasy / TMeasyNr tire parameters (dof18.h:23-152, 263-399), after tireInit.
struct TireP {
    double jw, rr, mu, r0, width, rim_radius, pn, pnmax, cx, cy, kt, dx, dy;
    double rdyncoPn, rdyncoP2n, fzRdynco, rdyncoCrit;
    double dfx0Pn, dfx0P2n, fxmPn, fxmP2n, fxsPn, fxsP2n, sxmPn, sxmP2n, sxsPn, sxsP2n;
    double dfy0Pn, dfy0P2n, fymPn, fymP2n, fysPn, fysP2n, symPn, symP2n, sysPn, sysP2n;
    double vcoulomb, frblend_begin, frblend_end, bearingCapacity, li, p_li, p_use;
};

// Parameter-only subexpressions of one set, evaluated once on the host with the same operands and
// operation order as the reference evaluates them every step (so the same bits).  The first row is
// used by the 18-DOF chassis only.
struct Derived {
    double two_ab, E2c, Uc, A1, A1sq, jxz2, kro, bro, inv_jw, inv_ci;
    double qx1, qx2, qfxm1, qfxm2, qfym1, qfym2, qfxs1, qfxs2, qfys1, qfys2;  // InterpQ pairs (2w1 - w2/2, w1 - w2/2)
    double lsxm, lsym, lsxs, lsys, lrd;                                       // InterpL slopes (w2 - w1)
    double delta_rear;      // steering angle of the rear tires (steering input 0)
    double tan_delta_rear;  // its tangent
    // reciprocals and quotients of parameters
    double inv_pn, inv_kt, inv_two_ab, krof_cf, kror_cr, brof_cf, bror_cr, inv_cf, inv_cr, inv_mu;
};

// Vehicle parameters of one set (11-DOF dof11.h:496-620, 18-DOF dof18.h:489-799, 24-DOF
// dof24.h:580-764 plus SuspensionParam :766-780) with the tire set and the group's driver table.
// One layout serves the three models; a model reads only its own fields (11-DOF: a b h m jz muf
// mur; 24-DOF adds jx jy cf cr hrcf hrcr ks bs; 18-DOF: jx jxz cf cr hrcf hrcr krof kror brof bror).
struct PSet {
    double a, b, h, m, jz, jx, jy, jxz, cf, cr, muf, mur, hrcf, hrcr, krof, kror, brof, bror, ks, bs;
    double maxSteer, crankInertia, maxBrakeTorque;
    int nonLinearSteer, tcbool, throttleMod, driveType, whichWheels, ngears, tire_nr, dof;
    double gears[kMaxGears];
    MapXY shift[kMaxGears];
    Map steer, torque, losses, cfmap, trmap;
    TireP tire;
    Derived d;
    int32_t drv_off, drv_rows;  // rows of this group's driver table in the table pool
};

// One driver-table row (utils.h:19-31): time, steering, throttle, braking.
struct DrvRow {
    double t, steering, throttle, braking;
};

// Control sources of a launch.
enum CtrlMode : int {
    kCtrlTable = 0,  // group driver table, reference GetDriverInput semantics
    kCtrlConst = 1,  // one (steering, throttle, braking) for every vehicle (SolveStep)
    kCtrlKnots = 2,  // per-vehicle knot program (steering lerp grid, throttle/brake step grid)
    kCtrlSteps = 3,  // per-vehicle per-step arrays [step][3][slot]
    kCtrlFunctor = 4,  // a device functor evaluated per vehicle and step (kernel instantiated by the caller)
};

}  // namespace lfdm
