// 11-DOF single-track vehicle (TMeasy / TMeasyNr axle tires) as register-resident device code.
//
// Follows wheeled_vehicle_models/11dof/dof11.cpp (physics) and dof11_halfImplicit.cpp:172-235
// (step) with the same IEEE operation order.  The tire cores are the shared ones (dof11.cpp:312-606
// is the 18-DOF tire code); the axle tire's doubled parameters come from the loader.
#pragma once
#include "dev_tire.h"

namespace lfdm {
namespace dev {

// Carried state: front (0) and rear (1) tire, chassis, crank and gear (udot is read back by the
// next step's load transfer, dof11.cpp:237-239).
struct S11 {
    double xe[2], ye[2], xedot[2], yedot[2], omega[2];
    double x, y, u, v, psi, wz, udot, crank;
    int gear;
};

// Quantities of the last right-hand side that the full state record also holds.
struct X11 {
    double dOmega[2], xt[2], rStat[2], fx[2], fy[2], fz[2], vsx[2], vsy[2], My[2], engTor[2];
    double vdot, wzdot, dcrank;
};

// Per-vehicle constants (m, jz, cornering stiffness, torque scale, brake torque and what depends
// on parameters only).
struct K11 {
    double m, jz, tqs, maxBrake, qy1, qy2;
    double Z1, Z3, Z4n, two_ab, mt;
};

LFDM_HD void k11_init(const PSet& P, double m, double jz, double dfy0Pn, double dfy0P2n,
                                         double tqs, double maxBrake, K11& k) {
    const double a = P.a, b = P.b, muf = P.muf, mur = P.mur;
    const double huf = P.tire.r0, hur = P.tire.r0;
    k.m = m; k.jz = jz; k.tqs = tqs; k.maxBrake = maxBrake;
    k.qy1 = 2. * dfy0Pn - 0.5 * dfy0P2n;
    k.qy2 = dfy0Pn - 0.5 * dfy0P2n;
    // computeTireLoads, dof11.cpp:229-247
    k.Z1 = (m * kG * b) / (2. * (a + b)) + (muf * kG) / 2.;
    k.Z3 = (m * kG * a) / (2. * (a + b)) + (mur * kG) / 2.;
    k.Z4n = m * P.h + muf * huf + mur * hur;
    k.two_ab = 2. * (a + b);
    // computeVehRHS, dof11.cpp:893-905
    k.mt = m + 2 * (muf + mur);
}

// The step's last part, the position update from the new velocities at the step's yaw angle, then
// the yaw angle (dof11_halfImplicit.cpp:214-219).  Nothing else in a step reads x, y or psi: a step loop may
// leave it out (POS = false) and apply it at the start of the next step or before the state is read,
// the same operations on the same values.
template <bool STRICT = false>
LFDM_HD void position11(S11& s, double dt) {
    double sp, cp;
    if (STRICT) sincos(s.psi, &sp, &cp);
    else sincos_fast(s.psi, sp, cp);
    s.x = s.x + (s.u * cp - s.v * sp) * dt;
    s.y = s.y + (s.u * sp + s.v * cp) * dt;
    s.psi = s.psi + s.wz * dt;
}

// One half-implicit step (rhsFun order, dof11_halfImplicit.cpp:560-605; update :172-235).
// STRICT: the reference's own quotients and transcendentals throughout (see dev_tire.h).
// POS = false leaves out the closing position update (position11).
template <bool NR, bool FULL, bool STRICT = false, bool POS = true>
LFDM_HD void step11(const PSet& P, const Derived& u, const K11& k, S11& s, double steering,
                                       double throttle, double braking, double dt, X11* x) {
    const TireP& t = P.tire;
    // computeTireLoads
    const double Z2 = xdiv<STRICT>(k.Z4n * (s.udot - s.wz * s.v), k.two_ab);
    double fz[2];
    fz[0] = k.Z1 - Z2;
    fz[1] = k.Z3 + Z2;
    fz[0] = fz[0] > 0. ? fz[0] : 0.;
    fz[1] = fz[1] > 0. ? fz[1] : 0.;
    // vehToTireTransform (dof11.cpp:45-71)
    const double delta = P.nonLinearSteer ? map_y<STRICT>(P.steer, steering) : steering * P.maxSteer;
    double sd, cd;
    if (STRICT) sincos(delta, &sd, &cd);
    else sincos_fast(delta, sd, cd);
    double vsx[2], vsy[2];
    vsy[0] = s.v + s.wz * P.a;
    vsx[0] = s.u * cd + vsy[0] * sd;
    vsy[1] = s.v - s.wz * P.b;
    vsx[1] = s.u;
    // computeTireRHS, front and rear (dof11.cpp:312-606)
    double fx[2], fy[2], My[2], xt[2], rStat[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double dl = i == 0 ? delta : u.delta_rear;
        const double d = i == 0 ? xdiv<STRICT>(sd, cd) : u.tan_delta_rear;
        if (fz[i] <= 0) {
            if (!NR) { s.xe[i] = 0; s.ye[i] = 0; }
            s.omega[i] = 0; xt[i] = 0; rStat[i] = t.r0; fx[i] = 0; fy[i] = 0; My[i] = 0;
            continue;
        }
        xt[i] = STRICT ? qdiv(fz[i], t.kt) : fz[i] * u.inv_kt;
        rStat[i] = t.r0 - xt[i];
        if (!NR)
            tmeasy_core<STRICT>(t, u, k.qy1, k.qy2, dl, d, fz[i], rStat[i], vsx[i], vsy[i], s.xe[i], s.ye[i], s.omega[i],
                        s.xedot[i], s.yedot[i], fx[i], fy[i], My[i]);
        else
            tmeasy_nr_core<STRICT>(t, u, k.qy1, k.qy2, dl, d, fz[i], rStat[i], vsx[i], vsy[i], s.omega[i], fx[i], fy[i],
                                   My[i]);
    }
    // computePowertrainRHS (dof11.cpp:607-749)
    double torque_t;
    double dcrank = 0.0;
    const double gr = pick(P.gears, P.ngears, s.gear);
    double shift_w;
    if (P.tcbool) {
        const double omega_out = xdiv<STRICT>(0.5 * (s.omega[0] + s.omega[1]), gr);
        const double omega_in = s.crank;
        double sr = 0;
        bool rev = false;
        if (!(omega_out < 1e-9 || omega_in < 1e-9)) {
            sr = xdiv<STRICT>(omega_out, omega_in);
            if (sr > 1.) {
                sr = 1. - (sr - 1.);
                rev = true;
            }
            if (sr < 0) sr = 0;
        }
        const double cf = map_y<STRICT>(P.cfmap, sr);
        const double tr = map_y<STRICT>(P.trmap, sr);
        const double q = xdiv<STRICT>(omega_in, cf);
        double torque_in = -(q * q);
        if (rev) torque_in = -torque_in;
        const double torque_out = rev ? -torque_in : -tr * torque_in;
        torque_t = xdiv<STRICT>(torque_out, gr);
        if (s.u < 1e-9 && torque_t < 0) torque_t = 0;
        dcrank = u.inv_ci * (drive_torque<STRICT>(P, k.tqs, throttle, s.crank) + torque_in);
        shift_w = omega_out;
    } else {
        s.crank = xdiv<STRICT>(0.5 * (s.omega[0] + s.omega[1]), gr);
        torque_t = xdiv<STRICT>(drive_torque<STRICT>(P, k.tqs, throttle, s.crank), gr);
        if (s.u < 1e-9 && torque_t < 0) torque_t = 0;
        shift_w = s.crank;
    }
    {
        double up = P.shift[0].y, dn = P.shift[0].x;
        for (int g = 1; g < P.ngears; ++g) {
            up = (s.gear == g) ? P.shift[g].y : up;
            dn = (s.gear == g) ? P.shift[g].x : dn;
        }
        if (shift_w > up) {
            if (s.gear < P.ngears - 1) s.gear++;
        } else if (shift_w < dn) {
            if (s.gear > 0) s.gear--;
        }
    }
    // differentialSplit between the axles (dof11.cpp:271-311): 4WD splits with the bias rule
    // (rear axle in the "left" slot), else whichWheels picks the driven axle
    double eng[2];
    if (P.driveType) {
        diff_split<STRICT>(torque_t, s.omega[1], s.omega[0], eng[1], eng[0]);
    } else if (P.whichWheels) {
        eng[1] = torque_t;
        eng[0] = 0;
    } else {
        eng[1] = 0;
        eng[0] = torque_t;
    }
    const double bt = k.maxBrake * braking;
    double dOm[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
        dOm[i] = u.inv_jw * (eng[i] + My[i] - sgn(s.omega[i]) * bt - fx[i] * rStat[i]);
    // tireToVehTransform (front tire only, dof11.cpp:100-124)
    {
        const double gx = fx[0] * cd - fy[0] * sd;
        const double gy = fx[0] * sd + fy[0] * cd;
        fx[0] = gx;
        fy[0] = gy;
    }
    // computeVehRHS (dof11.cpp:893-905)
    const double vdot = -s.u * s.wz + xdiv<STRICT>(fy[0] + fy[1], k.mt);
    const double udot = s.v * s.wz + xdiv<STRICT>(fx[0] + fx[1], k.mt);
    const double wzdot = xdiv<STRICT>(P.a * fy[0] - P.b * fy[1], k.jz);
    // Half-implicit update (dof11_halfImplicit.cpp:190-226)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (!NR) {
            s.xe[i] = s.xe[i] + s.xedot[i] * dt;
            s.ye[i] = s.ye[i] + s.yedot[i] * dt;
        }
        s.omega[i] = s.omega[i] + dOm[i] * dt;
    }
    if (P.tcbool) s.crank = s.crank + dcrank * dt;
    s.u = s.u + udot * dt;
    s.v = s.v + vdot * dt;
    s.wz = s.wz + wzdot * dt;
    s.udot = udot;
    if (POS) position11<STRICT>(s, dt);
    if (FULL) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            x->dOmega[i] = dOm[i]; x->xt[i] = xt[i]; x->rStat[i] = rStat[i]; x->fx[i] = fx[i]; x->fy[i] = fy[i];
            x->fz[i] = fz[i]; x->vsx[i] = vsx[i]; x->vsy[i] = vsy[i]; x->My[i] = My[i]; x->engTor[i] = eng[i];
        }
        x->vdot = vdot;
        x->wzdot = wzdot;
        x->dcrank = dcrank;
    }
}

}  // namespace dev
}  // namespace lfdm
