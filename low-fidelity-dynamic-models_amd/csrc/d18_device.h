// 18-DOF vehicle (TMeasy / TMeasyNr tires) as register-resident device code.
//
// Follows wheeled_vehicle_models/18dof/dof18.cpp (physics) and dof18_halfImplicit.cpp:245-318
// (step) with the same IEEE operation order.  What differs from the reference is only where
// work happens: parameter-only subexpressions are evaluated once per launch (same operands and
// order, hence the same bits), the steering angle's sine/cosine once per step instead of eight
// times, and the state lives in VGPRs across all steps of a launch.
#pragma once
#include "dev_tire.h"

namespace lfdm {
namespace dev {

// Carried state: everything step n+1 reads from step n (xedot/yedot are read back only when a
// tire is off the ground, dof18.cpp:399-409).
struct S18 {
    double xe[4], ye[4], xedot[4], yedot[4], omega[4];
    double x, y, u, v, psi, wz, phi, wx, udot, vdot, crank;
    int gear;
};

// Quantities of the last right-hand side that the full state record also holds.
struct X18 {
    double dOmega[4], xt[4], rStat[4], fx[4], fy[4], fz[4], vsx[4], vsy[4], My[4], engTor[4];
    double wxdot, wzdot, dcrank;
};

// Per-vehicle constants: per-vehicle parameters (m, jz, jx, cornering stiffness, torque scale,
// brake torque) and every expression that depends on parameters only.
struct K18 {
    double m, jz, jx, tqs, maxBrake;
    double Z1f, Z2fc, Z1r, Z2rc, Z4n;
    double neg_mt, E3c, hm, A2, A3, inv_mt, two_hrc_m, rdenom, A1A2, A1A3, A3sq;
    double qy1, qy2;  // InterpQ coefficients of dfy0 (2 w1 - w2/2, w1 - w2/2)
};

LFDM_HD void k18_init(const PSet& P, double m, double jz, double jx, double dfy0Pn, double dfy0P2n,
                                         double tqs, double maxBrake, K18& k) {
    const double a = P.a, b = P.b, h = P.h, cf = P.cf, cr = P.cr, muf = P.muf, mur = P.mur;
    const double huf = P.tire.r0, hur = P.tire.r0;
    k.m = m; k.jz = jz; k.jx = jx; k.tqs = tqs; k.maxBrake = maxBrake;
    // computeTireLoads, dof18.cpp:308-342
    k.Z1f = (m * kG * b) / (2. * (a + b)) + (muf * kG) / 2.;
    k.Z2fc = (muf * huf) / cf + m * b * (h - P.hrcf) / (cf * (a + b));
    k.Z1r = (m * kG * a) / (2. * (a + b)) + (mur * kG) / 2.;
    k.Z2rc = (mur * hur) / cr + m * a * (h - P.hrcr) / (cr * (a + b));
    k.Z4n = m * h + muf * huf + mur * hur;
    // computeVehRHS, dof18.cpp:1019-1078
    const double mt = m + 2 * (muf + mur);
    const double hrc = (P.hrcf * b + P.hrcr * a) / (a + b);
    const double A1 = mur * b - muf * a;
    const double A2 = jx + m * (hrc * hrc);
    const double A3 = hrc * m;
    k.neg_mt = -mt;
    k.E3c = m * kG * hrc;
    k.hm = hrc * m;
    k.A2 = A2;
    k.A3 = A3;
    k.inv_mt = 1 / mt;
    k.two_hrc_m = 2. * hrc * m;
    // the chassis solve's determinant, kept as its reciprocal (the fast step multiplies by it)
    k.rdenom = 1. / (A2 * (A1 * A1) - 2. * A1 * A3 * P.jxz + jz * (A3 * A3) + mt * (P.jxz * P.jxz) - A2 * jz * mt);
    k.A1A2 = A1 * A2;
    k.A1A3 = A1 * A3;
    k.A3sq = A3 * A3;
    k.qy1 = 2. * dfy0Pn - 0.5 * dfy0P2n;
    k.qy2 = dfy0Pn - 0.5 * dfy0P2n;
}

// Tire RHS, TMeasy (dof18.cpp:380-539): vertical load from the load transfer, static radius from the
// tire compression, then the force core.  Off the ground the tire resets and keeps its deflection
// rates (dof18.cpp:399-409).  delta / tan_delta: the tire's steering angle and its tangent.
template <bool STRICT = false>
LFDM_HD void tire_tmeasy(const TireP& t, const Derived& u, const K18& k, double delta, double tan_delta, double fz,
                                            double vsx, double vsy, double& xe, double& ye, double& xedot,
                                            double& yedot, double& omega, double& fx, double& fy, double& My,
                                            double& xt, double& rStat) {
    if (fz <= 0) {
        xe = 0; ye = 0; omega = 0; xt = 0; rStat = t.r0; fx = 0; fy = 0; My = 0;
        return;
    }
    xt = STRICT ? qdiv(fz, t.kt) : fz * u.inv_kt;
    rStat = t.r0 - xt;
    tmeasy_core<STRICT>(t, u, k.qy1, k.qy2, delta, tan_delta, fz, rStat, vsx, vsy, xe, ye, omega, xedot, yedot, fx,
                        fy, My);
}

// Tire RHS, TMeasyNr (dof18.cpp:541-673).
template <bool STRICT = false>
LFDM_HD void tire_tmeasy_nr(const TireP& t, const Derived& u, const K18& k, double delta, double tan_delta,
                                               double fz, double vsx, double vsy, double& omega, double& fx,
                                               double& fy, double& My, double& xt, double& rStat) {
    if (fz <= 0) {
        omega = 0; xt = 0; rStat = t.r0; fx = 0; fy = 0; My = 0;
        return;
    }
    xt = STRICT ? qdiv(fz, t.kt) : fz * u.inv_kt;
    rStat = t.r0 - xt;
    tmeasy_nr_core<STRICT>(t, u, k.qy1, k.qy2, delta, tan_delta, fz, rStat, vsx, vsy, omega, fx, fy, My);
}

// The step's last part, the position update from the new velocities at the step's yaw angle, then
// the yaw angle (dof18_halfImplicit.cpp:301-305).  Nothing else in a step reads x, y or psi: a step
// loop may leave it out (POS = false) and apply it at the start of the next step or before the state
// is read, the same operations on the same values.
template <bool STRICT = false>
LFDM_HD void position18(S18& s, double dt) {
    double sp, cp;
    if (STRICT) sincos(s.psi, &sp, &cp);
    else sincos_fast(s.psi, sp, cp);
    s.x = s.x + (s.u * cp - s.v * sp) * dt;
    s.y = s.y + (s.u * sp + s.v * cp) * dt;
    s.psi = s.psi + s.wz * dt;
}

// Tire right-hand-side outputs of one step (before the front forces are rotated), captured for the
// adjoint's reverse sweep.
struct TireOut18 {
    double fx[4], fy[4], My[4], rStat[4], xedot[4], yedot[4];
};

// One half-implicit step (rhsFun order, dof18_halfImplicit.cpp:618-673; update :245-318).
// STRICT: the reference's own quotients and transcendentals throughout (see dev_tire.h).  POS = false
// leaves out the closing position update (position18).
template <bool NR, bool FULL, bool CAPTURE = false, bool STRICT = false, bool POS = true>
LFDM_HD void step18(const PSet& P, const Derived& u, const K18& k, S18& s, double steering,
                                       double throttle, double braking, double dt, X18* x,
                                       TireOut18* to = nullptr) {
    const TireP& t = P.tire;
    // computeTireLoads
    const double lat = s.vdot + s.wz * s.u;
    const double Z2f = k.Z2fc * lat;
    const double Z3f = STRICT ? qdiv(P.krof * s.phi + P.brof * s.wx, P.cf) : (P.krof * s.phi + P.brof * s.wx) * u.inv_cf;
    const double Z4 = STRICT ? qdiv(k.Z4n * (s.udot - s.wz * s.v), u.two_ab)
                             : (k.Z4n * (s.udot - s.wz * s.v)) * u.inv_two_ab;
    const double Z2r = k.Z2rc * lat;
    const double Z3r = STRICT ? qdiv(P.kror * s.phi + P.bror * s.wx, P.cr) : (P.kror * s.phi + P.bror * s.wx) * u.inv_cr;
    double fz[4];
    fz[0] = k.Z1f - Z2f - Z3f - Z4;
    fz[1] = k.Z1f + Z2f + Z3f - Z4;
    fz[2] = k.Z1r - Z2r - Z3r + Z4;
    fz[3] = k.Z1r + Z2r + Z3r + Z4;
#pragma unroll
    for (int i = 0; i < 4; ++i) fz[i] = fz[i] > 0. ? fz[i] : 0.;
    // vehToTireTransform
    const double delta = P.nonLinearSteer ? map_y<STRICT>(P.steer, steering) : steering * P.maxSteer;
    double sd, cd;
    if (STRICT) sincos(delta, &sd, &cd);
    else sincos_fast(delta, sd, cd);
    double vsx[4], vsy[4];
    vsy[0] = s.v + s.wz * P.a;
    vsx[0] = (s.u - (s.wz * P.cf) / 2.) * cd + vsy[0] * sd;
    vsy[1] = vsy[0];
    vsx[1] = (s.u + (s.wz * P.cf) / 2.) * cd + vsy[1] * sd;
    vsy[2] = s.v - s.wz * P.b;
    vsx[2] = s.u - (s.wz * P.cr) / 2.;
    vsy[3] = vsy[2];
    vsx[3] = s.u + (s.wz * P.cr) / 2.;
    // computeTireRHS x4
    double fx[4], fy[4], My[4], xt[4], rStat[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double dl = i < 2 ? delta : u.delta_rear;
        const double d = i < 2 ? xdiv<STRICT>(sd, cd) : u.tan_delta_rear;
        if (!NR)
            tire_tmeasy<STRICT>(t, u, k, dl, d, fz[i], vsx[i], vsy[i], s.xe[i], s.ye[i], s.xedot[i], s.yedot[i],
                                s.omega[i], fx[i], fy[i], My[i], xt[i], rStat[i]);
        else
            tire_tmeasy_nr<STRICT>(t, u, k, dl, d, fz[i], vsx[i], vsy[i], s.omega[i], fx[i], fy[i], My[i], xt[i],
                                   rStat[i]);
    }
    if (CAPTURE) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            to->fx[i] = fx[i]; to->fy[i] = fy[i]; to->My[i] = My[i]; to->rStat[i] = rStat[i];
            to->xedot[i] = s.xedot[i]; to->yedot[i] = s.yedot[i];
        }
    }
    // computePowertrainRHS
    double dOm[4], eng[4], dcrank;
    powertrain4<STRICT>(P, u, k.tqs, k.maxBrake, throttle, braking, s.u, s.crank, s.gear, s.omega, fx, rStat, My, dOm, eng,
                dcrank);
    // tireToVehTransform (front tires only)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double gx = fx[i] * cd - fy[i] * sd;
        const double gy = fx[i] * sd + fy[i] * cd;
        fx[i] = gx;
        fy[i] = gy;
    }
    // computeVehRHS
    const double wzu = s.wz * s.u;
    const double E1 = k.neg_mt * s.wz * s.u + (fy[0] + fy[1] + fy[2] + fy[3]);
    const double E2 = (fy[0] + fy[1]) * P.a - (fy[2] + fy[3]) * P.b + (fx[1] - fx[0]) * P.cf / 2 +
                      (fx[3] - fx[2]) * P.cr / 2 + u.E2c * s.wz * s.u;
    const double E3 = k.E3c * s.phi - u.kro * s.phi - u.bro * s.wx + k.hm * s.wz * s.u;
    (void)wzu;
    const double udot = s.wz * s.v + k.inv_mt * ((fx[0] + fx[1] + fx[2] + fx[3]) + u.Uc * (s.wz * s.wz) -
                                                 k.two_hrc_m * s.wz * s.wx);
    const double A1 = u.A1, A2 = k.A2, A3 = k.A3, jz = k.jz, jxz = P.jxz, mt = -k.neg_mt;
    const double A3E1 = A3 * E1;
    const double nv = E1 * u.jxz2 - k.A1A2 * E2 + A1 * E3 * jxz + A3 * E2 * jxz - A2 * E1 * jz - A3 * E3 * jz;
    const double nwx = u.A1sq * E3 - k.A1A3 * E2 + A1 * E1 * jxz - A3E1 * jz + E2 * jxz * mt - E3 * jz * mt;
    const double nwz = k.A3sq * E2 - k.A1A2 * E1 - k.A1A3 * E3 + A3E1 * jxz - A2 * E2 * mt + E3 * jxz * mt;
    // the strict step divides by the determinant, formed as k18_init forms it (A1sq, A3sq and jxz2 are
    // the products it squares)
    const double denom = STRICT ? A2 * u.A1sq - 2. * A1 * A3 * jxz + jz * k.A3sq + mt * u.jxz2 - A2 * jz * mt : 0.0;
    const double vdot = STRICT ? qdiv(nv, denom) : nv * k.rdenom;
    const double wxdot = STRICT ? qdiv(nwx, denom) : nwx * k.rdenom;
    const double wzdot = STRICT ? qdiv(nwz, denom) : nwz * k.rdenom;
    // Half-implicit update (dof18_halfImplicit.cpp:262-305)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!NR) {
            s.xe[i] = s.xe[i] + s.xedot[i] * dt;
            s.ye[i] = s.ye[i] + s.yedot[i] * dt;
        }
        s.omega[i] = s.omega[i] + dOm[i] * dt;
    }
    if (P.tcbool) s.crank = s.crank + dcrank * dt;
    s.u = s.u + udot * dt;
    s.v = s.v + vdot * dt;
    s.wx = s.wx + wxdot * dt;
    s.wz = s.wz + wzdot * dt;
    s.phi = s.phi + s.wx * dt;
    s.udot = udot;
    s.vdot = vdot;
    if (POS) position18<STRICT>(s, dt);
    if (FULL) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            x->dOmega[i] = dOm[i]; x->xt[i] = xt[i]; x->rStat[i] = rStat[i]; x->fx[i] = fx[i]; x->fy[i] = fy[i];
            x->fz[i] = fz[i]; x->vsx[i] = vsx[i]; x->vsy[i] = vsy[i]; x->My[i] = My[i]; x->engTor[i] = eng[i];
        }
        x->wxdot = wxdot;
        x->wzdot = wzdot;
        x->dcrank = dcrank;
    }
}

}  // namespace dev
}  // namespace lfdm
