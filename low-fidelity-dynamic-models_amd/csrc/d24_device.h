// 24-DOF vehicle (heave, pitch, roll, four suspension struts, tire compression; TMeasy / TMeasyNr)
// as register-resident device code.
//
// Follows wheeled_vehicle_models/24dof/dof24.cpp (physics) and dof24_halfImplicit.cpp:188-297
// (step, right-hand-side order :594-706) with the same IEEE operation order.  Every sin/cos of the
// pitch and roll angles is evaluated once per step (the reference evaluates the same arguments
// many times; the values are identical), and the Euler-angle rates of the chassis RHS, which the
// update step recomputes before any use, are not evaluated twice.
#pragma once
#include "dev_tire.h"

namespace lfdm {
namespace dev {

// Carried state.  Per tire LF RF LR RR: deflections (TMeasy), compression xt, wheel speed and the
// static radius of the last step (read by vehToTireTransform, dof24.cpp:340-364); per strut:
// deflection xs, unsprung heave velocity wu, last deflection rate dxs and damper transfer force fzd
// (read by computeSusRHS, :1199-1228); chassis velocities and angles with the accelerations of the
// last step (read by vehToSusTransform, :199-229).  The record's per-strut values that Initialize
// sets and no step changes (xsi, lsi, the tire's vsz) and the tires' deflection rates are in K24.
struct S24 {
    double xe[4], ye[4], xt[4], omega[4], rStat[4];
    double xs[4], wu[4], dxs[4], fzd[4];
    double x, y, u, v, w, psi, phi, theta, wx, wy, wz;
    double udot, vdot, wdot, wxdot, wydot, wzdot, crank;
    int gear;
};

// Quantities of the last right-hand side that the full state record also holds.
struct X24 {
    double dOmega[4], dxt[4], fx[4], fy[4], fz[4], fxg[4], fyg[4], fzg[4], vgx[4], vgy[4], vsx[4], vsy[4];
    double My[4], engTor[4];
    double ls[4], us[4], vs[4], ws[4], dus[4], dvs[4], uu[4], vu[4], duu[4], dvu[4], dwu[4];
    double fxs[4], fys[4], fzs[4], mx[4], my[4];
    double dtheta, dphi, dpsi, dcrank;
};

// Per-vehicle values the integration kernel keeps in LDS rather than in registers: the constants from
// the parameters and overrides (k24_init), the record's per-strut values no step changes (initial strut
// deflection xsi and length lsi, tire compression velocity vsz), and the TMeasy deflection rates xedot,
// yedot (state, but written once per step by the tire and read once by the update); the last two
// groups are read from the record with the state.
struct K24 {
    double xsi[4], lsi[4], vsz[4], xedot[4], yedot[4];
    double m, tqs, maxBrake, qy1, qy2;
    double inv_m, inv_jx, inv_jy, inv_jz, mu_g[2], inv_mu[2], half_cf, half_cr;
};

LFDM_HD void k24_init(const PSet& P, double m, double jz, double jx, double dfy0Pn, double dfy0P2n,
                                         double tqs, double maxBrake, K24& k) {
    k.m = m; k.tqs = tqs; k.maxBrake = maxBrake;
    k.qy1 = 2. * dfy0Pn - 0.5 * dfy0P2n;
    k.qy2 = dfy0Pn - 0.5 * dfy0P2n;
    k.inv_m = 1 / m;
    k.inv_jx = 1 / jx;
    k.inv_jy = 1 / P.jy;
    k.inv_jz = 1 / jz;
    k.mu_g[0] = P.muf * kG;
    k.mu_g[1] = P.mur * kG;
    k.inv_mu[0] = 1 / P.muf;
    k.inv_mu[1] = 1 / P.mur;
    k.half_cf = P.cf / 2.;
    k.half_cr = P.cr / 2.;
}

// vehToSusTransform (dof24.cpp:148-230) of strut i: strut-top velocities, strut length and
// unsprung-mass velocities from the chassis velocities and the strut deflection.
struct Kin24 {
    double us, vs, ls, uu, vu;
};

LFDM_HD Kin24 kin24(const PSet& P, const S24& s, const K24& k, int i) {
    const double a = P.a, b = P.b, cf = P.cf, cr = P.cr;
    Kin24 q;
    if (i == 0) {
        q.us = (-cf * s.wz / 2.) + s.u;
        q.vs = a * s.wz + s.v;
    } else if (i == 1) {
        q.us = (cf * s.wz / 2.) + s.u;
        q.vs = a * s.wz + s.v;
    } else if (i == 2) {
        q.us = (-cr * s.wz / 2.) + s.u;
        q.vs = -b * s.wz + s.v;
    } else {
        q.us = (cr * s.wz / 2.) + s.u;
        q.vs = -b * s.wz + s.v;
    }
    q.ls = k.lsi[i] - (s.xs[i] - k.xsi[i]);
    q.uu = q.us - q.ls * s.wy;
    q.vu = q.vs + q.ls * s.wx;
    return q;
}

// The step's last part, the position update from the new velocities and yaw angle
// (dof24_halfImplicit.cpp:275-280).  Nothing else in a step reads x and y, and psi only here: a step
// loop may leave it out (POS = false) and apply it at the start of the next step or before the state
// is read, the same operations on the same values.
template <bool STRICT = false>
LFDM_HD void position24(S24& s, double dt) {
    double sp, cp;
    if (STRICT) sincos(s.psi, &sp, &cp);
    else sincos_fast(s.psi, sp, cp);
    s.x = s.x + (s.u * cp - s.v * sp) * dt;
    s.y = s.y + (s.u * sp + s.v * cp) * dt;
}

// One half-implicit step (rhsFun order, dof24_halfImplicit.cpp:594-706; update :188-297).
// STRICT: the reference's own quotients and transcendentals throughout (see dev_tire.h): the static
// radius divided by cos(theta) cos(phi), the jacking forces by the track widths, the Euler-angle
// rates by cos(theta) with tan(theta) from the libm.  Where a reference function uses only the sine
// or only the cosine of an angle (computeTireRHS: cos theta and cos phi; computeForcesThroughSus:
// sin phi; the Euler-angle rates of the step: cos theta), the strict step calls that function alone;
// everywhere else the reference's sin and cos of one angle compile to one sincos, as here.
// Each strut's kinematics (vehToSusTransform) are computed where a stage uses them: before its tire
// and again after the tires, from the same state and with the same operations (the same values),
// so they are not held in registers across the four tire evaluations.
// POS = false leaves out the closing position update (position24).
template <bool NR, bool FULL, bool STRICT = false, bool POS = true>
LFDM_HD void step24(const PSet& P, const Derived& u, K24& k, S24& s, double steering,
                                       double throttle, double braking, double dt, X24* x) {
    const TireP& t = P.tire;
    const double a = P.a, b = P.b, cf = P.cf, cr = P.cr;
    double sth, cth, sph, cph;
    if (STRICT) {
        sincos(s.theta, &sth, &cth);
        sincos(s.phi, &sph, &cph);
    } else {
        sincos_fast(s.theta, sth, cth);
        sincos_fast(s.phi, sph, cph);
    }
    const double delta = P.nonLinearSteer ? map_y<STRICT>(P.steer, steering) : steering * P.maxSteer;
    double sd, cd;
    if (STRICT) sincos(delta, &sd, &cd);
    else sincos_fast(delta, sd, cd);
    // per corner: vehToTireTransform (dof24.cpp:313-385) with last step's static radius, then
    // computeTireRHS (dof24.cpp:827-1121): load from the compression, radius from the attitude
    double fx[4], fy[4], fz[4], My[4];
    const double cc = cth * cph;
    const double rcc = frcp(cc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const Kin24 q = kin24(P, s, k, i);
        const double vgx = cth * (q.uu - s.wy * s.rStat[i]) + sth * (s.wu[i] * cph + sph * (s.wx * s.rStat[i] + q.vu));
        const double vgy = cph * (q.vu + s.wx * s.rStat[i]) - s.wu[i] * sph;
        const double vsy = vgy;
        const double vsx = i < 2 ? vgx * cd + vgy * sd : vgx;
        if (FULL) {
            x->vgx[i] = vgx; x->vgy[i] = vgy; x->vsx[i] = vsx; x->vsy[i] = vsy;
        }
        const double dl = i < 2 ? delta : u.delta_rear;
        const double d = i < 2 ? xdiv<STRICT>(sd, cd) : u.tan_delta_rear;
        fz[i] = s.xt[i] * t.kt;
        s.rStat[i] = STRICT ? qdiv(t.r0 - s.xt[i], cos(s.theta) * cos(s.phi)) : (t.r0 - s.xt[i]) * rcc;
        if (fz[i] <= 0) {
            if (!NR) { s.xe[i] = 0; s.ye[i] = 0; }
            s.omega[i] = 0; s.xt[i] = 0; s.rStat[i] = t.r0; fz[i] = 0; fx[i] = 0; fy[i] = 0; My[i] = 0;
            continue;
        }
        if (!NR)
            tmeasy_core<STRICT>(t, u, k.qy1, k.qy2, dl, d, fz[i], s.rStat[i], vsx, vsy, s.xe[i], s.ye[i],
                                s.omega[i], k.xedot[i], k.yedot[i], fx[i], fy[i], My[i]);
        else
            tmeasy_nr_core<STRICT>(t, u, k.qy1, k.qy2, dl, d, fz[i], s.rStat[i], vsx, vsy, s.omega[i], fx[i],
                                   fy[i], My[i]);
    }
    // the strut kinematics again, recomputed rather than carried over the tires
    opaque(s.wz);
#pragma unroll
    for (int i = 0; i < 4; ++i) opaque(s.xs[i]);
    double ls[4], uu[4], vu[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const Kin24 q = kin24(P, s, k, i);
        ls[i] = q.ls;
        uu[i] = q.uu;
        vu[i] = q.vu;
        if (FULL) {
            x->us[i] = q.us; x->vs[i] = q.vs; x->ls[i] = q.ls; x->uu[i] = q.uu; x->vu[i] = q.vu;
        }
    }
    // computeTireCompressionVelocity (dof24.cpp:1122-1151).  Each state is advanced (the
    // half-implicit update, dof24_halfImplicit.cpp:214-281) as soon as nothing later in the step reads
    // its old value, with the same operations as an update at the end of the step.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double dxt = k.vsz[i] - (cth * (s.wu[i] * cph + vu[i] * sph) - uu[i] * sth);
        if (FULL) x->dxt[i] = dxt;
        if (!NR) {
            s.xe[i] = s.xe[i] + k.xedot[i] * dt;
            s.ye[i] = s.ye[i] + k.yedot[i] * dt;
        }
        s.xt[i] = s.xt[i] + dxt * dt;
    }
    // computePowertrainRHS (dof24.cpp:1338-1509)
    double dOm[4], eng[4], dcrank;
    powertrain4<STRICT>(P, u, k.tqs, k.maxBrake, throttle, braking, s.u, s.crank, s.gear, s.omega, fx, s.rStat, My, dOm, eng,
                dcrank);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (FULL) {
            x->dOmega[i] = dOm[i]; x->engTor[i] = eng[i]; x->fz[i] = fz[i]; x->My[i] = My[i];
        }
        s.omega[i] = s.omega[i] + dOm[i] * dt;
    }
    if (P.tcbool) s.crank = s.crank + dcrank * dt;
    // Per strut, front axle then rear: tireToVehTransform (dof24.cpp:459-528), computeSusRHS
    // (:1182-1229: last step's fzd and dxs, then the new dxs) and computeForcesThroughSus (:599-680;
    // the rear-right lateral force has no cos(theta) factor in the reference (:639) and keeps that
    // form here), then the axle's jacking force.  The sums of computeVehicleRHS (dof24.cpp:1682-1733;
    // the strut moments mz are zero) accumulate as the struts come, in the reference's order.
    const double sph_f = STRICT ? sin(s.phi) : sph;
    double sfxs = 0, sfys = 0, sfzs = 0, smx = 0, smy = 0;
    double tF = 0, sF = 0, tA = 0, tC = 0, tR = 0, sR = 0, tB = 0, tD = 0;
#pragma unroll
    for (int ax = 0; ax < 2; ++ax) {
        const double mu = ax == 0 ? P.muf : P.mur;
        const double hrc = ax == 0 ? P.hrcf : P.hrcr;
        double fyg_a[2], fxs_a[2], fys_a[2], fzs_a[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int i = 2 * ax + j;
            double fxi = fx[i], fyi = fy[i];
            if (ax == 0) {
                fxi = fx[i] * cd - fy[i] * sd;
                fyi = fx[i] * sd + fy[i] * cd;
            }
            const double fxg = fxi * cth - fz[i] * sth;
            const double fyg = fxi * sph * sth + fyi * cph + fz[i] * sph * cth;
            const double fzg = fxi * sth * cph - fyi * sph + fz[i] * cph * cth;
            const double dwu = k.inv_mu[ax] * (cph * (cth * (fz[i] - k.mu_g[ax]) + sth * fxi) - sph * fyi - s.fzd[i] -
                                               s.xs[i] * P.ks - s.dxs[i] * P.bs - mu * (vu[i] * s.wx - uu[i] * s.wy));
            double dus, dvs, ws;
            if (i == 0) {
                dus = (-cf * s.wzdot) / 2 + s.udot;
                dvs = a * s.wzdot + s.vdot;
                ws = (cf * s.wx / 2.) - (a * s.wy) + s.w;
            } else if (i == 1) {
                dus = (cf * s.wzdot) / 2 + s.udot;
                dvs = a * s.wzdot + s.vdot;
                ws = (-cf * s.wx / 2.) - (a * s.wy) + s.w;
            } else if (i == 2) {
                dus = (-cr * s.wzdot) / 2 + s.udot;
                dvs = -b * s.wzdot + s.vdot;
                ws = (cr * s.wx / 2.) + b * s.wy + s.w;
            } else {
                dus = (cr * s.wzdot) / 2 + s.udot;
                dvs = (-b * s.wzdot) + s.vdot;
                ws = (-cr * s.wx / 2.) + b * s.wy + s.w;
            }
            const double duu = dus - (-s.dxs[i] * s.wy + ls[i] * s.wydot);
            const double dvu = dvs + (-s.dxs[i] * s.wx + ls[i] * s.wxdot);
            s.dxs[i] = -ws + s.wu[i];
            const double fxs = fxg + k.mu_g[ax] * sth - mu * duu + mu * s.wz * vu[i] - mu * s.wy * s.wu[i];
            const double fys = i < 3 ? fyg - k.mu_g[ax] * sph_f * cth - mu * dvu + mu * s.wx * s.wu[i] - mu * s.wz * uu[i]
                                     : fyg - k.mu_g[ax] * sph_f - mu * dvu + mu * s.wx * s.wu[i] - mu * s.wz * uu[i];
            const double fzs = s.xs[i] * P.ks + s.dxs[i] * P.bs;
            const double mx = fys * hrc;
            const double my = -(fxg * s.rStat[i] + fxs * ls[i]);
            s.wu[i] = s.wu[i] + dwu * dt;
            s.xs[i] = s.xs[i] + s.dxs[i] * dt;
            sfxs = i == 0 ? fxs : sfxs + fxs;
            sfys = i == 0 ? fys : sfys + fys;
            sfzs = i == 0 ? fzs : sfzs + fzs;
            smx = i == 0 ? mx : smx + mx;
            smy = i == 0 ? my : smy + my;
            fyg_a[j] = fyg;
            fxs_a[j] = fxs;
            fys_a[j] = fys;
            fzs_a[j] = fzs;
            if (FULL) {
                x->fx[i] = fxi; x->fy[i] = fyi; x->fxg[i] = fxg; x->fyg[i] = fyg; x->fzg[i] = fzg;
                x->ws[i] = ws; x->dus[i] = dus; x->dvs[i] = dvs; x->duu[i] = duu; x->dvu[i] = dvu; x->dwu[i] = dwu;
                x->fxs[i] = fxs; x->fys[i] = fys; x->fzs[i] = fzs; x->mx[i] = mx; x->my[i] = my;
            }
        }
        // the axle's jacking force, from this step's static radii
        const int i0 = 2 * ax, i1 = 2 * ax + 1;
        const double jk = fyg_a[1] * s.rStat[i1] + fys_a[1] * ls[i1] + fyg_a[0] * s.rStat[i0] + fys_a[0] * ls[i0] -
                          (fys_a[1] + fys_a[0]) * hrc;
        if (ax == 0) {
            s.fzd[1] = STRICT ? qdiv(jk, cf) : jk * u.inv_cf;
            s.fzd[0] = -s.fzd[1];
            tF = (fzs_a[0] - fzs_a[1]) * k.half_cf;
            sF = fzs_a[0] + fzs_a[1];
            tA = (fys_a[0] + fys_a[1]) * a;
            tC = (-fxs_a[0] + fxs_a[1]) * k.half_cf;
        } else {
            s.fzd[3] = STRICT ? qdiv(jk, cr) : jk * u.inv_cr;
            s.fzd[2] = -s.fzd[3];
            tR = (fzs_a[0] - fzs_a[1]) * k.half_cr;
            sR = fzs_a[0] + fzs_a[1];
            tB = (fys_a[0] + fys_a[1]) * b;
            tD = (-fxs_a[0] + fxs_a[1]) * k.half_cr;
        }
    }
    const double udot = s.wz * s.v - s.wy * s.w + k.inv_m * sfxs + kG * sth;
    const double vdot = s.wx * s.w - s.wz * s.u + k.inv_m * sfys - kG * sph * cth;
    const double wdot = s.wy * s.u - s.wx * s.v + k.inv_m * (sfzs + s.fzd[0] + s.fzd[1] + s.fzd[2] + s.fzd[3]) -
                        kG * cph * cth;
    const double wxdot = k.inv_jx * (smx + tF + tR);
    const double wydot = k.inv_jy * (smy + sR * b - sF * a);
    const double wzdot = k.inv_jz * (0.0 + tA - tB + tC + tD);
    s.u = s.u + udot * dt;
    s.v = s.v + vdot * dt;
    s.w = s.w + wdot * dt;
    s.wx = s.wx + wxdot * dt;
    s.wy = s.wy + wydot * dt;
    s.wz = s.wz + wzdot * dt;
    // Euler-angle rates at the step's (old) theta and phi: tan(theta) and 1/cos(theta) from the
    // step's sine and cosine (same values, other rounding than tan() and two divisions); the fast
    // form's 1/cos(theta) is cos(phi) / (cos(theta) cos(phi)), from the tires' reciprocal
    const double rcth = cph * rcc;
    const double tth = STRICT ? tan(s.theta) : sth * rcth;
    const double dtheta = s.wy * cph - s.wz * sph;
    const double cth_e = STRICT ? cos(s.theta) : cth;
    const double dpsi =
        STRICT ? qdiv(s.wy * sph, cth_e) + qdiv(s.wz * cph, cth_e) : (s.wy * sph * rcth) + (s.wz * cph * rcth);
    const double dphi = s.wx + s.wy * sph * tth + s.wz * cph * tth;
    s.theta = s.theta + dtheta * dt;
    s.psi = s.psi + dpsi * dt;
    s.phi = s.phi + dphi * dt;
    s.udot = udot;
    s.vdot = vdot;
    s.wdot = wdot;
    s.wxdot = wxdot;
    s.wydot = wydot;
    s.wzdot = wzdot;
    if (POS) position24<STRICT>(s, dt);
    if (FULL) {
        x->dtheta = dtheta;
        x->dphi = dphi;
        x->dpsi = dpsi;
        x->dcrank = dcrank;
    }
}

}  // namespace dev
}  // namespace lfdm
