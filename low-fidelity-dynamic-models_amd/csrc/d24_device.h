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
// (read by computeSusRHS, :1199-1228) plus the constants xsi, lsi and the tire's vsz set by
// Initialize; chassis velocities and angles with the accelerations of the last step (read by
// vehToSusTransform, :199-229).
struct S24 {
    double xe[4], ye[4], xedot[4], yedot[4], xt[4], omega[4], rStat[4];
    double xs[4], wu[4], dxs[4], fzd[4], xsi[4], lsi[4], vsz[4];
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

// Per-vehicle constants.
struct K24 {
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

// One half-implicit step (rhsFun order, dof24_halfImplicit.cpp:594-706; update :188-297).
// STRICT: the reference's own quotients and transcendentals throughout (see dev_tire.h): the static
// radius divided by cos(theta) cos(phi), the jacking forces by the track widths, the Euler-angle
// rates by cos(theta) with tan(theta) from the libm.  Where a reference function uses only the sine
// or only the cosine of an angle (computeTireRHS: cos theta and cos phi; computeForcesThroughSus:
// sin phi; the Euler-angle rates of the step: cos theta), the strict step calls that function alone;
// everywhere else the reference's sin and cos of one angle compile to one sincos, as here.
template <bool NR, bool FULL, bool STRICT = false>
LFDM_HD void step24(const PSet& P, const Derived& u, const K24& k, S24& s, double steering,
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
    // vehToSusTransform (dof24.cpp:148-230)
    double us[4], vs[4], ws[4], ls[4], uu[4], vu[4], dus[4], dvs[4], duu[4], dvu[4];
    us[0] = (-cf * s.wz / 2.) + s.u;
    vs[0] = a * s.wz + s.v;
    ws[0] = (cf * s.wx / 2.) - (a * s.wy) + s.w;
    us[1] = (cf * s.wz / 2.) + s.u;
    vs[1] = a * s.wz + s.v;
    ws[1] = (-cf * s.wx / 2.) - (a * s.wy) + s.w;
    us[2] = (-cr * s.wz / 2.) + s.u;
    vs[2] = -b * s.wz + s.v;
    ws[2] = (cr * s.wx / 2.) + b * s.wy + s.w;
    us[3] = (cr * s.wz / 2.) + s.u;
    vs[3] = -b * s.wz + s.v;
    ws[3] = (-cr * s.wx / 2.) + b * s.wy + s.w;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        ls[i] = s.lsi[i] - (s.xs[i] - s.xsi[i]);
        uu[i] = us[i] - ls[i] * s.wy;
        vu[i] = vs[i] + ls[i] * s.wx;
    }
    dus[0] = (-cf * s.wzdot) / 2 + s.udot;
    dvs[0] = a * s.wzdot + s.vdot;
    dus[1] = (cf * s.wzdot) / 2 + s.udot;
    dvs[1] = a * s.wzdot + s.vdot;
    dus[2] = (-cr * s.wzdot) / 2 + s.udot;
    dvs[2] = -b * s.wzdot + s.vdot;
    dus[3] = (cr * s.wzdot) / 2 + s.udot;
    dvs[3] = (-b * s.wzdot) + s.vdot;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        duu[i] = dus[i] - (-s.dxs[i] * s.wy + ls[i] * s.wydot);
        dvu[i] = dvs[i] + (-s.dxs[i] * s.wx + ls[i] * s.wxdot);
    }
    // vehToTireTransform (dof24.cpp:313-385) with last step's static radius
    const double delta = P.nonLinearSteer ? map_y<STRICT>(P.steer, steering) : steering * P.maxSteer;
    double sd, cd;
    if (STRICT) sincos(delta, &sd, &cd);
    else sincos_fast(delta, sd, cd);
    double vgx[4], vgy[4], vsx[4], vsy[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        vgx[i] = cth * (uu[i] - s.wy * s.rStat[i]) + sth * (s.wu[i] * cph + sph * (s.wx * s.rStat[i] + vu[i]));
        vgy[i] = cph * (vu[i] + s.wx * s.rStat[i]) - s.wu[i] * sph;
        vsy[i] = vgy[i];
        vsx[i] = i < 2 ? vgx[i] * cd + vgy[i] * sd : vgx[i];
    }
    // computeTireRHS (dof24.cpp:827-1121): load from the compression, radius from the attitude
    double fx[4], fy[4], fz[4], My[4];
    const double cc = cth * cph;
    const double rcc = frcp(cc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
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
            tmeasy_core<STRICT>(t, u, k.qy1, k.qy2, dl, d, fz[i], s.rStat[i], vsx[i], vsy[i], s.xe[i], s.ye[i],
                                s.omega[i], s.xedot[i], s.yedot[i], fx[i], fy[i], My[i]);
        else
            tmeasy_nr_core<STRICT>(t, u, k.qy1, k.qy2, dl, d, fz[i], s.rStat[i], vsx[i], vsy[i], s.omega[i], fx[i],
                                   fy[i], My[i]);
    }
    // computeTireCompressionVelocity (dof24.cpp:1122-1151)
    double dxt[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) dxt[i] = s.vsz[i] - (cth * (s.wu[i] * cph + vu[i] * sph) - uu[i] * sth);
    // computePowertrainRHS (dof24.cpp:1338-1509)
    double dOm[4], eng[4], dcrank;
    powertrain4<STRICT>(P, u, k.tqs, k.maxBrake, throttle, braking, s.u, s.crank, s.gear, s.omega, fx, s.rStat, My, dOm, eng,
                dcrank);
    // tireToVehTransform (dof24.cpp:459-528)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double gx = fx[i] * cd - fy[i] * sd;
        const double gy = fx[i] * sd + fy[i] * cd;
        fx[i] = gx;
        fy[i] = gy;
    }
    double fxg[4], fyg[4], fzg[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        fxg[i] = fx[i] * cth - fz[i] * sth;
        fyg[i] = fx[i] * sph * sth + fy[i] * cph + fz[i] * sph * cth;
        fzg[i] = fx[i] * sth * cph - fy[i] * sph + fz[i] * cph * cth;
    }
    // computeSusRHS (dof24.cpp:1182-1229): last step's fzd and dxs, then the new dxs
    double dwu[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int ax = i < 2 ? 0 : 1;
        const double mu = ax == 0 ? P.muf : P.mur;
        dwu[i] = k.inv_mu[ax] * (cph * (cth * (fz[i] - k.mu_g[ax]) + sth * fx[i]) - sph * fy[i] - s.fzd[i] -
                                 s.xs[i] * P.ks - s.dxs[i] * P.bs - mu * (vu[i] * s.wx - uu[i] * s.wy));
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) s.dxs[i] = -ws[i] + s.wu[i];
    // computeForcesThroughSus (dof24.cpp:599-680); the rear-right lateral force has no cos(theta)
    // factor in the reference (:639) and keeps that form here
    double fxs[4], fys[4], fzs[4];
    const double sph_f = STRICT ? sin(s.phi) : sph;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int ax = i < 2 ? 0 : 1;
        const double mu = ax == 0 ? P.muf : P.mur;
        fxs[i] = fxg[i] + k.mu_g[ax] * sth - mu * duu[i] + mu * s.wz * vu[i] - mu * s.wy * s.wu[i];
        if (i < 3)
            fys[i] = fyg[i] - k.mu_g[ax] * sph_f * cth - mu * dvu[i] + mu * s.wx * s.wu[i] - mu * s.wz * uu[i];
        else
            fys[i] = fyg[i] - k.mu_g[ax] * sph_f - mu * dvu[i] + mu * s.wx * s.wu[i] - mu * s.wz * uu[i];
        fzs[i] = s.xs[i] * P.ks + s.dxs[i] * P.bs;
    }
    const double jf = fyg[1] * s.rStat[1] + fys[1] * ls[1] + fyg[0] * s.rStat[0] + fys[0] * ls[0] -
                      (fys[1] + fys[0]) * P.hrcf;
    s.fzd[1] = STRICT ? qdiv(jf, cf) : jf * u.inv_cf;
    s.fzd[0] = -s.fzd[1];
    const double jr = fyg[3] * s.rStat[3] + fys[3] * ls[3] + fyg[2] * s.rStat[2] + fys[2] * ls[2] -
                      (fys[3] + fys[2]) * P.hrcr;
    s.fzd[3] = STRICT ? qdiv(jr, cr) : jr * u.inv_cr;
    s.fzd[2] = -s.fzd[3];
    double mx[4], my[4];
    mx[0] = fys[0] * P.hrcf;
    mx[1] = fys[1] * P.hrcf;
    mx[2] = fys[2] * P.hrcr;
    mx[3] = fys[3] * P.hrcr;
#pragma unroll
    for (int i = 0; i < 4; ++i) my[i] = -(fxg[i] * s.rStat[i] + fxs[i] * ls[i]);
    // computeVehicleRHS (dof24.cpp:1682-1733); the strut moments mz are zero
    const double udot = s.wz * s.v - s.wy * s.w + k.inv_m * (fxs[0] + fxs[1] + fxs[2] + fxs[3]) + kG * sth;
    const double vdot = s.wx * s.w - s.wz * s.u + k.inv_m * (fys[0] + fys[1] + fys[2] + fys[3]) - kG * sph * cth;
    const double wdot =
        s.wy * s.u - s.wx * s.v +
        k.inv_m * (fzs[0] + fzs[1] + fzs[2] + fzs[3] + s.fzd[0] + s.fzd[1] + s.fzd[2] + s.fzd[3]) -
        kG * cph * cth;
    const double wxdot =
        k.inv_jx * ((mx[0] + mx[1] + mx[2] + mx[3]) + (fzs[0] - fzs[1]) * k.half_cf + (fzs[2] - fzs[3]) * k.half_cr);
    const double wydot = k.inv_jy * ((my[0] + my[1] + my[2] + my[3]) + (fzs[2] + fzs[3]) * b - (fzs[0] + fzs[1]) * a);
    const double wzdot = k.inv_jz * (0.0 + (fys[0] + fys[1]) * a - (fys[2] + fys[3]) * b +
                                     (-fxs[0] + fxs[1]) * k.half_cf + (-fxs[2] + fxs[3]) * k.half_cr);
    // Half-implicit update (dof24_halfImplicit.cpp:214-281)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!NR) {
            s.xe[i] = s.xe[i] + s.xedot[i] * dt;
            s.ye[i] = s.ye[i] + s.yedot[i] * dt;
        }
        s.xt[i] = s.xt[i] + dxt[i] * dt;
        s.omega[i] = s.omega[i] + dOm[i] * dt;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        s.wu[i] = s.wu[i] + dwu[i] * dt;
        s.xs[i] = s.xs[i] + s.dxs[i] * dt;
    }
    if (P.tcbool) s.crank = s.crank + dcrank * dt;
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
    double sp, cp;
    if (STRICT) sincos(s.psi, &sp, &cp);
    else sincos_fast(s.psi, sp, cp);
    s.x = s.x + (s.u * cp - s.v * sp) * dt;
    s.y = s.y + (s.u * sp + s.v * cp) * dt;
    s.udot = udot;
    s.vdot = vdot;
    s.wdot = wdot;
    s.wxdot = wxdot;
    s.wydot = wydot;
    s.wzdot = wzdot;
    if (FULL) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            x->dOmega[i] = dOm[i]; x->dxt[i] = dxt[i]; x->fx[i] = fx[i]; x->fy[i] = fy[i]; x->fz[i] = fz[i];
            x->fxg[i] = fxg[i]; x->fyg[i] = fyg[i]; x->fzg[i] = fzg[i]; x->vgx[i] = vgx[i]; x->vgy[i] = vgy[i];
            x->vsx[i] = vsx[i]; x->vsy[i] = vsy[i]; x->My[i] = My[i]; x->engTor[i] = eng[i];
            x->ls[i] = ls[i]; x->us[i] = us[i]; x->vs[i] = vs[i]; x->ws[i] = ws[i]; x->dus[i] = dus[i];
            x->dvs[i] = dvs[i]; x->uu[i] = uu[i]; x->vu[i] = vu[i]; x->duu[i] = duu[i]; x->dvu[i] = dvu[i];
            x->dwu[i] = dwu[i]; x->fxs[i] = fxs[i]; x->fys[i] = fys[i]; x->fzs[i] = fzs[i]; x->mx[i] = mx[i];
            x->my[i] = my[i];
        }
        x->dtheta = dtheta;
        x->dphi = dphi;
        x->dpsi = dpsi;
        x->dcrank = dcrank;
    }
}

}  // namespace dev
}  // namespace lfdm
