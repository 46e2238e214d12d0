// 18-DOF vehicle (TMeasy / TMeasyNr tires) as register-resident device code.
//
// Follows wheeled_vehicle_models/18dof/dof18.cpp (physics) and dof18_halfImplicit.cpp:245-318
// (step) with the same IEEE operation order.  What differs from the reference is only where
// work happens: parameter-only subexpressions are evaluated once per launch (same operands and
// order, hence the same bits), the steering angle's sine/cosine once per step instead of eight
// times, and the state lives in VGPRs across all steps of a launch.
#pragma once
#include "dev_common.h"

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
    double neg_mt, E3c, hm, A2, A3, inv_mt, two_hrc_m, denom, A1A2, A1A3, A3sq;
    double qy1, qy2;  // InterpQ coefficients of dfy0 (2 w1 - w2/2, w1 - w2/2)
};

__device__ __forceinline__ void k18_init(const P18& P, double m, double jz, double jx, double dfy0Pn, double dfy0P2n,
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
    k.denom = (A2 * (A1 * A1) - 2. * A1 * A3 * P.jxz + jz * (A3 * A3) + mt * (P.jxz * P.jxz) - A2 * jz * mt);
    k.A1A2 = A1 * A2;
    k.A1A3 = A1 * A3;
    k.A3sq = A3 * A3;
    k.qy1 = 2. * dfy0Pn - 0.5 * dfy0P2n;
    k.qy2 = dfy0Pn - 0.5 * dfy0P2n;
}

// tmxy_combined (dof18.cpp:247-288); returns fos, and f through `f`.
__device__ __forceinline__ double tmxy_fos(double s, double df0, double sm, double fm, double ss, double fs,
                                           double& f) {
    double df0loc = 0.0;
    if (sm > 0.0) {
        const double c = 2.0 * fm / sm;
        df0loc = (c < df0) ? df0 : c;
    }
    double fos;
    if (s > 0.0 && df0loc > 0.0) {
        if (s > ss) {
            f = fs;
            fos = f / s;
        } else if (s < sm) {
            const double p = df0loc * sm / fm - 2.0;
            const double sn = s / sm;
            const double dn = 1.0 + (sn + p) * sn;
            f = df0loc * sm * sn / dn;
            fos = df0loc / dn;
        } else {
            const double q = fm / sm;
            const double a = q * q / (df0loc * sm);
            const double sstar = sm + (fm - fs) / (a * (ss - sm));
            if (sstar <= ss) {
                if (s <= sstar) {
                    f = fm - a * (s - sm) * (s - sm);
                } else {
                    const double b = a * (sstar - sm) / (ss - sstar);
                    f = fs + b * (ss - s) * (ss - s);
                }
            } else {
                const double sn = (s - sm) / (ss - sm);
                f = fm - (fm - fs) * sn * sn * (3.0 - 2.0 * sn);
            }
            fos = f / s;
        }
    } else {
        f = 0.0;
        fos = 0.0;
    }
    return fos;
}

// Load-dependent curve parameters shared by both tire models.
struct Curve {
    double dfx0, dfy0, fxm, fym, fxs, fys, sxm, sym, sxs, sys;
};

__device__ __forceinline__ void curve_at(const TireP& t, const D18& u, const K18& k, double fz, Curve& c) {
    const double fzn = fz / t.pn;
    const double fzn1 = fzn - 1.;
    // InterpQ(fz, w1, w2, pn) = (fz/pn) * (2 w1 - 0.5 w2 - (w1 - 0.5 w2) (fz/pn)), utils.h:98-100
    c.dfx0 = fzn * (u.qx1 - u.qx2 * fzn);
    c.dfy0 = fzn * (k.qy1 - k.qy2 * fzn);
    c.fxm = fzn * (u.qfxm1 - u.qfxm2 * fzn);
    c.fym = fzn * (u.qfym1 - u.qfym2 * fzn);
    c.fxs = fzn * (u.qfxs1 - u.qfxs2 * fzn);
    c.fys = fzn * (u.qfys1 - u.qfys2 * fzn);
    // InterpL(fz, w1, w2, pn) = w1 + (w2 - w1) (fz/pn - 1), utils.h:93-95
    c.sxm = t.sxmPn + u.lsxm * fzn1;
    c.sym = t.symPn + u.lsym * fzn1;
    c.sxs = t.sxsPn + u.lsxs * fzn1;
    c.sys = t.sysPn + u.lsys * fzn1;
}

// Tire RHS, TMeasy (dof18.cpp:380-539).  Tire i of the carried state; writes xedot/yedot into the
// carried state and the forces and kinematics into the locals passed by reference.
__device__ __forceinline__ void tire_tmeasy(const TireP& t, const D18& u, const K18& k, double delta, double fz,
                                            double vsx, double vsy, double& xe, double& ye, double& xedot,
                                            double& yedot, double& omega, double& fx, double& fy, double& My,
                                            double& xt, double& rStat) {
    if (fz <= 0) {
        xe = 0; ye = 0; omega = 0; xt = 0; rStat = t.r0; fx = 0; fy = 0; My = 0;
        return;
    }
    xt = fz / t.kt;
    rStat = t.r0 - xt;
    double rdynco;
    if (fz <= t.fzRdynco)
        rdynco = t.rdyncoPn + u.lrd * (fz / t.pn - 1.);
    else
        rdynco = t.rdyncoCrit;
    const double r_eff = rdynco * t.r0 + (1. - rdynco) * rStat;
    vsx = vsx - (omega * r_eff);
    const double rw = r_eff * fabs(omega);
    const double vta = rw + 0.01;
    const double sx = -vsx / vta;
    const double alpha = atan2(vsy, vta) - delta;
    const double sy = -tan(alpha);
    if (fz > t.pnmax) fz = t.pnmax;
    Curve c;
    curve_at(t, u, k, fz, c);
    const double rx = c.fxm / c.dfx0, ry = c.fym / c.dfy0;
    const double ssum = c.sxm + c.sym, rsum = rx + ry;
    const double hsxn = c.sxm / ssum + rx / rsum;
    const double hsyn = c.sym / ssum + ry / rsum;
    const double sxn = sx / hsxn;
    const double syn = sy / hsyn;
    const double sc = hypot(sxn, syn);
    double ca, sa;
    if (sc > 0) {
        ca = sxn / sc;
        sa = syn / sc;
    } else {
        ca = kHalfSqrt2;
        sa = kHalfSqrt2;
    }
    const double df0 = hypot(c.dfx0 * ca * hsxn, c.dfy0 * sa * hsyn);
    const double fm = hypot(c.fxm * ca, c.fym * sa);
    const double sm = hypot(c.sxm * ca / hsxn, c.sym * sa / hsyn);
    const double fs = hypot(c.fxs * ca, c.fys * sa);
    const double ss = hypot(c.sxs * ca / hsxn, c.sys * sa / hsyn);
    double f;
    const double fos = tmxy_fos(sc, df0, sm, fm, ss, fs, f);
    // -sineStep(vta, 0, 0, 0, 1) * rr * fz * rStat * sgn(omega); the step is 0 for vta <= 0, else 1
    const double ssv = (vta <= 0.0) ? 0.0 : 1.0;
    My = -ssv * t.rr * fz * rStat * sgn(omega);
    const double vtxs = rw * hsxn + 0.01;
    const double vtys = rw * hsyn + 0.01;
    const double vy = -sy * vta;
    const double numx = -vtxs * t.cx * xe - fos * vsx;
    const double denx = vtxs * t.dx + fos;
    const double numy = -vtys * t.cy * ye - fos * vy;
    const double deny = vtys * t.dy + fos;
    xedot = numx / denx;
    yedot = numy / deny;
    const double fxdyn = t.dx * numx / denx + t.cx * xe;
    const double fydyn = t.dy * yedot + (t.cy * ye);
    const double fxstr = clampd(xe * t.cx + xedot * t.dx, -t.fxmP2n, t.fxmP2n);
    const double fystr = clampd(ye * t.cy + yedot * t.dy, -t.fymP2n, t.fymP2n);
    const double wx = sine_step_weight(fabs(vsx));
    const double wy = sine_step_weight(fabs(vy));
    fx = wx * fxstr + (1. - wx) * fxdyn;
    fy = wy * fystr + (1. - wy) * fydyn;
}

// Tire RHS, TMeasyNr (dof18.cpp:541-673) with the Coulomb blend (:290-306).
__device__ __forceinline__ void tire_tmeasy_nr(const TireP& t, const D18& u, const K18& k, double delta, double fz,
                                               double vsx, double vsy, double& omega, double& fx, double& fy,
                                               double& My, double& xt, double& rStat) {
    if (fz <= 0) {
        omega = 0; xt = 0; rStat = t.r0; fx = 0; fy = 0; My = 0;
        return;
    }
    const double vsx_in = vsx;
    xt = fz / t.kt;
    rStat = t.r0 - xt;
    double rdynco;
    if (fz <= t.fzRdynco)
        rdynco = t.rdyncoPn + u.lrd * (fz / t.pn - 1.);
    else
        rdynco = t.rdyncoCrit;
    const double r_eff = rdynco * t.r0 + (1. - rdynco) * rStat;
    vsx = vsx - (omega * r_eff);
    const double vta = r_eff * fabs(omega) + 0.01;
    double Fx0 = tanh(-2.0 * vsx / t.vcoulomb) * fz * t.mu;
    double Fy0 = tanh(-2.0 * vsy / t.vcoulomb) * fz * t.mu;
    if (hypot(Fx0, Fy0) > fz * t.mu) {
        const double fsc = fz * t.mu / hypot(Fx0, Fy0);
        Fx0 = Fx0 * fsc;
        Fy0 = Fy0 * fsc;
    }
    const double sx = -vsx / vta;
    const double alpha = atan2(vsy, vta) - delta;
    const double sy = -tan(alpha);
    if (fz > t.pnmax) fz = t.pnmax;
    Curve c;
    curve_at(t, u, k, fz, c);
    const double frblend = sine_step(fabs(vsx_in), t.frblend_begin, 0., t.frblend_end, 1.);
    const double sc = hypot(sx, sy);
    double ca, sa;
    if (sc > 0) {
        ca = sx / sc;
        sa = sy / sc;
    } else {
        ca = kHalfSqrt2;
        sa = kHalfSqrt2;
    }
    const double df0 = hypot(c.dfx0 * ca, c.dfy0 * sa);
    const double fm = t.mu * hypot(c.fxm * ca, c.fym * sa);
    const double sm = t.mu * hypot(c.sxm * ca, c.sym * sa);
    const double fs = t.mu * hypot(c.fxs * ca, c.fys * sa);
    const double ss = t.mu * hypot(c.sxs * ca, c.sys * sa);
    double f;
    tmxy_fos(sc, df0, sm, fm, ss, fs, f);
    double Fx, Fy;
    if (sc > 0.) {
        Fx = f * sx / sc;
        Fy = f * sy / sc;
    } else {
        Fx = 0.;
        Fy = 0.;
    }
    fx = (1.0 - frblend) * Fx0 + frblend * Fx;
    fy = (1.0 - frblend) * Fy0 + frblend * Fy;
    My = -t.rr * fz * rStat * tanh(omega);
}

// differentialSplit (dof18.cpp:46-73) with max_bias = 2.
__device__ __forceinline__ void diff_split(double torque, double sl, double sr, double& tl, double& tr) {
    const double max_bias = 2;
    const double diff = fabs(sl - sr);
    double bias = 1;
    if (diff > 0.5)
        bias = max_bias;
    else if (diff > 0.25)
        bias = 4 * (max_bias - 1) * diff + (2 - max_bias);
    const double alpha = bias / (1 + bias);
    const double slow = alpha * torque;
    const double fast = torque - slow;
    const bool left_slow = fabs(sl) < fabs(sr);
    tl = left_slow ? slow : fast;
    tr = left_slow ? fast : slow;
}

// driveTorque (dof18.cpp:15-42) with the per-vehicle torque scale on the map ordinates.
__device__ __forceinline__ double drive_torque(const P18& P, const K18& k, double throttle, double w) {
    if (P.throttleMod) {
        const double tq = map_y_scaled(P.torque, w, throttle, k.tqs, throttle);
        return tq + map_y(P.losses, w);
    }
    const double tq = map_y_scaled(P.torque, w, 1.0, k.tqs, 1.0);
    return tq * throttle + map_y(P.losses, w);
}

// One half-implicit step (rhsFun order, dof18_halfImplicit.cpp:618-673; update :245-318).
template <bool NR, bool FULL>
__device__ __forceinline__ void step18(const P18& P, const D18& u, const K18& k, S18& s, double steering,
                                       double throttle, double braking, double dt, X18* x) {
    const TireP& t = P.tire;
    // computeTireLoads
    const double lat = s.vdot + s.wz * s.u;
    const double Z2f = k.Z2fc * lat;
    const double Z3f = (P.krof * s.phi + P.brof * s.wx) / P.cf;
    const double Z4 = (k.Z4n * (s.udot - s.wz * s.v)) / u.two_ab;
    const double Z2r = k.Z2rc * lat;
    const double Z3r = (P.kror * s.phi + P.bror * s.wx) / P.cr;
    double fz[4];
    fz[0] = k.Z1f - Z2f - Z3f - Z4;
    fz[1] = k.Z1f + Z2f + Z3f - Z4;
    fz[2] = k.Z1r - Z2r - Z3r + Z4;
    fz[3] = k.Z1r + Z2r + Z3r + Z4;
#pragma unroll
    for (int i = 0; i < 4; ++i) fz[i] = fz[i] > 0. ? fz[i] : 0.;
    // vehToTireTransform
    const double delta = P.nonLinearSteer ? map_y(P.steer, steering) : steering * P.maxSteer;
    double sd, cd;
    sincos(delta, &sd, &cd);
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
        const double d = i < 2 ? delta : u.delta_rear;
        if (!NR)
            tire_tmeasy(t, u, k, d, fz[i], vsx[i], vsy[i], s.xe[i], s.ye[i], s.xedot[i], s.yedot[i], s.omega[i],
                        fx[i], fy[i], My[i], xt[i], rStat[i]);
        else
            tire_tmeasy_nr(t, u, k, d, fz[i], vsx[i], vsy[i], s.omega[i], fx[i], fy[i], My[i], xt[i], rStat[i]);
    }
    // computePowertrainRHS
    double torque_t;
    double dcrank = 0.0;
    const double osum = s.omega[0] + s.omega[1] + s.omega[2] + s.omega[3];
    const double gr = pick(P.gears, P.ngears, s.gear);
    double shift_w;
    if (P.tcbool) {
        const double omega_out = 0.25 * osum / gr;
        const double omega_in = s.crank;
        double sr = 0;
        bool rev = false;
        if (!(omega_out < 1e-9 || omega_in < 1e-9)) {
            sr = omega_out / omega_in;
            if (sr > 1.) {
                sr = 1. - (sr - 1.);
                rev = true;
            }
            if (sr < 0) sr = 0;
        }
        const double cf = map_y(P.cfmap, sr);
        const double tr = map_y(P.trmap, sr);
        const double q = omega_in / cf;
        double torque_in = -(q * q);
        if (rev) torque_in = -torque_in;
        const double torque_out = rev ? -torque_in : -tr * torque_in;
        torque_t = torque_out / gr;
        if (s.u < 1e-9 && torque_t < 0) torque_t = 0;
        dcrank = u.inv_ci * (drive_torque(P, k, throttle, s.crank) + torque_in);
        shift_w = omega_out;
    } else {
        s.crank = 0.25 * osum / gr;
        torque_t = drive_torque(P, k, throttle, s.crank) / gr;
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
    double tf, trr;
    if (P.driveType) {
        tf = torque_t * 0.5;
        trr = torque_t * 0.5;
    } else if (P.whichWheels) {
        tf = 0;
        trr = torque_t;
    } else {
        tf = torque_t;
        trr = 0;
    }
    double eng[4];
    diff_split(tf, s.omega[0], s.omega[1], eng[0], eng[1]);
    diff_split(trr, s.omega[2], s.omega[3], eng[2], eng[3]);
    const double bt = k.maxBrake * braking;
    double dOm[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        dOm[i] = u.inv_jw * (eng[i] + My[i] - sgn(s.omega[i]) * bt - fx[i] * rStat[i]);
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
    const double vdot = (E1 * u.jxz2 - k.A1A2 * E2 + A1 * E3 * jxz + A3 * E2 * jxz - A2 * E1 * jz - A3 * E3 * jz) /
                        k.denom;
    const double wxdot =
        (u.A1sq * E3 - k.A1A3 * E2 + A1 * E1 * jxz - A3E1 * jz + E2 * jxz * mt - E3 * jz * mt) / k.denom;
    const double wzdot =
        (k.A3sq * E2 - k.A1A2 * E1 - k.A1A3 * E3 + A3E1 * jxz - A2 * E2 * mt + E3 * jxz * mt) / k.denom;
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
    double sp, cp;
    sincos(s.psi, &sp, &cp);
    s.x = s.x + (s.u * cp - s.v * sp) * dt;
    s.y = s.y + (s.u * sp + s.v * cp) * dt;
    s.psi = s.psi + s.wz * dt;
    s.phi = s.phi + s.wx * dt;
    s.udot = udot;
    s.vdot = vdot;
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
