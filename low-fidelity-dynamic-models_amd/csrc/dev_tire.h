// Tire force cores and the four-wheel powertrain shared by the 11-, 18- and 24-DOF device models.
//
// The reference repeats these functions verbatim in dof11.cpp, dof18.cpp and dof24.cpp; here they
// exist once.  The IEEE operation order is the reference's (the library is built with
// -ffp-contract=off).
#pragma once
#include "dev_common.h"

namespace lfdm {
namespace dev {

// tmxy_combined (dof18.cpp:247-288); returns fos, and f through `f`.  STRICT divides as the
// reference writes it.  The fast form takes the reciprocals rs = 1 / s, rdf0 = 1 / df0, rfm = 1 / fm
// and rsm = 1 / sm from the caller (the tire core forms each with its square root), so that the
// quotients by s, sm, fm and df0loc become products, and divides by the adhesion branch's denominator
// once, as a reciprocal.  The strict form ignores them.
template <bool STRICT = false>
LFDM_HD double tmxy_fos(double s, double df0, double sm, double fm, double ss, double fs, double rs, double rdf0,
                        double rfm, double rsm, double& f) {
    double df0loc = 0.0, rdf0loc = 0.0;
    if (sm > 0.0) {
        const double c = STRICT ? qdiv(2.0 * fm, sm) : 2.0 * fm * rsm;
        df0loc = (c < df0) ? df0 : c;
        rdf0loc = (c < df0) ? rdf0 : 0.5 * sm * rfm;  // 1 / df0loc
    }
    double fos;
    if (s > 0.0 && df0loc > 0.0) {
        if (s > ss) {
            f = fs;
            fos = STRICT ? qdiv(f, s) : f * rs;
        } else if (s < sm) {
            const double p = (STRICT ? qdiv(df0loc * sm, fm) : df0loc * sm * rfm) - 2.0;
            const double sn = STRICT ? qdiv(s, sm) : s * rsm;
            const double dn = 1.0 + (sn + p) * sn;
            if (STRICT) {
                f = xdiv<STRICT>(df0loc * sm * sn, dn);
                fos = xdiv<STRICT>(df0loc, dn);
            } else {
                const double rdn = frcp(dn);
                f = df0loc * sm * sn * rdn;
                fos = df0loc * rdn;
            }
        } else {
            const double q = STRICT ? qdiv(fm, sm) : fm * rsm;
            const double a = STRICT ? qdiv(q * q, df0loc * sm) : q * q * rdf0loc * rsm;
            const double sstar = sm + xdiv<STRICT>(fm - fs, a * (ss - sm));
            if (sstar <= ss) {
                if (s <= sstar) {
                    f = fm - a * (s - sm) * (s - sm);
                } else {
                    const double b = xdiv<STRICT>(a * (sstar - sm), ss - sstar);
                    f = fs + b * (ss - s) * (ss - s);
                }
            } else {
                const double sn = xdiv<STRICT>(s - sm, ss - sm);
                f = fm - (fm - fs) * sn * sn * (3.0 - 2.0 * sn);
            }
            fos = STRICT ? qdiv(f, s) : f * rs;
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

// qy1, qy2: InterpQ coefficients of the (per-vehicle) initial lateral slope dfy0.  STRICT forms
// fz / pn as the reference does; otherwise fz times the set's 1 / pn.
template <bool STRICT = false>
LFDM_HD void curve_at(const TireP& t, const Derived& u, double qy1, double qy2, double fz,
                                         Curve& c) {
    const double fzn = STRICT ? qdiv(fz, t.pn) : fz * u.inv_pn;
    const double fzn1 = fzn - 1.;
    // InterpQ(fz, w1, w2, pn) = (fz/pn) * (2 w1 - 0.5 w2 - (w1 - 0.5 w2) (fz/pn)), utils.h:98-100
    c.dfx0 = fzn * (u.qx1 - u.qx2 * fzn);
    c.dfy0 = fzn * (qy1 - qy2 * fzn);
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

// Where the tire cores evaluate the reference's functions in another form (same function, other
// rounding; the differences are of the size of device libm's last-bit differences from glibc, far
// inside the 1e-9 parity bar):
//  - lateral slip sy = -tan(atan2(vsy, vta) - delta) (dof18.cpp:430-431) as the tangent difference
//    -(r - tan delta) / (1 + r tan delta), r = vsy / vta (vta > 0, so tan(atan2(vsy, vta)) = r):
//    no atan2 and no tan per tire; tan delta comes once per step from the step's sin/cos of the
//    steering angle;
//  - hypot(a, b) as sqrt(a a + b b) (hyp): the tire quantities stay hundreds of orders of magnitude
//    away from the overflow and underflow that hypot's scaling guards against;
//  - divisions by a shared denominator (vta, hsxn, hsyn, sc, sxm + sym, rx + ry, pn, kt) as
//    multiplications by its reciprocal.
// The STRICT instantiations evaluate all three as the reference writes them: atan2 and tan of the
// steering angle, the device libm's hypot, and every quotient as its own (correctly rounded)
// division, so that they differ from the reference only where the device libm and glibc round
// differently.
LFDM_HD double hyp(double a, double b) { return fsqrt(a * a + b * b); }

template <bool STRICT>
LFDM_HD double hyp2(double a, double b) {
    if (STRICT) return hypot(a, b);
    return hyp(a, b);
}

// Intermediate values of the TMeasy force core (dof18.cpp:412-539; the same code is
// dof11.cpp:335-472 and dof24.cpp:851-987) for a tire on the ground (fz > 0) with static radius
// rStat: the step reads the outputs (deflection rates, forces, rolling moment), the 18-DOF adjoint
// the rest, so both see the same values.
struct TireEval {
    double rdynco, drd, r_eff, vsx2, aom, som, rw, vta, sx, sy;
    bool clip;
    double fzc;
    Curve c;
    double fzn, rx, ry, ssum, rsum, rssum, rrsum, hsxn, hsyn, rhsxn, rhsyn, sxn, syn, sc, rsc, ca, sa;
    double A, B, C, D, E, F, G, H, I, J, df0, fm, sm, fs, ss, f, fos;
    double ssv, vtxs, vtys, vy, numx, denx, numy, deny, xedot, yedot;
    double fxdyn, fydyn, fxr, fyr, fxstr, fystr, avx, avy, wxw, wyw, fx, fy, My;
};

// vsx, vsy are the contact-patch velocities of the transform, delta the tire's steering angle and
// tan_delta its tangent (the fast form reads only the tangent, the strict one only the angle).
template <bool STRICT = false>
LFDM_HD void tmeasy_eval(const TireP& t, const Derived& u, double qy1, double qy2, double delta, double tan_delta,
                         double fz, double rStat, double vsx, double vsy, double xe, double ye, double omega,
                         TireEval& w) {
    if (fz <= t.fzRdynco) {
        w.rdynco = t.rdyncoPn + u.lrd * ((STRICT ? qdiv(fz, t.pn) : fz * u.inv_pn) - 1.);
        w.drd = u.lrd * u.inv_pn;
    } else {
        w.rdynco = t.rdyncoCrit;
        w.drd = 0.0;
    }
    w.r_eff = w.rdynco * t.r0 + (1. - w.rdynco) * rStat;
    w.vsx2 = vsx - (omega * w.r_eff);
    w.aom = fabs(omega);
    w.som = sgn(omega);
    w.rw = w.r_eff * w.aom;
    w.vta = w.rw + 0.01;
    if (STRICT) {
        w.sx = qdiv(-w.vsx2, w.vta);
        w.sy = -tan(atan2(vsy, w.vta) - delta);
    } else {
        // sy = -(r - tan d) / (1 + r tan d), r = vsy / vta = tan(atan2(vsy, vta)), is
        // -(vsy - vta tan d) / (vta + vsy tan d); sx and sy share one reciprocal
        const double den = w.vta + vsy * tan_delta;
        const double rq = frcp(w.vta * den);
        w.sx = -w.vsx2 * den * rq;
        w.sy = -(vsy - w.vta * tan_delta) * w.vta * rq;
    }
    w.clip = fz > t.pnmax;
    w.fzc = w.clip ? t.pnmax : fz;
    const Curve& c = w.c;
    curve_at<STRICT>(t, u, qy1, qy2, w.fzc, w.c);
    w.fzn = w.fzc * u.inv_pn;
    if (STRICT) {
        w.rx = qdiv(c.fxm, c.dfx0);
        w.ry = qdiv(c.fym, c.dfy0);
    } else {
        // the fast form's pairs of quotients share one reciprocal: 1 / a = b / (a b), 1 / b = a / (a b)
        const double rd = frcp(c.dfx0 * c.dfy0);
        w.rx = c.fxm * c.dfy0 * rd;
        w.ry = c.fym * c.dfx0 * rd;
    }
    w.ssum = c.sxm + c.sym;
    w.rsum = w.rx + w.ry;
    w.rsc = 0.;
    double rdf0 = 0., rfm = 0., rsm = 0.;
    if (STRICT) {
        w.hsxn = qdiv(c.sxm, w.ssum) + qdiv(w.rx, w.rsum);
        w.hsyn = qdiv(c.sym, w.ssum) + qdiv(w.ry, w.rsum);
        w.sxn = qdiv(w.sx, w.hsxn);
        w.syn = qdiv(w.sy, w.hsyn);
        w.sc = hypot(w.sxn, w.syn);
        if (w.sc > 0) {
            w.ca = qdiv(w.sxn, w.sc);
            w.sa = qdiv(w.syn, w.sc);
        } else {
            w.ca = kHalfSqrt2;
            w.sa = kHalfSqrt2;
        }
    } else {
        const double rp = frcp(w.ssum * w.rsum);
        w.rssum = w.rsum * rp;
        w.rrsum = w.ssum * rp;
        w.hsxn = c.sxm * w.rssum + w.rx * w.rrsum;
        w.hsyn = c.sym * w.rssum + w.ry * w.rrsum;
        const double rh = frcp(w.hsxn * w.hsyn);
        w.rhsxn = w.hsyn * rh;
        w.rhsyn = w.hsxn * rh;
        w.sxn = w.sx * w.rhsxn;
        w.syn = w.sy * w.rhsyn;
        double rsc;
        fsqrt_rcp(w.sxn * w.sxn + w.syn * w.syn, w.sc, rsc);
        if (w.sc > 0) {
            w.rsc = rsc;
            w.ca = w.sxn * w.rsc;
            w.sa = w.syn * w.rsc;
        } else {
            w.ca = kHalfSqrt2;
            w.sa = kHalfSqrt2;
        }
    }
    w.A = c.dfx0 * w.ca * w.hsxn;
    w.B = c.dfy0 * w.sa * w.hsyn;
    w.C = c.fxm * w.ca;
    w.D = c.fym * w.sa;
    if (STRICT) {
        w.E = qdiv(c.sxm * w.ca, w.hsxn);
        w.F = qdiv(c.sym * w.sa, w.hsyn);
    } else {
        w.E = c.sxm * w.ca * w.rhsxn;
        w.F = c.sym * w.sa * w.rhsyn;
    }
    w.G = c.fxs * w.ca;
    w.H = c.fys * w.sa;
    if (STRICT) {
        w.I = qdiv(c.sxs * w.ca, w.hsxn);
        w.J = qdiv(c.sys * w.sa, w.hsyn);
    } else {
        w.I = c.sxs * w.ca * w.rhsxn;
        w.J = c.sys * w.sa * w.rhsyn;
    }
    if (STRICT) {
        w.df0 = hypot(w.A, w.B);
        w.fm = hypot(w.C, w.D);
        w.sm = hypot(w.E, w.F);
    } else {
        fsqrt_rcp(w.A * w.A + w.B * w.B, w.df0, rdf0);
        fsqrt_rcp(w.C * w.C + w.D * w.D, w.fm, rfm);
        fsqrt_rcp(w.E * w.E + w.F * w.F, w.sm, rsm);
    }
    w.fs = hyp2<STRICT>(w.G, w.H);
    w.ss = hyp2<STRICT>(w.I, w.J);
    w.fos = tmxy_fos<STRICT>(w.sc, w.df0, w.sm, w.fm, w.ss, w.fs, w.rsc, rdf0, rfm, rsm, w.f);
    // -sineStep(vta, 0, 0, 0, 1) * rr * fz * rStat * sgn(omega) with the limited fz; the step is 0
    // for vta <= 0, else 1
    w.ssv = (w.vta <= 0.0) ? 0.0 : 1.0;
    w.My = -w.ssv * t.rr * w.fzc * rStat * w.som;
    w.vtxs = w.rw * w.hsxn + 0.01;
    w.vtys = w.rw * w.hsyn + 0.01;
    w.vy = -w.sy * w.vta;
    w.numx = -w.vtxs * t.cx * xe - w.fos * w.vsx2;
    w.denx = w.vtxs * t.dx + w.fos;
    w.numy = -w.vtys * t.cy * ye - w.fos * w.vy;
    w.deny = w.vtys * t.dy + w.fos;
    if (STRICT) {
        w.xedot = qdiv(w.numx, w.denx);
        w.yedot = qdiv(w.numy, w.deny);
    } else {
        w.xedot = fdiv(w.numx, w.denx);
        w.yedot = fdiv(w.numy, w.deny);
    }
    // the reference forms dx * num / den as (dx * num) / den here, but dy * (num / den) below
    w.fxdyn = (STRICT ? qdiv(t.dx * w.numx, w.denx) : t.dx * w.xedot) + t.cx * xe;
    w.fydyn = t.dy * w.yedot + (t.cy * ye);
    w.fxr = xe * t.cx + w.xedot * t.dx;
    w.fyr = ye * t.cy + w.yedot * t.dy;
    w.fxstr = clampd(w.fxr, -t.fxmP2n, t.fxmP2n);
    w.fystr = clampd(w.fyr, -t.fymP2n, t.fymP2n);
    w.avx = fabs(w.vsx2);
    w.avy = fabs(w.vy);
    w.wxw = sine_step_weight<STRICT>(w.avx);
    w.wyw = sine_step_weight<STRICT>(w.avy);
    w.fx = w.wxw * w.fxstr + (1. - w.wxw) * w.fxdyn;
    w.fy = w.wyw * w.fystr + (1. - w.wyw) * w.fydyn;
}

// TMeasy force core: deflection rates, forces and rolling moment of a tire on the ground.
template <bool STRICT = false>
LFDM_HD void tmeasy_core(const TireP& t, const Derived& u, double qy1, double qy2, double delta, double tan_delta,
                         double fz, double rStat, double vsx, double vsy, double xe, double ye, double omega,
                         double& xedot, double& yedot, double& fx, double& fy, double& My) {
    TireEval w;
    tmeasy_eval<STRICT>(t, u, qy1, qy2, delta, tan_delta, fz, rStat, vsx, vsy, xe, ye, omega, w);
    xedot = w.xedot;
    yedot = w.yedot;
    fx = w.fx;
    fy = w.fy;
    My = w.My;
}

// TMeasyNr force core (dof18.cpp:571-673 with the Coulomb blend of :290-306; the same code is
// dof11.cpp:503-606 and dof24.cpp:1014-1121) for a tire on the ground with static radius rStat.
template <bool STRICT = false>
LFDM_HD void tmeasy_nr_core(const TireP& t, const Derived& u, double qy1, double qy2, double delta,
                                               double tan_delta, double fz, double rStat, double vsx, double vsy,
                                               double omega, double& fx, double& fy, double& My) {
    const double vsx_in = vsx;
    double rdynco;
    if (fz <= t.fzRdynco)
        rdynco = t.rdyncoPn + u.lrd * ((STRICT ? qdiv(fz, t.pn) : fz * u.inv_pn) - 1.);
    else
        rdynco = t.rdyncoCrit;
    const double r_eff = rdynco * t.r0 + (1. - rdynco) * rStat;
    vsx = vsx - (omega * r_eff);
    const double vta = r_eff * fabs(omega) + 0.01;
    double Fx0 = tanh(xdiv<STRICT>(-2.0 * vsx, t.vcoulomb)) * fz * t.mu;
    double Fy0 = tanh(xdiv<STRICT>(-2.0 * vsy, t.vcoulomb)) * fz * t.mu;
    const double F0 = hyp2<STRICT>(Fx0, Fy0);
    if (F0 > fz * t.mu) {
        const double fsc = xdiv<STRICT>(fz * t.mu, F0);
        Fx0 = Fx0 * fsc;
        Fy0 = Fy0 * fsc;
    }
    double sx, sy;
    if (STRICT) {
        sx = qdiv(-vsx, vta);
        sy = -tan(atan2(vsy, vta) - delta);
    } else {
        // as in tmeasy_eval: -(vsy - vta tan d) / (vta + vsy tan d), one reciprocal with sx
        const double den = vta + vsy * tan_delta;
        const double rq = frcp(vta * den);
        sx = -vsx * den * rq;
        sy = -(vsy - vta * tan_delta) * vta * rq;
    }
    if (fz > t.pnmax) fz = t.pnmax;
    Curve c;
    curve_at<STRICT>(t, u, qy1, qy2, fz, c);
    const double frblend = sine_step<STRICT>(fabs(vsx_in), t.frblend_begin, 0., t.frblend_end, 1.);
    double sc, rsc = 0.;
    if (STRICT) sc = hypot(sx, sy);
    else fsqrt_rcp(sx * sx + sy * sy, sc, rsc);
    double ca, sa;
    if (sc > 0) {
        ca = STRICT ? qdiv(sx, sc) : sx * rsc;
        sa = STRICT ? qdiv(sy, sc) : sy * rsc;
    } else {
        ca = kHalfSqrt2;
        sa = kHalfSqrt2;
    }
    double df0, fm, sm, rdf0 = 0., rfm = 0., rsm = 0.;
    if (STRICT) {
        df0 = hypot(c.dfx0 * ca, c.dfy0 * sa);
        fm = t.mu * hypot(c.fxm * ca, c.fym * sa);
        sm = t.mu * hypot(c.sxm * ca, c.sym * sa);
    } else {
        const double ax = c.dfx0 * ca, ay = c.dfy0 * sa, bx = c.fxm * ca, by = c.fym * sa, ex = c.sxm * ca,
                     ey = c.sym * sa;
        double hf, hs;
        fsqrt_rcp(ax * ax + ay * ay, df0, rdf0);
        fsqrt_rcp(bx * bx + by * by, hf, rfm);
        fsqrt_rcp(ex * ex + ey * ey, hs, rsm);
        fm = t.mu * hf;
        sm = t.mu * hs;
        rfm = rfm * u.inv_mu;
        rsm = rsm * u.inv_mu;
    }
    const double fs = t.mu * hyp2<STRICT>(c.fxs * ca, c.fys * sa);
    const double ss = t.mu * hyp2<STRICT>(c.sxs * ca, c.sys * sa);
    double f;
    tmxy_fos<STRICT>(sc, df0, sm, fm, ss, fs, rsc, rdf0, rfm, rsm, f);
    double Fx, Fy;
    if (sc > 0. && STRICT) {
        Fx = qdiv(f * sx, sc);
        Fy = qdiv(f * sy, sc);
    } else if (sc > 0.) {
        Fx = f * sx * rsc;
        Fy = f * sy * rsc;
    } else {
        Fx = 0.;
        Fy = 0.;
    }
    fx = (1.0 - frblend) * Fx0 + frblend * Fx;
    fy = (1.0 - frblend) * Fy0 + frblend * Fy;
    My = -t.rr * fz * rStat * tanh(omega);
}

// differentialSplit (dof18.cpp:46-73) with max_bias = 2.
template <bool STRICT>
LFDM_HD void diff_split(double torque, double sl, double sr, double& tl, double& tr) {
    const double max_bias = 2;
    const double diff = fabs(sl - sr);
    double bias = 1;
    if (diff > 0.5)
        bias = max_bias;
    else if (diff > 0.25)
        bias = 4 * (max_bias - 1) * diff + (2 - max_bias);
    const double alpha = xdiv<STRICT>(bias, 1 + bias);
    const double slow = alpha * torque;
    const double fast = torque - slow;
    const bool left_slow = fabs(sl) < fabs(sr);
    tl = left_slow ? slow : fast;
    tr = left_slow ? fast : slow;
}

// driveTorque (dof18.cpp:15-42, dof11.cpp:15-42, dof24.cpp:92-119) with the per-vehicle torque
// scale tqs on the map ordinates.
template <bool STRICT>
LFDM_HD double drive_torque(const PSet& P, double tqs, double throttle, double w) {
    if (P.throttleMod) {
        const double tq = map_y_scaled<STRICT>(P.torque, w, throttle, tqs, throttle);
        return tq + map_y<STRICT>(P.losses, w);
    }
    const double tq = map_y_scaled<STRICT>(P.torque, w, 1.0, tqs, 1.0);
    return tq * throttle + map_y<STRICT>(P.losses, w);
}

// computePowertrainRHS of the four-wheel models (dof18.cpp:675-845, dof24.cpp:1338-1509): torque
// converter or slaved crank, gear shift for the next step, axle and differential split, wheel
// accelerations.  Updates crank (without a converter) and gear; returns dOm, eng and dcrank.
template <bool STRICT>
LFDM_HD void powertrain4(const PSet& P, const Derived& u, double tqs, double maxBrake,
                                            double throttle, double braking, double vel_u, double& crank, int& gear,
                                            const double* omega, const double* fx, const double* rStat,
                                            const double* My, double* dOm, double* eng, double& dcrank) {
    double torque_t;
    dcrank = 0.0;
    const double osum = omega[0] + omega[1] + omega[2] + omega[3];
    const double gr = pick(P.gears, P.ngears, gear);
    double shift_w;
    // the quotients by the gear ratio stay divisions in the fast form too: the converter's speed ratio
    // omega_out / crank is exactly 1 at a crank speed set from the same quotient, and the model's
    // reverse-flow branch (sr > 1) flips with one ulp there
    if (P.tcbool) {
        const double omega_out = xdiv<STRICT>(0.25 * osum, gr);
        const double omega_in = crank;
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
        if (vel_u < 1e-9 && torque_t < 0) torque_t = 0;
        dcrank = u.inv_ci * (drive_torque<STRICT>(P, tqs, throttle, crank) + torque_in);
        shift_w = omega_out;
    } else {
        crank = xdiv<STRICT>(0.25 * osum, gr);
        torque_t = xdiv<STRICT>(drive_torque<STRICT>(P, tqs, throttle, crank), gr);
        if (vel_u < 1e-9 && torque_t < 0) torque_t = 0;
        shift_w = crank;
    }
    {
        double up = P.shift[0].y, dn = P.shift[0].x;
        for (int g = 1; g < P.ngears; ++g) {
            up = (gear == g) ? P.shift[g].y : up;
            dn = (gear == g) ? P.shift[g].x : dn;
        }
        if (shift_w > up) {
            if (gear < P.ngears - 1) gear++;
        } else if (shift_w < dn) {
            if (gear > 0) gear--;
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
    diff_split<STRICT>(tf, omega[0], omega[1], eng[0], eng[1]);
    diff_split<STRICT>(trr, omega[2], omega[3], eng[2], eng[3]);
    const double bt = maxBrake * braking;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        dOm[i] = u.inv_jw * (eng[i] + My[i] - sgn(omega[i]) * bt - fx[i] * rStat[i]);
}

}  // namespace dev
}  // namespace lfdm
