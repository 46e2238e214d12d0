// Reverse-mode adjoint of the 18-DOF half-implicit step (dof18.cpp / dof18_halfImplicit.cpp as
// restated in d18_device.h), written by hand block by block.
//
// step18_adj(state before step k, controls of step k, adjoint of the state after step k) returns
// the adjoint of the state before step k, the adjoints of the step's controls and accumulates the
// adjoints of the per-vehicle constants.  Each block recomputes its own forward values and then
// runs backwards through them; discrete events (gear, clamps, branch of every piecewise formula,
// off-ground tires, sgn) are frozen at the forward pass's choice, so the result is the derivative
// of the piecewise-smooth map on the piece the trajectory is on (map lookups use the slope of the
// segment the forward pass used).
//
// Supported configuration: TMeasy tires, torque converter on or off, linear steering, no
// throttleMod (the bench and MPC vehicle, HMMWV); lfdm_adjoint_* rejects the others.
#pragma once
#include "d18_device.h"

namespace lfdm {
namespace dev {

// Adjoint of the carried state S18 (the gear is discrete).
struct S18b {
    double xe[4], ye[4], xedot[4], yedot[4], omega[4];
    double x, y, u, v, psi, wz, phi, wx, udot, vdot, crank;
};

LFDM_HD void s18b_zero(S18b& b) {
#pragma unroll
    for (int i = 0; i < 4; ++i) b.xe[i] = b.ye[i] = b.xedot[i] = b.yedot[i] = b.omega[i] = 0.0;
    b.x = b.y = b.u = b.v = b.psi = b.wz = b.phi = b.wx = b.udot = b.vdot = b.crank = 0.0;
}

// Adjoint of the per-vehicle constants the step reads (K18).
struct K18b {
    double Z1f, Z2fc, Z1r, Z2rc, Z4n, neg_mt, E3c, hm, A2, A3, inv_mt, two_hrc_m, denom, A1A2, A1A3, A3sq;
    double qy1, qy2, jz, tqs, maxBrake;
};

LFDM_HD void k18b_zero(K18b& b) {
    b.Z1f = b.Z2fc = b.Z1r = b.Z2rc = b.Z4n = b.neg_mt = b.E3c = b.hm = b.A2 = b.A3 = b.inv_mt = b.two_hrc_m = 0.0;
    b.denom = b.A1A2 = b.A1A3 = b.A3sq = b.qy1 = b.qy2 = b.jz = b.tqs = b.maxBrake = 0.0;
}

// Slope of a clamped piecewise-linear map at x on the segment map_y uses (0 outside the knots).
LFDM_HD double map_slope(const Map& m, double x) {
    const int n = m.n;
    if (x <= m.e[0].x || x >= m.e[n - 1].x) return 0.0;
    int r = 0;
    for (int k = 1; k < n - 1; ++k) r += (m.e[k].x < x) ? 1 : 0;
    r += 1;
    return fdiv(m.e[r].y - m.e[r - 1].y, m.e[r].x - m.e[r - 1].x);
}

// map_y(m, x) and the slope of the segment it uses (0 where it clamps), from one scan of the knots.
LFDM_HD double map_y_slope(const Map& m, double x, double& slope) {
    const int n = m.n;
    slope = 0.0;
    if (x <= m.e[0].x) return m.e[0].y;
    if (x >= m.e[n - 1].x) return m.e[n - 1].y;
    int r = 0;
    for (int k = 1; k < n - 1; ++k) r += (m.e[k].x < x) ? 1 : 0;
    r += 1;
    const MapXY L = m.e[r - 1];
    const MapXY R = m.e[r];
    slope = fdiv(R.y - L.y, R.x - L.x);
    const double mbar = fdiv(x - L.x, R.x - L.x);
    return L.y + mbar * (R.y - L.y);
}

// d sineStep(x, 1, 1, 1.5, 0) / dx; the cosine of the turn fraction v = (x - 1) / dx as the sine a
// quarter turn on (sin_2pi_unit of v + 1/4, whose rounding moves the value by ~1e-16)
LFDM_HD double sine_step_weight_d(double x) {
    if (x <= 1.0 || x >= 1.5) return 0.0;
    const double dx = 0.5, dy = -1.0;
    return dy / dx - (dy / dx) * sin_2pi_unit((x - 1.0) / dx + 0.25);
}

LFDM_HD void hypot_adj(double a, double b, double h, double hb, double& ab, double& bb) {
    if (h > 0.0) {
        const double r = fdiv(hb, h);
        ab += r * a;
        bb += r * b;
    }
}

// The reverse sweeps below multiply by reciprocals of repeated denominators instead of dividing
// each time (a correctly rounded FP64 quotient is 8 instructions, a product 1): the adjoint's last
// bits change, the forward values and branch choices it recomputes do not.

// Adjoint of tmxy_combined (dof18.cpp:247-288): bars of (s, df0, sm, fm, ss, fs) from bars of (f, fos).
LFDM_HD void tmxy_adj(double s, double df0, double sm, double fm, double ss, double fs, double fb, double fosb,
                      double& sb, double& df0b, double& smb, double& fmb, double& ssb, double& fsb) {
    double df0loc = 0.0;
    int which = 0;  // 1: df0loc = df0, 2: df0loc = 2 fm / sm
    if (sm > 0.0) {
        const double c = fdiv(2.0 * fm, sm);
        if (c < df0) {
            df0loc = df0;
            which = 1;
        } else {
            df0loc = c;
            which = 2;
        }
    }
    if (!(s > 0.0 && df0loc > 0.0)) return;
    const double rs = frcp(s), rsm = frcp(sm), rfm = frcp(fm);
    double dlb = 0.0;
    if (s > ss) {
        fsb += fb + fosb * rs;
        sb -= fosb * fs * rs * rs;
    } else if (s < sm) {
        const double p = fdiv(df0loc * sm, fm) - 2.0;
        const double sn = fdiv(s, sm);
        const double dn = 1.0 + (sn + p) * sn;
        const double f = fdiv(df0loc * sm * sn, dn);
        const double rdn = frcp(dn);
        dlb += fosb * rdn;
        double dnb = -fosb * df0loc * rdn * rdn;
        dlb += fb * sm * sn * rdn;
        smb += fb * df0loc * sn * rdn;
        double snb = fb * df0loc * sm * rdn;
        dnb -= fb * f * rdn;
        snb += dnb * (2.0 * sn + p);
        const double pb = dnb * sn;
        sb += snb * rsm;
        smb -= snb * sn * rsm;
        dlb += pb * sm * rfm;
        smb += pb * df0loc * rfm;
        fmb -= pb * df0loc * sm * rfm * rfm;
    } else {
        const double q = fdiv(fm, sm);
        const double a = fdiv(q * q, df0loc * sm);
        const double h = a * (ss - sm);
        const double sstar = sm + fdiv(fm - fs, h);
        double f, ab = 0.0, sstarb = 0.0;
        bool use_sstar = false;
        if (sstar <= ss) {
            use_sstar = true;
            if (s <= sstar) {
                const double d = s - sm;
                f = fm - a * d * d;
            } else {
                const double b = fdiv(a * (sstar - sm), ss - sstar);
                f = fs + b * (ss - s) * (ss - s);
            }
        } else {
            const double sn = fdiv(s - sm, ss - sm);
            f = fm - (fm - fs) * sn * sn * (3.0 - 2.0 * sn);
        }
        const double ftb = fb + fosb * rs;
        sb -= fosb * f * rs * rs;
        if (use_sstar) {
            if (s <= sstar) {
                const double d = s - sm;
                fmb += ftb;
                ab -= ftb * d * d;
                sb -= ftb * 2.0 * a * d;
                smb += ftb * 2.0 * a * d;
            } else {
                const double g = ss - sstar;
                const double rg = frcp(g);
                const double b = a * (sstar - sm) * rg;
                const double e = ss - s;
                fsb += ftb;
                const double bb = ftb * e * e;
                ssb += ftb * 2.0 * b * e;
                sb -= ftb * 2.0 * b * e;
                ab += bb * (sstar - sm) * rg;
                sstarb += bb * (a * rg + a * (sstar - sm) * rg * rg);
                smb -= bb * a * rg;
                ssb -= bb * a * (sstar - sm) * rg * rg;
            }
            // sstar = sm + (fm - fs) / h, h = a (ss - sm)
            const double rh = frcp(h);
            smb += sstarb;
            fmb += sstarb * rh;
            fsb -= sstarb * rh;
            const double hb = -sstarb * (fm - fs) * rh * rh;
            ab += hb * (ss - sm);
            ssb += hb * a;
            smb -= hb * a;
        } else {
            const double w = ss - sm;
            const double rw = frcp(w);
            const double sn = (s - sm) * rw;
            const double g = sn * sn * (3.0 - 2.0 * sn);
            fmb += ftb * (1.0 - g);
            fsb += ftb * g;
            const double snb = -ftb * (fm - fs) * (6.0 * sn - 6.0 * sn * sn);
            sb += snb * rw;
            smb += snb * (sn - 1.0) * rw;
            ssb -= snb * sn * rw;
        }
        // a = fm^2 / (sm^3 df0loc)
        fmb += ab * 2.0 * a * rfm;
        smb -= ab * 3.0 * a * rsm;
        dlb -= fdiv(ab * a, df0loc);
    }
    if (which == 1) {
        df0b += dlb;
    } else if (which == 2) {
        fmb += dlb * 2.0 * rsm;
        smb -= dlb * 2.0 * fm * rsm * rsm;
    }
}

// Output bars of one TMeasy tire and the input bars they produce.
struct TireBars {
    double xedot, yedot, fx, fy, My, rStat;                  // in
    double fz, vsx, vsy, delta, xe, ye, omega, qy1, qy2;      // out (accumulated)
};

// Adjoint of tire_tmeasy (dof18.cpp:380-539) for a tire on the ground (fz > 0).
LFDM_HD void tire_tmeasy_adj(const TireP& t, const Derived& u, double qy1, double qy2, double tan_delta, double fz,
                             double vsx, double vsy, double xe, double ye, double omega, TireBars& B) {
    // ---- forward values (the step's own evaluation) ----
    const double rStat = t.r0 - fz * u.inv_kt;
    TireEval w;
    tmeasy_eval(t, u, qy1, qy2, 0.0, tan_delta, fz, rStat, vsx, vsy, xe, ye, omega, w);
    const double rdynco = w.rdynco, drd = w.drd, r_eff = w.r_eff, vsx2 = w.vsx2, aom = w.aom, som = w.som;
    const double rw = w.rw, vta = w.vta, sx = w.sx, sy = w.sy, ta = -sy;  // ta: tan of the slip angle
    const bool clip = w.clip;
    const double fzc = w.fzc;
    const Curve& c = w.c;
    const double fzn = w.fzn, rx = w.rx, ry = w.ry, hsxn = w.hsxn, hsyn = w.hsyn;
    const double sxn = w.sxn, syn = w.syn, sc = w.sc, ca = w.ca, sa = w.sa;
    const double A = w.A, Bq = w.B, C = w.C, D = w.D, E = w.E, F = w.F, G = w.G, H = w.H, I = w.I, J = w.J;
    const double df0 = w.df0, fm = w.fm, sm = w.sm, fs = w.fs, ss = w.ss, fos = w.fos, ssv = w.ssv;
    const double vtxs = w.vtxs, vtys = w.vtys, vy = w.vy, numx = w.numx, denx = w.denx, numy = w.numy;
    const double deny = w.deny, xedot = w.xedot, yedot = w.yedot, fxdyn = w.fxdyn, fydyn = w.fydyn;
    const double fxr = w.fxr, fyr = w.fyr, fxstr = w.fxstr, fystr = w.fystr, avx = w.avx, avy = w.avy;
    const double wxw = w.wxw, wyw = w.wyw;

    // ---- reverse ----
    const double rhsxn = w.rhsxn, rhsyn = w.rhsyn, rvta = frcp(vta);
    double vsx2b = 0, vyb = 0, xedotb = B.xedot, yedotb = B.yedot, xeb = 0, yeb = 0;
    // fx = wxw fxstr + (1 - wxw) fxdyn
    const double wxwb = B.fx * (fxstr - fxdyn), fxstrb = B.fx * wxw, fxdynb = B.fx * (1. - wxw);
    const double wywb = B.fy * (fystr - fydyn), fystrb = B.fy * wyw, fydynb = B.fy * (1. - wyw);
    vsx2b += wxwb * sine_step_weight_d(avx) * sgn(vsx2);
    vyb += wywb * sine_step_weight_d(avy) * sgn(vy);
    const double fxrb = (fxr > -t.fxmP2n && fxr < t.fxmP2n) ? fxstrb : 0.0;
    const double fyrb = (fyr > -t.fymP2n && fyr < t.fymP2n) ? fystrb : 0.0;
    xeb += fxrb * t.cx + fxdynb * t.cx;
    xedotb += fxrb * t.dx + fxdynb * t.dx;
    yeb += fyrb * t.cy + fydynb * t.cy;
    yedotb += fyrb * t.dy + fydynb * t.dy;
    // xedot = numx / denx, yedot = numy / deny
    const double numxb = fdiv(xedotb, denx), denxb = -numxb * xedot;
    const double numyb = fdiv(yedotb, deny), denyb = -numyb * yedot;
    double vtxsb = -numxb * t.cx * xe + denxb * t.dx;
    xeb += -numxb * vtxs * t.cx;
    double fosb = -numxb * vsx2 + denxb;
    vsx2b += -numxb * fos;
    double vtysb = -numyb * t.cy * ye + denyb * t.dy;
    yeb += -numyb * vtys * t.cy;
    fosb += -numyb * vy + denyb;
    vyb += -numyb * fos;
    // vy = -sy vta
    double syb = -vyb * vta;
    double vtab = -vyb * sy;
    // vtxs = rw hsxn + 0.01, vtys = rw hsyn + 0.01
    double rwb = vtxsb * hsxn + vtysb * hsyn;
    double hsxnb = vtxsb * rw, hsynb = vtysb * rw;
    // My = -ssv rr fzc rStat sgn(omega), fzc = min(fz, pnmax)
    double fzb = clip ? 0.0 : B.My * (-ssv * t.rr * rStat * som);
    double rStatb = B.rStat + B.My * (-ssv * t.rr * fzc * som);
    // tmxy
    double scb = 0, df0b = 0, smb = 0, fmb = 0, ssb = 0, fsb = 0;
    tmxy_adj(sc, df0, sm, fm, ss, fs, 0.0, fosb, scb, df0b, smb, fmb, ssb, fsb);
    double Ab = 0, Bb = 0, Cb = 0, Db = 0, Eb = 0, Fb = 0, Gb = 0, Hb = 0, Ib = 0, Jb = 0;
    hypot_adj(A, Bq, df0, df0b, Ab, Bb);
    hypot_adj(C, D, fm, fmb, Cb, Db);
    hypot_adj(E, F, sm, smb, Eb, Fb);
    hypot_adj(G, H, fs, fsb, Gb, Hb);
    hypot_adj(I, J, ss, ssb, Ib, Jb);
    double cab = 0, sab = 0;
    double dfx0b = 0, dfy0b = 0, fxmb = 0, fymb = 0, fxsb = 0, fysb = 0, sxmb = 0, symb = 0, sxsb = 0, sysb = 0;
    // I = sxs ca / hsxn, J = sys sa / hsyn
    const double Ibr = Ib * rhsxn, Jbr = Jb * rhsyn;
    sxsb += Ibr * ca; cab += Ibr * c.sxs; hsxnb -= Ibr * I;
    sysb += Jbr * sa; sab += Jbr * c.sys; hsynb -= Jbr * J;
    // G = fxs ca, H = fys sa
    fxsb += Gb * ca; cab += Gb * c.fxs;
    fysb += Hb * sa; sab += Hb * c.fys;
    // E = sxm ca / hsxn, F = sym sa / hsyn
    const double Ebr = Eb * rhsxn, Fbr = Fb * rhsyn;
    sxmb += Ebr * ca; cab += Ebr * c.sxm; hsxnb -= Ebr * E;
    symb += Fbr * sa; sab += Fbr * c.sym; hsynb -= Fbr * F;
    // C = fxm ca, D = fym sa
    fxmb += Cb * ca; cab += Cb * c.fxm;
    fymb += Db * sa; sab += Db * c.fym;
    // A = dfx0 ca hsxn, B = dfy0 sa hsyn
    dfx0b += Ab * ca * hsxn; cab += Ab * c.dfx0 * hsxn; hsxnb += Ab * c.dfx0 * ca;
    dfy0b += Bb * sa * hsyn; sab += Bb * c.dfy0 * hsyn; hsynb += Bb * c.dfy0 * sa;
    double sxnb = 0, synb = 0;
    if (sc > 0) {
        // ca = sxn / sc, sa = syn / sc, sc = hypot(sxn, syn)
        const double rsc = frcp(sc);
        scb -= (cab * ca + sab * sa) * rsc;
        sxnb += (cab + scb * sxn) * rsc;
        synb += (sab + scb * syn) * rsc;
    }
    // sxn = sx / hsxn, syn = sy / hsyn
    const double sxnbr = sxnb * rhsxn, synbr = synb * rhsyn;
    const double sxb = sxnbr;
    hsxnb -= sxnbr * sxn;
    syb += synbr;
    hsynb -= synbr * syn;
    // hsxn = sxm/ssum + rx/rsum, hsyn = sym/ssum + ry/rsum
    const double rssum = w.rssum, rrsum = w.rrsum;
    sxmb += hsxnb * rssum;
    symb += hsynb * rssum;
    const double ssumb = -(hsxnb * c.sxm + hsynb * c.sym) * rssum * rssum;
    double rxb = hsxnb * rrsum, ryb = hsynb * rrsum;
    const double rsumb = -(hsxnb * rx + hsynb * ry) * rrsum * rrsum;
    rxb += rsumb;
    ryb += rsumb;
    sxmb += ssumb;
    symb += ssumb;
    // rx = fxm / dfx0, ry = fym / dfy0
    const double rxbr = fdiv(rxb, c.dfx0), rybr = fdiv(ryb, c.dfy0);
    fxmb += rxbr;
    dfx0b -= rxbr * rx;
    fymb += rybr;
    dfy0b -= rybr * ry;
    // curve_at(fzc): InterpQ v = fzn (q1 - q2 fzn), InterpL v = w1 + l (fzn - 1)
    double fznb = dfx0b * (u.qx1 - 2.0 * u.qx2 * fzn) + dfy0b * (qy1 - 2.0 * qy2 * fzn) +
                  fxmb * (u.qfxm1 - 2.0 * u.qfxm2 * fzn) + fymb * (u.qfym1 - 2.0 * u.qfym2 * fzn) +
                  fxsb * (u.qfxs1 - 2.0 * u.qfxs2 * fzn) + fysb * (u.qfys1 - 2.0 * u.qfys2 * fzn) + sxmb * u.lsxm +
                  symb * u.lsym + sxsb * u.lsxs + sysb * u.lsys;
    B.qy1 += dfy0b * fzn;
    B.qy2 -= dfy0b * fzn * fzn;
    if (!clip) fzb += fznb * u.inv_pn;
    // sy = -tan(alpha), alpha = atan2(vsy, vta) - delta (evaluated as a tangent difference)
    const double alphab = -syb * (1.0 + ta * ta);
    B.delta -= alphab;
    const double ar2 = fdiv(alphab, vsy * vsy + vta * vta);
    B.vsy += ar2 * vta;
    vtab -= ar2 * vsy;
    // sx = -vsx2 / vta
    vsx2b -= sxb * rvta;
    vtab += sxb * vsx2 * rvta * rvta;
    // vta = rw + 0.01, rw = r_eff |omega|
    rwb += vtab;
    double r_effb = rwb * aom;
    double omegab = rwb * r_eff * som;
    // vsx2 = vsx - omega r_eff
    B.vsx += vsx2b;
    omegab -= vsx2b * r_eff;
    r_effb -= vsx2b * omega;
    // r_eff = rdynco r0 + (1 - rdynco) rStat
    const double rdyncob = r_effb * (t.r0 - rStat);
    rStatb += r_effb * (1. - rdynco);
    fzb += rdyncob * drd;
    // rStat = r0 - fz / kt
    fzb -= rStatb * u.inv_kt;
    B.fz += fzb;
    B.xe += xeb;
    B.ye += yeb;
    B.omega += omegab;
}

// Adjoint of differentialSplit (dof18.cpp:46-73): bars of (torque, sl, sr) from bars of (tl, tr).
LFDM_HD void diff_split_adj(double torque, double sl, double sr, double tlb, double trb, double& torqueb, double& slb,
                            double& srb) {
    const double max_bias = 2;
    const double diff = fabs(sl - sr);
    double bias = 1, dbias = 0;
    if (diff > 0.5) {
        bias = max_bias;
    } else if (diff > 0.25) {
        bias = 4 * (max_bias - 1) * diff + (2 - max_bias);
        dbias = 4 * (max_bias - 1);
    }
    const double alpha = fdiv(bias, 1 + bias);
    const bool left_slow = fabs(sl) < fabs(sr);
    const double slowb = left_slow ? tlb : trb;
    const double fastb = left_slow ? trb : tlb;
    // fast = torque - slow, slow = alpha torque
    torqueb += fastb;
    const double sb = slowb - fastb;
    torqueb += sb * alpha;
    const double alphab = sb * torque;
    const double biasb = fdiv(alphab, (1 + bias) * (1 + bias));
    const double diffb = biasb * dbias;
    const double sd = sgn(sl - sr);
    slb += diffb * sd;
    srb -= diffb * sd;
}

// Adjoint of powertrain4 with throttleMod off.  Inputs as powertrain4 (omega after the tires'
// off-ground resets, crank and gear before the step); bars in: dOm, eng (recorded engine torque is
// not a state, pass 0), dcrank, crank_out (the slaved crank without a converter); bars out:
// omega, crank_in, fx, rStat, My, throttle, braking, tqs, maxBrake.
struct PtBars {
    double dOm[4], dcrank, crank_out;                                                 // in
    double omega[4], crank, fx[4], rStat[4], My[4], throttle, braking, tqs, maxBrake;  // out (accumulated)
};

LFDM_HD void powertrain4_adj(const PSet& P, const Derived& u, double tqs, double maxBrake, double throttle,
                             double braking, double vel_u, double crank, int gear, const double* omega,
                             const double* fx, const double* rStat, const double* My, PtBars& B) {
    const double osum = omega[0] + omega[1] + omega[2] + omega[3];
    const double gr = pick(P.gears, P.ngears, gear);
    // ---- forward recompute ----
    double torque_t, torque_in = 0, torque_out = 0, sr0 = 0, q = 0, cf = 1, tr = 0, cfs = 0, trs = 0;
    bool active = false, rev = false, clip = false;
    const double omega_out = fdiv(0.25 * osum, gr);
    double w_dt;  // crank speed the drive torque is evaluated at
    if (P.tcbool) {
        const double omega_in = crank;
        double sr = 0;
        if (!(omega_out < 1e-9 || omega_in < 1e-9)) {
            sr0 = fdiv(omega_out, omega_in);
            sr = sr0;
            if (sr > 1.) {
                sr = 1. - (sr - 1.);
                rev = true;
            }
            active = true;
            if (sr < 0) {
                sr = 0;
                active = false;
            }
        }
        cf = map_y_slope(P.cfmap, sr, cfs);
        tr = map_y_slope(P.trmap, sr, trs);
        q = fdiv(omega_in, cf);
        torque_in = -(q * q);
        if (rev) torque_in = -torque_in;
        torque_out = rev ? -torque_in : -tr * torque_in;
        torque_t = fdiv(torque_out, gr);
        w_dt = crank;
    } else {
        w_dt = fdiv(0.25 * osum, gr);
        torque_t = fdiv(drive_torque<false>(P, tqs, throttle, w_dt), gr);
    }
    if (vel_u < 1e-9 && torque_t < 0) {
        torque_t = 0;
        clip = true;
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
    // ---- reverse ----
    double engb[4];
    double btb = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double cm = B.dOm[i] * u.inv_jw;
        engb[i] = cm;
        B.My[i] += cm;
        btb -= cm * sgn(omega[i]);
        B.fx[i] -= cm * rStat[i];
        B.rStat[i] -= cm * fx[i];
    }
    B.maxBrake += btb * braking;
    B.braking += btb * maxBrake;
    double tfb = 0, trrb = 0;
    diff_split_adj(tf, omega[0], omega[1], engb[0], engb[1], tfb, B.omega[0], B.omega[1]);
    diff_split_adj(trr, omega[2], omega[3], engb[2], engb[3], trrb, B.omega[2], B.omega[3]);
    double ttb;
    if (P.driveType)
        ttb = 0.5 * (tfb + trrb);
    else if (P.whichWheels)
        ttb = trrb;
    else
        ttb = tfb;
    if (clip) ttb = 0;
    // drive torque at w_dt: tqs * T(w) * throttle + L(w)
    double Ts;
    const double Tw = map_y_slope(P.torque, w_dt, Ts), Ls = map_slope(P.losses, w_dt);
    const double rgr = frcp(gr);
    if (P.tcbool) {
        const double omega_in = crank;
        double toutb = ttb * rgr;
        double tinb = rev ? -toutb : -tr * toutb;
        const double trb = rev ? 0.0 : -toutb * torque_in;
        // dcrank = inv_ci (DT + torque_in)
        const double DTb = B.dcrank * u.inv_ci;
        tinb += B.dcrank * u.inv_ci;
        B.tqs += DTb * Tw * throttle;
        B.throttle += DTb * tqs * Tw;
        B.crank += DTb * (tqs * throttle * Ts + Ls);
        // torque_in = -q^2 (negated when reversed)
        const double qb = tinb * (rev ? 2.0 * q : -2.0 * q);
        const double qbr = fdiv(qb, cf);
        double omega_inb = qbr;
        const double cfb = -qbr * q;
        const double srb = cfb * cfs + trb * trs;
        if (active) {
            const double sr0b = rev ? -srb : srb;
            const double omega_outb = fdiv(sr0b, omega_in);
            omega_inb -= omega_outb * sr0;
            const double ob = omega_outb * 0.25 * rgr;
#pragma unroll
            for (int i = 0; i < 4; ++i) B.omega[i] += ob;
        }
        B.crank += omega_inb;
    } else {
        // crank_out = w_dt = 0.25 osum / gr; torque_t = DT(w_dt) / gr
        const double DTb = ttb * rgr;
        B.tqs += DTb * Tw * throttle;
        B.throttle += DTb * tqs * Tw;
        const double wb = (B.crank_out + DTb * (tqs * throttle * Ts + Ls)) * 0.25 * rgr;
#pragma unroll
        for (int i = 0; i < 4; ++i) B.omega[i] += wb;
    }
}

// Adjoint of one half-implicit step.  s0: state before the step (gear included); sb: in = adjoint
// of the state after the step, out = adjoint of the state before it; controls' bars are returned,
// constants' bars accumulated into kb.
// `to` (optional) holds the step's tire outputs from a forward pass (step18<..., CAPTURE>); without
// it the tires' forward is recomputed here.
LFDM_HD void step18_adj(const PSet& P, const Derived& u, const K18& k, const S18& s0, double steering,
                        double throttle, double braking, double dt, S18b& sb, double& steerb, double& thrb,
                        double& brkb, K18b& kb, const TireOut18* to = nullptr) {
    const TireP& t = P.tire;
    // ================= forward recompute =================
    const double lat = s0.vdot + s0.wz * s0.u;
    const double Z2f = k.Z2fc * lat;
    const double Z3f = (P.krof * s0.phi + P.brof * s0.wx) * u.inv_cf;
    const double Z4 = (k.Z4n * (s0.udot - s0.wz * s0.v)) * u.inv_two_ab;
    const double Z2r = k.Z2rc * lat;
    const double Z3r = (P.kror * s0.phi + P.bror * s0.wx) * u.inv_cr;
    double fzr[4], fz[4];
    fzr[0] = k.Z1f - Z2f - Z3f - Z4;
    fzr[1] = k.Z1f + Z2f + Z3f - Z4;
    fzr[2] = k.Z1r - Z2r - Z3r + Z4;
    fzr[3] = k.Z1r + Z2r + Z3r + Z4;
#pragma unroll
    for (int i = 0; i < 4; ++i) fz[i] = fzr[i] > 0. ? fzr[i] : 0.;
    const double delta = steering * P.maxSteer;
    double sd, cd;
    sincos_fast(delta, sd, cd);
    double vsx[4], vsy[4];
    vsy[0] = s0.v + s0.wz * P.a;
    vsx[0] = (s0.u - (s0.wz * P.cf) / 2.) * cd + vsy[0] * sd;
    vsy[1] = vsy[0];
    vsx[1] = (s0.u + (s0.wz * P.cf) / 2.) * cd + vsy[1] * sd;
    vsy[2] = s0.v - s0.wz * P.b;
    vsx[2] = s0.u - (s0.wz * P.cr) / 2.;
    vsy[3] = vsy[2];
    vsx[3] = s0.u + (s0.wz * P.cr) / 2.;
    // tires: on-ground outputs; off-ground tires reset xe, ye, omega and keep xedot, yedot
    double fx[4], fy[4], My[4], rStat[4], xedot[4], yedot[4], xe[4], ye[4], om[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        xe[i] = s0.xe[i];
        ye[i] = s0.ye[i];
        om[i] = s0.omega[i];
        xedot[i] = s0.xedot[i];
        yedot[i] = s0.yedot[i];
        if (to) {
            if (fz[i] <= 0) {
                xe[i] = 0;
                ye[i] = 0;
                om[i] = 0;
            }
            fx[i] = to->fx[i];
            fy[i] = to->fy[i];
            My[i] = to->My[i];
            rStat[i] = to->rStat[i];
            xedot[i] = to->xedot[i];
            yedot[i] = to->yedot[i];
        } else {
            double xt;
            tire_tmeasy(t, u, k, 0.0, i < 2 ? fdiv(sd, cd) : u.tan_delta_rear, fz[i], vsx[i], vsy[i], xe[i], ye[i],
                        xedot[i], yedot[i], om[i], fx[i], fy[i], My[i], xt, rStat[i]);
        }
    }
    double crank = s0.crank;
    int gear = s0.gear;
    double dOm[4], eng[4], dcrank;
    powertrain4<false>(P, u, k.tqs, k.maxBrake, throttle, braking, s0.u, crank, gear, om, fx, rStat, My, dOm, eng, dcrank);
    double gfx[4], gfy[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        gfx[i] = fx[i];
        gfy[i] = fy[i];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        gfx[i] = fx[i] * cd - fy[i] * sd;
        gfy[i] = fx[i] * sd + fy[i] * cd;
    }
    const double jxz = P.jxz, jz = k.jz, A1 = u.A1, A2 = k.A2, A3 = k.A3, mt = -k.neg_mt;
    const double E1 = k.neg_mt * s0.wz * s0.u + (gfy[0] + gfy[1] + gfy[2] + gfy[3]);
    const double E2 = (gfy[0] + gfy[1]) * P.a - (gfy[2] + gfy[3]) * P.b + (gfx[1] - gfx[0]) * P.cf / 2 +
                      (gfx[3] - gfx[2]) * P.cr / 2 + u.E2c * s0.wz * s0.u;
    const double E3 = k.E3c * s0.phi - u.kro * s0.phi - u.bro * s0.wx + k.hm * s0.wz * s0.u;
    const double inner = (gfx[0] + gfx[1] + gfx[2] + gfx[3]) + u.Uc * (s0.wz * s0.wz) - k.two_hrc_m * s0.wz * s0.wx;
    const double vdot = (E1 * u.jxz2 - k.A1A2 * E2 + A1 * E3 * jxz + A3 * E2 * jxz - A2 * E1 * jz - A3 * E3 * jz) *
                        k.rdenom;
    const double wxdot =
        (u.A1sq * E3 - k.A1A3 * E2 + A1 * E1 * jxz - A3 * E1 * jz + E2 * jxz * mt - E3 * jz * mt) * k.rdenom;
    const double wzdot =
        (k.A3sq * E2 - k.A1A2 * E1 - k.A1A3 * E3 + A3 * E1 * jxz - A2 * E2 * mt + E3 * jxz * mt) * k.rdenom;
    const double udot = s0.wz * s0.v + k.inv_mt * inner;
    const double un = s0.u + udot * dt, vn = s0.v + vdot * dt, wzn = s0.wz + wzdot * dt;
    double sp, cp;
    sincos_fast(s0.psi, sp, cp);

    // ================= reverse =================
    S18b o = sb;  // bars of the state after the step
    s18b_zero(sb);
    // x' = x + (u' cp - v' sp) dt, y' = y + (u' sp + v' cp) dt
    double unb = o.u, vnb = o.v, wznb = o.wz, wxnb = o.wx;
    sb.x += o.x;
    sb.y += o.y;
    unb += o.x * cp * dt + o.y * sp * dt;
    vnb += -o.x * sp * dt + o.y * cp * dt;
    sb.psi += o.x * (-un * sp - vn * cp) * dt + o.y * (un * cp - vn * sp) * dt;
    // psi' = psi + wz' dt, phi' = phi + wx' dt
    sb.psi += o.psi;
    wznb += o.psi * dt;
    sb.phi += o.phi;
    wxnb += o.phi * dt;
    // u' = u + udot dt (and udot' = udot), likewise v, wx, wz
    sb.u += unb;
    double udotb = unb * dt + o.udot;
    sb.v += vnb;
    double vdotb = vnb * dt + o.vdot;
    sb.wx += wxnb;
    const double wxdotb = wxnb * dt;
    sb.wz += wznb;
    const double wzdotb = wznb * dt;
    // crank' = crank + dcrank dt (converter) or the slaved crank
    PtBars pb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pb.dOm[i] = o.omega[i] * dt;
        pb.omega[i] = 0;
        pb.fx[i] = 0;
        pb.rStat[i] = 0;
        pb.My[i] = 0;
    }
    pb.crank = 0;
    pb.throttle = pb.braking = pb.tqs = pb.maxBrake = 0;
    if (P.tcbool) {
        sb.crank += o.crank;
        pb.dcrank = o.crank * dt;
        pb.crank_out = 0;
    } else {
        pb.dcrank = 0;
        pb.crank_out = o.crank;
    }
    // tire deflections: xe' = xe(reset) + xedot dt, xedot' = xedot
    double xedotb[4], yedotb[4], omb[4], xeb[4], yeb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        xeb[i] = o.xe[i];
        yeb[i] = o.ye[i];
        xedotb[i] = o.xe[i] * dt + o.xedot[i];
        yedotb[i] = o.ye[i] * dt + o.yedot[i];
        omb[i] = o.omega[i];
    }
    // ---- chassis (computeVehRHS) ----
    const double rden = k.rdenom;
    const double Nvb = vdotb * rden, Nxb = wxdotb * rden, Nzb = wzdotb * rden;
    kb.denom -= (Nvb * vdot + Nxb * wxdot + Nzb * wzdot);
    double E1b = Nvb * (u.jxz2 - A2 * jz) + Nxb * (A1 * jxz - A3 * jz) + Nzb * (-k.A1A2 + A3 * jxz);
    double E2b = Nvb * (-k.A1A2 + A3 * jxz) + Nxb * (-k.A1A3 + jxz * mt) + Nzb * (k.A3sq - A2 * mt);
    double E3b = Nvb * (A1 * jxz - A3 * jz) + Nxb * (u.A1sq - jz * mt) + Nzb * (-k.A1A3 + jxz * mt);
    kb.A1A2 -= Nvb * E2 + Nzb * E1;
    kb.A1A3 -= Nxb * E2 + Nzb * E3;
    kb.A3sq += Nzb * E2;
    kb.A3 += Nvb * (E2 * jxz - E3 * jz) - Nxb * E1 * jz + Nzb * jxz * E1;
    kb.A2 -= Nvb * E1 * jz + Nzb * mt * E2;
    kb.jz -= Nvb * (A2 * E1 + A3 * E3) + Nxb * (A3 * E1 + mt * E3);
    const double mtb = Nxb * (jxz * E2 - jz * E3) + Nzb * (-A2 * E2 + jxz * E3);
    kb.neg_mt -= mtb;
    // udot = wz v + inv_mt inner
    sb.v += udotb * s0.wz;
    sb.wz += udotb * s0.v;
    kb.inv_mt += udotb * inner;
    const double innerb = udotb * k.inv_mt;
    double gfxb[4], gfyb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        gfxb[i] = innerb;
        gfyb[i] = E1b;
    }
    sb.wz += innerb * (2.0 * u.Uc * s0.wz - k.two_hrc_m * s0.wx);
    sb.wx -= innerb * k.two_hrc_m * s0.wz;
    kb.two_hrc_m -= innerb * s0.wz * s0.wx;
    // E1 = neg_mt wz u + sum fy
    kb.neg_mt += E1b * s0.wz * s0.u;
    sb.wz += E1b * k.neg_mt * s0.u;
    sb.u += E1b * k.neg_mt * s0.wz;
    // E2
    gfyb[0] += E2b * P.a;
    gfyb[1] += E2b * P.a;
    gfyb[2] -= E2b * P.b;
    gfyb[3] -= E2b * P.b;
    gfxb[1] += E2b * P.cf / 2;
    gfxb[0] -= E2b * P.cf / 2;
    gfxb[3] += E2b * P.cr / 2;
    gfxb[2] -= E2b * P.cr / 2;
    sb.wz += E2b * u.E2c * s0.u;
    sb.u += E2b * u.E2c * s0.wz;
    // E3
    kb.E3c += E3b * s0.phi;
    sb.phi += E3b * (k.E3c - u.kro);
    sb.wx -= E3b * u.bro;
    kb.hm += E3b * s0.wz * s0.u;
    sb.wz += E3b * k.hm * s0.u;
    sb.u += E3b * k.hm * s0.wz;
    // ---- tireToVehTransform (front tires rotated by delta) ----
    double fxb[4], fyb[4];
    double sdb = 0, cdb = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        fxb[i] = gfxb[i];
        fyb[i] = gfyb[i];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        fxb[i] = gfxb[i] * cd + gfyb[i] * sd;
        fyb[i] = -gfxb[i] * sd + gfyb[i] * cd;
        sdb += -gfxb[i] * fy[i] + gfyb[i] * fx[i];
        cdb += gfxb[i] * fx[i] + gfyb[i] * fy[i];
    }
    // ---- powertrain ----
    powertrain4_adj(P, u, k.tqs, k.maxBrake, throttle, braking, s0.u, s0.crank, s0.gear, om, fx, rStat, My, pb);
    thrb += pb.throttle;
    brkb += pb.braking;
    kb.tqs += pb.tqs;
    kb.maxBrake += pb.maxBrake;
    sb.crank += pb.crank;
    // ---- tires ----
    double fzb[4], vsxb[4], vsyb[4];
    double deltab = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        fzb[i] = 0;
        vsxb[i] = 0;
        vsyb[i] = 0;
        omb[i] += pb.omega[i];
        if (fz[i] <= 0) {
            // reset: xe, ye, omega become 0 (their incoming bars stop here); xedot, yedot pass through
            sb.xedot[i] += xedotb[i];
            sb.yedot[i] += yedotb[i];
            continue;
        }
        TireBars tb;
        tb.xedot = xedotb[i];
        tb.yedot = yedotb[i];
        tb.fx = fxb[i] + pb.fx[i];
        tb.fy = fyb[i];
        tb.My = pb.My[i];
        tb.rStat = pb.rStat[i];
        tb.fz = tb.vsx = tb.vsy = tb.delta = tb.xe = tb.ye = tb.omega = tb.qy1 = tb.qy2 = 0.0;
        tire_tmeasy_adj(t, u, k.qy1, k.qy2, i < 2 ? fdiv(sd, cd) : u.tan_delta_rear, fz[i], vsx[i], vsy[i], s0.xe[i],
                        s0.ye[i], s0.omega[i], tb);
        sb.xe[i] += xeb[i] + tb.xe;
        sb.ye[i] += yeb[i] + tb.ye;
        sb.omega[i] += omb[i] + tb.omega;
        fzb[i] = tb.fz;
        vsxb[i] = tb.vsx;
        vsyb[i] = tb.vsy;
        if (i < 2) deltab += tb.delta;
        kb.qy1 += tb.qy1;
        kb.qy2 += tb.qy2;
    }
    // ---- vehToTireTransform ----
    const double m0b = vsxb[0] * cd, m1b = vsxb[1] * cd;
    cdb += vsxb[0] * (s0.u - (s0.wz * P.cf) / 2.) + vsxb[1] * (s0.u + (s0.wz * P.cf) / 2.);
    sdb += (vsxb[0] + vsxb[1]) * vsy[0];
    const double vsy0b = vsyb[0] + vsyb[1] + (vsxb[0] + vsxb[1]) * sd;
    sb.u += m0b + m1b + vsxb[2] + vsxb[3];
    sb.wz += (m1b - m0b) * P.cf / 2. + (vsxb[3] - vsxb[2]) * P.cr / 2.;
    sb.v += vsy0b + vsyb[2] + vsyb[3];
    sb.wz += vsy0b * P.a - (vsyb[2] + vsyb[3]) * P.b;
    deltab += sdb * cd - cdb * sd;
    steerb += deltab * P.maxSteer;
    // ---- computeTireLoads ----
    double rb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) rb[i] = fzr[i] > 0. ? fzb[i] : 0.;
    kb.Z1f += rb[0] + rb[1];
    kb.Z1r += rb[2] + rb[3];
    const double Z2fb = -rb[0] + rb[1], Z3fb = -rb[0] + rb[1];
    const double Z2rb = -rb[2] + rb[3], Z3rb = -rb[2] + rb[3];
    const double Z4b = -rb[0] - rb[1] + rb[2] + rb[3];
    kb.Z2fc += Z2fb * lat;
    kb.Z2rc += Z2rb * lat;
    const double latb = Z2fb * k.Z2fc + Z2rb * k.Z2rc;
    sb.phi += Z3fb * u.krof_cf + Z3rb * u.kror_cr;
    sb.wx += Z3fb * u.brof_cf + Z3rb * u.bror_cr;
    kb.Z4n += Z4b * (s0.udot - s0.wz * s0.v) * u.inv_two_ab;
    const double ddb = Z4b * k.Z4n * u.inv_two_ab;
    sb.udot += ddb;
    sb.wz -= ddb * s0.v;
    sb.v -= ddb * s0.wz;
    sb.vdot += latb;
    sb.wz += latb * s0.u;
    sb.u += latb * s0.wz;
}

// Adjoint of k18_init: bars of the per-vehicle parameters (m, jz, jx, dfy0Pn, dfy0P2n, tqs, maxBrake)
// from the accumulated bars of the constants.
LFDM_HD void k18_init_adj(const PSet& P, const K18& k, const K18b& kb, double* pb) {
    const double a = P.a, b = P.b, h = P.h, cf = P.cf, cr = P.cr;
    const double hrc = (P.hrcf * b + P.hrcr * a) / (a + b);
    const double A1 = P.d.A1, jxz = P.jxz, mt = -k.neg_mt, jz = k.jz, A2 = k.A2, A3 = k.A3;
    const double A2b = kb.A2 + kb.A1A2 * A1 + kb.denom * (A1 * A1 - jz * mt);
    const double A3b = kb.A3 + kb.A1A3 * A1 + kb.A3sq * 2.0 * A3 + kb.denom * (-2.0 * A1 * jxz + 2.0 * jz * A3);
    const double mtb = -kb.neg_mt - kb.inv_mt / (mt * mt) + kb.denom * (jxz * jxz - A2 * jz);
    const double jzb = kb.jz + kb.denom * (A3 * A3 - A2 * mt);
    const double mb = kb.Z1f * kG * b / (2. * (a + b)) + kb.Z2fc * b * (h - P.hrcf) / (cf * (a + b)) +
                      kb.Z1r * kG * a / (2. * (a + b)) + kb.Z2rc * a * (h - P.hrcr) / (cr * (a + b)) + kb.Z4n * h +
                      mtb + kb.E3c * kG * hrc + kb.hm * hrc + A2b * hrc * hrc + A3b * hrc + kb.two_hrc_m * 2. * hrc;
    pb[0] += mb;
    pb[1] += jzb;
    pb[2] += A2b;                               // jx
    pb[3] += 2. * kb.qy1 + kb.qy2;              // dfy0Pn
    pb[4] += -0.5 * kb.qy1 - 0.5 * kb.qy2;      // dfy0P2n
    pb[5] += kb.tqs;
    pb[6] += kb.maxBrake;
}

}  // namespace dev
}  // namespace lfdm
