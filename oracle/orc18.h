// ORACLE — test infrastructure only (see orc_common.h).
//
// CPU restatement of the reference 18-DOF model and its half-implicit step, following the
// operation order of wheeled_vehicle_models/18dof/dof18.cpp and dof18_halfImplicit.cpp so that
// results are bit-identical to the reference solver compiled with the same flags.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "orc_common.h"

namespace orc {
namespace d18 {

// dof18.h:489-799 (VehicleParam), 23-152 (TMeasyParam), 263-399 (TMeasyNrParam)
template <class T>
struct VehParam {
    T a, b, h, m, jz, jx, jxz, cf, cr, muf, mur, hrcf, hrcr, krof, kror, brof, bror, maxSteer, crankInertia,
        maxBrakeTorque;
    bool nonLinearSteer, tcbool, throttleMod, driveType, whichWheels;
    std::vector<T> gears;
    std::vector<MapE<T>> shift, steer, torque, losses, cfmap, trmap;
};

template <class T>
struct TireParam {
    T jw, rr, mu, r0, width, rim_radius, pn, pnmax, cx, cy, kt, dx, dy, rdyncoPn, rdyncoP2n, fzRdynco, rdyncoCrit;
    T dfx0Pn, dfx0P2n, fxmPn, fxmP2n, fxsPn, fxsP2n, sxmPn, sxmP2n, sxsPn, sxsP2n;
    T dfy0Pn, dfy0P2n, fymPn, fymP2n, fysPn, fysP2n, symPn, symP2n, sysPn, sysP2n;
    T vcoulomb, frblend_begin, frblend_end, bearingCapacity, li, p_li, p_use;
};

// dof18.h:155-249 (TMeasyState; the Nr state is the subset without xe, ye, xedot, yedot)
template <class T>
struct TireState {
    T xe, ye, xedot, yedot, omega, dOmega, xt, rStat, fx, fy, fz, vsx, vsy, My, engTor;
};

// dof18_gpu.cuh:719-810 (the GPU VehicleState: no dx, dy, tor)
template <class T>
struct VehState {
    T x, y, u, v, psi, wz, phi, wx, udot, vdot, wxdot, wzdot, crankOmega, dOmega_crank;
    int gear;
};

template <class T>
struct Vehicle {
    VehParam<T> vp;
    TireParam<T> tp;
    bool nr = false;
    TireState<T> tire[4];  // LF RF LR RR
    VehState<T> vs;
};

// Reads the flat parameter vector: 25 vehicle scalars, gears, 6 maps, then 35 (TMeasy) or
// 40 (TMeasyNr) tire scalars.  The tire scalars are taken as already passed through tireInit.
template <class T>
void read_params(const double* flat, bool nr, VehParam<T>& v, TireParam<T>& t) {
    Flat f{flat};
    T* vs[] = {&v.a, &v.b, &v.h, &v.m, &v.jz, &v.jx, &v.jxz, &v.cf, &v.cr, &v.muf, &v.mur, &v.hrcf, &v.hrcr,
               &v.krof, &v.kror, &v.brof, &v.bror, &v.maxSteer, &v.crankInertia, &v.maxBrakeTorque};
    for (T* p : vs) *p = T(f.next());
    v.nonLinearSteer = f.next() != 0;
    v.tcbool = f.next() != 0;
    v.throttleMod = f.next() != 0;
    v.driveType = f.next() != 0;
    v.whichWheels = f.next() != 0;
    int ng = (int)f.next();
    v.gears.resize(ng);
    for (int i = 0; i < ng; ++i) v.gears[i] = T(f.next());
    f.map(v.shift);
    f.map(v.steer);
    f.map(v.torque);
    f.map(v.losses);
    f.map(v.cfmap);
    f.map(v.trmap);
    t = TireParam<T>{};
    if (!nr) {
        T* ts[] = {&t.jw, &t.rr, &t.mu, &t.r0, &t.pn, &t.pnmax, &t.cx, &t.cy, &t.kt, &t.dx, &t.dy, &t.rdyncoPn,
                   &t.rdyncoP2n, &t.fzRdynco, &t.rdyncoCrit, &t.dfx0Pn, &t.dfx0P2n, &t.fxmPn, &t.fxmP2n, &t.fxsPn,
                   &t.fxsP2n, &t.sxmPn, &t.sxmP2n, &t.sxsPn, &t.sxsP2n, &t.dfy0Pn, &t.dfy0P2n, &t.fymPn,
                   &t.fymP2n, &t.fysPn, &t.fysP2n, &t.symPn, &t.symP2n, &t.sysPn, &t.sysP2n};
        for (T* p : ts) *p = T(f.next());
    } else {
        T* ts[] = {&t.jw, &t.rr, &t.mu, &t.r0, &t.width, &t.rim_radius, &t.pn, &t.pnmax, &t.kt, &t.rdyncoPn,
                   &t.rdyncoP2n, &t.fzRdynco, &t.rdyncoCrit, &t.dfx0Pn, &t.dfx0P2n, &t.fxmPn, &t.fxmP2n,
                   &t.fxsPn, &t.fxsP2n, &t.sxmPn, &t.sxmP2n, &t.sxsPn, &t.sxsP2n, &t.dfy0Pn, &t.dfy0P2n,
                   &t.fymPn, &t.fymP2n, &t.fysPn, &t.fysP2n, &t.symPn, &t.symP2n, &t.sysPn, &t.sysP2n,
                   &t.vcoulomb, &t.frblend_begin, &t.frblend_end, &t.bearingCapacity, &t.li, &t.p_li, &t.p_use};
        for (T* p : ts) *p = T(f.next());
    }
}

// dof18.cpp:15-42
template <class T, class VP>
T drive_torque(const VP& v, const T& throttle, const T& omega) {
    if (v.throttleMod) {
        std::vector<MapE<T>> m = v.torque;
        for (size_t i = 0; i < m.size(); ++i) {
            m[i].x = v.torque[i].x * throttle;
            m[i].y = v.torque[i].y * throttle;
        }
        T tq = map_y(m, omega);
        T loss = map_y(v.losses, omega);
        return tq + loss;
    }
    T tq = map_y(v.torque, omega);
    T loss = map_y(v.losses, omega);
    return tq * throttle + loss;
}

// dof18.cpp:46-73
template <class T>
void differential_split(const T& torque, double max_bias, const T& sl, const T& sr, T& tl, T& tr) {
    using std::fabs;
    T diff = fabs(sl - sr);
    T bias = T(1.0);
    if (val(diff) > 0.5)
        bias = T(max_bias);
    else if (val(diff) > 0.25)
        bias = 4 * (max_bias - 1) * diff + (2 - max_bias);
    T alpha = bias / (1 + bias);
    T slow = alpha * torque;
    T fast = torque - slow;
    if (std::fabs(val(sl)) < std::fabs(val(sr))) {
        tl = slow;
        tr = fast;
    } else {
        tl = fast;
        tr = slow;
    }
}

template <class T, class VP>
T steer_angle(const VP& v, const T& steering) {
    if (v.nonLinearSteer) return map_y(v.steer, steering);
    return steering * v.maxSteer;
}

// dof18.cpp:75-114 (identical for the Nr tire, :117-156)
template <class T>
void veh_to_tire(Vehicle<T>& s, const T* loads, const T& steering) {
    using std::cos;
    using std::sin;
    const auto& v = s.vp;
    const auto& st = s.vs;
    T delta = steer_angle(v, steering);
    TireState<T>* t = s.tire;
    t[0].fz = loads[0];
    t[0].vsy = st.v + st.wz * v.a;
    t[0].vsx = (st.u - (st.wz * v.cf) / 2.) * cos(delta) + t[0].vsy * sin(delta);
    t[1].fz = loads[1];
    t[1].vsy = st.v + st.wz * v.a;
    t[1].vsx = (st.u + (st.wz * v.cf) / 2.) * cos(delta) + t[1].vsy * sin(delta);
    t[2].fz = loads[2];
    t[2].vsy = st.v - st.wz * v.b;
    t[2].vsx = st.u - (st.wz * v.cr) / 2.;
    t[3].fz = loads[3];
    t[3].vsy = st.v - st.wz * v.b;
    t[3].vsx = st.u + (st.wz * v.cr) / 2.;
}

// dof18.cpp:158-190
template <class T>
void tire_to_veh(Vehicle<T>& s, const T& steering) {
    using std::cos;
    using std::sin;
    T delta = steer_angle(s.vp, steering);
    for (int i = 0; i < 2; ++i) {
        TireState<T>& t = s.tire[i];
        T fx = t.fx * cos(delta) - t.fy * sin(delta);
        T fy = t.fx * sin(delta) + t.fy * cos(delta);
        t.fx = fx;
        t.fy = fy;
    }
}

// dof18.cpp:247-288
template <class T>
void tmxy_combined(T& f, T& fos, const T& s, const T& df0, const T& sm, const T& fm, const T& ss, const T& fs) {
    using std::max;
    T df0loc = T(0.0);
    if (val(sm) > 0.0) {
        T c = 2.0 * fm / sm;
        df0loc = val(c) < val(df0) ? df0 : c;  // std::max(a, b) returns a unless a < b
    }
    if (val(s) > 0.0 && val(df0loc) > 0.0) {
        if (val(s) > val(ss)) {
            f = fs;
            fos = f / s;
        } else {
            if (val(s) < val(sm)) {
                T p = df0loc * sm / fm - 2.0;
                T sn = s / sm;
                T dn = 1.0 + (sn + p) * sn;
                f = df0loc * sm * sn / dn;
                fos = df0loc / dn;
            } else {
                T a = (fm / sm) * (fm / sm) / (df0loc * sm);
                T sstar = sm + (fm - fs) / (a * (ss - sm));
                if (val(sstar) <= val(ss)) {
                    if (val(s) <= val(sstar)) {
                        f = fm - a * (s - sm) * (s - sm);
                    } else {
                        T b = a * (sstar - sm) / (ss - sstar);
                        f = fs + b * (ss - s) * (ss - s);
                    }
                } else {
                    T sn = (s - sm) / (ss - sm);
                    f = fm - (fm - fs) * sn * sn * (3.0 - 2.0 * sn);
                }
                fos = f / s;
            }
        }
    } else {
        f = T(0.0);
        fos = T(0.0);
    }
}

// dof18.cpp:290-306
template <class T>
void coulomb_force(T& fx, T& fy, const T& mu, const T& vsx, const T& vsy, const T& fz, const T& vcoulomb) {
    using std::hypot;
    using std::tanh;
    fx = tanh(-2.0 * vsx / vcoulomb) * fz * mu;
    fy = tanh(-2.0 * vsy / vcoulomb) * fz * mu;
    if (val(hypot(fx, fy)) > val(fz * mu)) {
        T f = fz * mu / hypot(fx, fy);
        fx = fx * f;
        fy = fy * f;
    }
}

// dof18.cpp:308-342 (identical for the Nr tire, :344-378)
template <class T>
void tire_loads(const Vehicle<T>& s, T* loads) {
    const auto& p = s.vp;
    const auto& st = s.vs;
    T huf = s.tp.r0;
    T hur = s.tp.r0;
    T Z1 = (p.m * kG * p.b) / (2. * (p.a + p.b)) + (p.muf * kG) / 2.;
    T Z2 = ((p.muf * huf) / p.cf + p.m * p.b * (p.h - p.hrcf) / (p.cf * (p.a + p.b))) * (st.vdot + st.wz * st.u);
    T Z3 = (p.krof * st.phi + p.brof * st.wx) / p.cf;
    T Z4 = ((p.m * p.h + p.muf * huf + p.mur * hur) * (st.udot - st.wz * st.v)) / (2. * (p.a + p.b));
    T l0 = Z1 - Z2 - Z3 - Z4;
    T l1 = Z1 + Z2 + Z3 - Z4;
    loads[0] = val(l0) > 0. ? l0 : T(0.);
    loads[1] = val(l1) > 0. ? l1 : T(0.);
    Z1 = (p.m * kG * p.a) / (2. * (p.a + p.b)) + (p.mur * kG) / 2.;
    Z2 = ((p.mur * hur) / p.cr + p.m * p.a * (p.h - p.hrcr) / (p.cr * (p.a + p.b))) * (st.vdot + st.wz * st.u);
    Z3 = (p.kror * st.phi + p.bror * st.wx) / p.cr;
    T l2 = Z1 - Z2 - Z3 + Z4;
    T l3 = Z1 + Z2 + Z3 + Z4;
    loads[2] = val(l2) > 0. ? l2 : T(0.);
    loads[3] = val(l3) > 0. ? l3 : T(0.);
}

// dof18.cpp:380-539.  The core (from the effective rolling radius on) is shared with the 11-DOF
// model (identical code, dof11.cpp:312-472) and the 24-DOF model (dof24.cpp:827-987), which
// differ only in how fz and rStat are formed before it.
// The TMeasy force and deflection-rate computation once fz and rStat are known.
template <class T>
void tire_core_tmeasy(TireState<T>& ts, const TireParam<T>& tp, const T& delta, T fz) {
    using std::atan2;
    using std::fabs;
    using std::hypot;
    using std::sqrt;
    using std::tan;
    T vsy = ts.vsy;
    T vsx = ts.vsx;
    T r_eff, rdynco;
    if (val(fz) <= val(tp.fzRdynco)) {
        rdynco = interp_l(fz, tp.rdyncoPn, tp.rdyncoP2n, tp.pn);
        r_eff = rdynco * tp.r0 + (1. - rdynco) * ts.rStat;
    } else {
        rdynco = tp.rdyncoCrit;
        r_eff = rdynco * tp.r0 + (1. - rdynco) * ts.rStat;
    }
    vsx = vsx - (ts.omega * r_eff);
    T vta = r_eff * fabs(ts.omega) + 0.01;
    T sx = -vsx / vta;
    T alpha = atan2(vsy, vta) - delta;
    T sy = -tan(alpha);
    if (val(fz) > val(tp.pnmax)) fz = tp.pnmax;
    T dfx0 = interp_q(fz, tp.dfx0Pn, tp.dfx0P2n, tp.pn);
    T dfy0 = interp_q(fz, tp.dfy0Pn, tp.dfy0P2n, tp.pn);
    T fxm = interp_q(fz, tp.fxmPn, tp.fxmP2n, tp.pn);
    T fym = interp_q(fz, tp.fymPn, tp.fymP2n, tp.pn);
    T fxs = interp_q(fz, tp.fxsPn, tp.fxsP2n, tp.pn);
    T fys = interp_q(fz, tp.fysPn, tp.fysP2n, tp.pn);
    T sxm = interp_l(fz, tp.sxmPn, tp.sxmP2n, tp.pn);
    T sym = interp_l(fz, tp.symPn, tp.symP2n, tp.pn);
    T sxs = interp_l(fz, tp.sxsPn, tp.sxsP2n, tp.pn);
    T sys = interp_l(fz, tp.sysPn, tp.sysP2n, tp.pn);
    T hsxn = sxm / (sxm + sym) + (fxm / dfx0) / (fxm / dfx0 + fym / dfy0);
    T hsyn = sym / (sxm + sym) + (fym / dfy0) / (fxm / dfx0 + fym / dfy0);
    T sxn = sx / hsxn;
    T syn = sy / hsyn;
    T sc = hypot(sxn, syn);
    T calpha, salpha;
    if (val(sc) > 0) {
        calpha = sxn / sc;
        salpha = syn / sc;
    } else {
        calpha = T(sqrt(2.) / 2.);
        salpha = T(sqrt(2.) / 2.);
    }
    T df0 = hypot(dfx0 * calpha * hsxn, dfy0 * salpha * hsyn);
    T fm = hypot(fxm * calpha, fym * salpha);
    T sm = hypot(sxm * calpha / hsxn, sym * salpha / hsyn);
    T fs = hypot(fxs * calpha, fys * salpha);
    T ss = hypot(sxs * calpha / hsxn, sys * salpha / hsyn);
    T f, fos;
    tmxy_combined(f, fos, sc, df0, sm, fm, ss, fs);
    // Fx, Fy (the structural force of :492-499) are computed by the reference but never used.
    ts.My = -sine_step(vta, 0., 0., 0., 1.) * tp.rr * fz * ts.rStat * (double)sgn(ts.omega);
    T vtxs = r_eff * fabs(ts.omega) * hsxn + 0.01;
    T vtys = r_eff * fabs(ts.omega) * hsyn + 0.01;
    ts.xedot = (-vtxs * tp.cx * ts.xe - fos * vsx) / (vtxs * tp.dx + fos);
    ts.yedot = (-vtys * tp.cy * ts.ye - fos * (-sy * vta)) / (vtys * tp.dy + fos);
    T fxdyn = tp.dx * (-vtxs * tp.cx * ts.xe - fos * vsx) / (vtxs * tp.dx + fos) + tp.cx * ts.xe;
    T fydyn = tp.dy * ((-vtys * tp.cy * ts.ye - fos * (-sy * vta)) / (vtys * tp.dy + fos)) + (tp.cy * ts.ye);
    T fxstr = clamp(ts.xe * tp.cx + ts.xedot * tp.dx, -tp.fxmP2n, tp.fxmP2n);
    T fystr = clamp(ts.ye * tp.cy + ts.yedot * tp.dy, -tp.fymP2n, tp.fymP2n);
    T weightx = sine_step(fabs(vsx), 1., 1., 1.5, 0.);
    T weighty = sine_step(fabs(-sy * vta), 1., 1., 1.5, 0.);
    ts.fx = weightx * fxstr + (1. - weightx) * fxdyn;
    ts.fy = weighty * fystr + (1. - weighty) * fydyn;
}

template <class T, class VP>
void tire_rhs_tmeasy(TireState<T>& ts, const TireParam<T>& tp, const VP& vp, const T& steering) {
    T delta = steer_angle(vp, steering);
    T fz = ts.fz;
    if (val(fz) <= 0) {
        ts.xe = T(0); ts.ye = T(0); ts.omega = T(0); ts.xt = T(0); ts.rStat = tp.r0;
        ts.fx = T(0); ts.fy = T(0); ts.My = T(0);
        return;
    }
    ts.xt = fz / tp.kt;
    ts.rStat = tp.r0 - ts.xt;
    tire_core_tmeasy(ts, tp, delta, fz);
}

// dof18.cpp:541-673 (shared like the TMeasy tire: dof11.cpp:473-606, dof24.cpp:988-1121)
template <class T>
void tire_core_tmeasy_nr(TireState<T>& ts, const TireParam<T>& tp, const T& delta, T fz) {
    using std::atan2;
    using std::fabs;
    using std::hypot;
    using std::sqrt;
    using std::tan;
    using std::tanh;
    T vsy = ts.vsy;
    T vsx = ts.vsx;
    T r_eff, rdynco;
    if (val(fz) <= val(tp.fzRdynco)) {
        rdynco = interp_l(fz, tp.rdyncoPn, tp.rdyncoP2n, tp.pn);
        r_eff = rdynco * tp.r0 + (1. - rdynco) * ts.rStat;
    } else {
        rdynco = tp.rdyncoCrit;
        r_eff = rdynco * tp.r0 + (1. - rdynco) * ts.rStat;
    }
    vsx = vsx - (ts.omega * r_eff);
    T vta = r_eff * fabs(ts.omega) + 0.01;
    T Fx0 = T(0), Fy0 = T(0);
    coulomb_force(Fx0, Fy0, tp.mu, vsx, vsy, fz, tp.vcoulomb);
    T sx = -vsx / vta;
    T alpha = atan2(vsy, vta) - delta;
    T sy = -tan(alpha);
    if (val(fz) > val(tp.pnmax)) fz = tp.pnmax;
    T dfx0 = interp_q(fz, tp.dfx0Pn, tp.dfx0P2n, tp.pn);
    T dfy0 = interp_q(fz, tp.dfy0Pn, tp.dfy0P2n, tp.pn);
    T fxm = interp_q(fz, tp.fxmPn, tp.fxmP2n, tp.pn);
    T fym = interp_q(fz, tp.fymPn, tp.fymP2n, tp.pn);
    T fxs = interp_q(fz, tp.fxsPn, tp.fxsP2n, tp.pn);
    T fys = interp_q(fz, tp.fysPn, tp.fysP2n, tp.pn);
    T sxm = interp_l(fz, tp.sxmPn, tp.sxmP2n, tp.pn);
    T sym = interp_l(fz, tp.symPn, tp.symP2n, tp.pn);
    T sxs = interp_l(fz, tp.sxsPn, tp.sxsP2n, tp.pn);
    T sys = interp_l(fz, tp.sysPn, tp.sysP2n, tp.pn);
    T frblend = sine_step(fabs(ts.vsx), val(tp.frblend_begin), 0., val(tp.frblend_end), 1.);
    T sc = hypot(sx, sy);
    T salpha, calpha;
    if (val(sc) > 0) {
        calpha = sx / sc;
        salpha = sy / sc;
    } else {
        calpha = T(sqrt(2.) / 2.);
        salpha = T(sqrt(2.) / 2.);
    }
    T df0 = hypot(dfx0 * calpha, dfy0 * salpha);
    T fm = tp.mu * hypot(fxm * calpha, fym * salpha);
    T sm = tp.mu * hypot(sxm * calpha, sym * salpha);
    T fs = tp.mu * hypot(fxs * calpha, fys * salpha);
    T ss = tp.mu * hypot(sxs * calpha, sys * salpha);
    T f, fos;
    tmxy_combined(f, fos, sc, df0, sm, fm, ss, fs);
    T Fx, Fy;
    if (val(sc) > 0.) {
        Fx = f * sx / sc;
        Fy = f * sy / sc;
    } else {
        Fx = T(0.);
        Fy = T(0.);
    }
    Fx = (1.0 - frblend) * Fx0 + frblend * Fx;
    Fy = (1.0 - frblend) * Fy0 + frblend * Fy;
    ts.My = -tp.rr * fz * ts.rStat * tanh(ts.omega);
    ts.fx = Fx;
    ts.fy = Fy;
}

template <class T, class VP>
void tire_rhs_tmeasy_nr(TireState<T>& ts, const TireParam<T>& tp, const VP& vp, const T& steering) {
    T delta = steer_angle(vp, steering);
    T fz = ts.fz;
    if (val(fz) <= 0) {
        ts.omega = T(0); ts.xt = T(0); ts.rStat = tp.r0; ts.fx = T(0); ts.fy = T(0); ts.My = T(0);
        return;
    }
    ts.xt = fz / tp.kt;
    ts.rStat = tp.r0 - ts.xt;
    tire_core_tmeasy_nr(ts, tp, delta, fz);
}

// dof18.cpp:675-845 (the Nr overload :847-1017 is the same code; the 24-DOF model uses it too,
// dof24.cpp:1338-1509)
template <class T, class V>
void powertrain_rhs(V& s, const T& throttle, const T& braking) {
    using std::fabs;
    const auto& v = s.vp;
    auto& st = s.vs;
    auto* t = s.tire;
    T torque_t = T(0);
    const double max_bias = 2;
    if (v.tcbool) {
        bool reverse = false;
        T omega_t = 0.25 * (t[0].omega + t[1].omega + t[2].omega + t[3].omega);
        T omega_out = omega_t / v.gears[st.gear];
        T omega_in = st.crankOmega;
        T sr, cf, tr;
        if (val(omega_out) < 1e-9 || val(omega_in) < 1e-9) {
            sr = T(0);
            cf = map_y(v.cfmap, sr);
            tr = map_y(v.trmap, sr);
        } else {
            sr = omega_out / omega_in;
            if (val(sr) > 1.) {
                sr = 1. - (sr - 1.);
                reverse = true;
            }
            if (val(sr) < 0) sr = T(0);
            cf = map_y(v.cfmap, sr);
            tr = map_y(v.trmap, sr);
        }
        T q = omega_in / cf;
        T torque_in = -(q * q);  // -std::pow(omega_in / cf, 2); g++ folds pow(x, 2) to x*x
        if (reverse) torque_in = -torque_in;
        T torque_out;
        if (reverse)
            torque_out = -torque_in;
        else
            torque_out = -tr * torque_in;
        torque_t = torque_out / v.gears[st.gear];
        if (val(st.u) < 1e-9 && val(torque_t) < 0) torque_t = T(0);
        st.dOmega_crank = (1. / v.crankInertia) * (drive_torque(v, throttle, st.crankOmega) + torque_in);
        if (val(omega_out) > val(v.shift[st.gear].y)) {
            if (st.gear < (int)v.gears.size() - 1) st.gear++;
        } else if (val(omega_out) < val(v.shift[st.gear].x)) {
            if (st.gear > 0) st.gear--;
        }
    } else {
        st.crankOmega = 0.25 * (t[0].omega + t[1].omega + t[2].omega + t[3].omega) / v.gears[st.gear];
        torque_t = drive_torque(v, throttle, st.crankOmega) / v.gears[st.gear];
        if (val(st.u) < 1e-9 && val(torque_t) < 0) torque_t = T(0);
        if (val(st.crankOmega) > val(v.shift[st.gear].y)) {
            if (st.gear < (int)v.gears.size() - 1) st.gear++;
        } else if (val(st.crankOmega) < val(v.shift[st.gear].x)) {
            if (st.gear > 0) st.gear--;
        }
    }
    T tf = T(0), trr = T(0);
    if (v.driveType) {
        tf = torque_t * 0.5;
        trr = torque_t * 0.5;
    } else if (v.whichWheels) {
        tf = T(0);
        trr = torque_t;
    } else {
        tf = torque_t;
        trr = T(0);
    }
    differential_split(tf, max_bias, t[0].omega, t[1].omega, t[0].engTor, t[1].engTor);
    differential_split(trr, max_bias, t[2].omega, t[3].omega, t[2].engTor, t[3].engTor);
    T bt = v.maxBrakeTorque * braking;  // brakeTorque, dof18.h:956-958
    for (int i = 0; i < 4; ++i) {
        t[i].dOmega = (1 / s.tp.jw) * (t[i].engTor + t[i].My - (double)sgn(t[i].omega) * bt - t[i].fx * t[i].rStat);
    }
}

// dof18.cpp:1019-1078
template <class T>
void veh_rhs(VehState<T>& st, const VehParam<T>& p, const T* fx, const T* fy) {
    T mt = p.m + 2 * (p.muf + p.mur);
    T hrc = (p.hrcf * p.b + p.hrcr * p.a) / (p.a + p.b);
    T E1 = -mt * st.wz * st.u + (fy[0] + fy[1] + fy[2] + fy[3]);
    T E2 = (fy[0] + fy[1]) * p.a - (fy[2] + fy[3]) * p.b + (fx[1] - fx[0]) * p.cf / 2 + (fx[3] - fx[2]) * p.cr / 2 +
           (-p.muf * p.a + p.mur * p.b) * st.wz * st.u;
    T E3 = p.m * kG * hrc * st.phi - (p.krof + p.kror) * st.phi - (p.brof + p.bror) * st.wx +
           hrc * p.m * st.wz * st.u;
    T A1 = p.mur * p.b - p.muf * p.a;
    T A2 = p.jx + p.m * (hrc * hrc);
    T A3 = hrc * p.m;
    st.udot = st.wz * st.v + (1 / mt) * ((fx[0] + fx[1] + fx[2] + fx[3]) +
                                         (-p.mur * p.b + p.muf * p.a) * (st.wz * st.wz) -
                                         2. * hrc * p.m * st.wz * st.wx);
    T denom = (A2 * (A1 * A1) - 2. * A1 * A3 * p.jxz + p.jz * (A3 * A3) + mt * (p.jxz * p.jxz) - A2 * p.jz * mt);
    st.vdot = (E1 * (p.jxz * p.jxz) - A1 * A2 * E2 + A1 * E3 * p.jxz + A3 * E2 * p.jxz - A2 * E1 * p.jz -
               A3 * E3 * p.jz) /
              denom;
    st.wxdot = ((A1 * A1) * E3 - A1 * A3 * E2 + A1 * E1 * p.jxz - A3 * E1 * p.jz + E2 * p.jxz * mt -
                E3 * p.jz * mt) /
               denom;
    st.wzdot = ((A3 * A3) * E2 - A1 * A2 * E1 - A1 * A3 * E3 + A3 * E1 * p.jxz - A2 * E2 * mt + E3 * p.jxz * mt) /
               denom;
}

// rhsFun order (dof18_halfImplicit.cpp:618-673), then the step update (:245-318).
template <class T>
void step(Vehicle<T>& s, const T& steering, const T& throttle, const T& braking, double dt) {
    using std::cos;
    using std::sin;
    T loads[4];
    tire_loads(s, loads);
    veh_to_tire(s, loads, steering);
    if (!s.nr) {
        tire_rhs_tmeasy(s.tire[0], s.tp, s.vp, steering);
        tire_rhs_tmeasy(s.tire[1], s.tp, s.vp, steering);
        tire_rhs_tmeasy(s.tire[2], s.tp, s.vp, T(0.0));
        tire_rhs_tmeasy(s.tire[3], s.tp, s.vp, T(0.0));
    } else {
        tire_rhs_tmeasy_nr(s.tire[0], s.tp, s.vp, steering);
        tire_rhs_tmeasy_nr(s.tire[1], s.tp, s.vp, steering);
        tire_rhs_tmeasy_nr(s.tire[2], s.tp, s.vp, T(0.0));
        tire_rhs_tmeasy_nr(s.tire[3], s.tp, s.vp, T(0.0));
    }
    powertrain_rhs<T>(s, throttle, braking);
    tire_to_veh(s, steering);
    T fx[4] = {s.tire[0].fx, s.tire[1].fx, s.tire[2].fx, s.tire[3].fx};
    T fy[4] = {s.tire[0].fy, s.tire[1].fy, s.tire[2].fy, s.tire[3].fy};
    veh_rhs(s.vs, s.vp, fx, fy);

    for (int i = 0; i < 4; ++i) {
        TireState<T>& t = s.tire[i];
        if (!s.nr) {
            t.xe = t.xe + t.xedot * dt;
            t.ye = t.ye + t.yedot * dt;
        }
        t.omega = t.omega + t.dOmega * dt;
    }
    VehState<T>& v = s.vs;
    if (s.vp.tcbool) v.crankOmega = v.crankOmega + v.dOmega_crank * dt;
    v.u = v.u + v.udot * dt;
    v.v = v.v + v.vdot * dt;
    v.wx = v.wx + v.wxdot * dt;
    v.wz = v.wz + v.wzdot * dt;
    v.x = v.x + (v.u * cos(v.psi) - v.v * sin(v.psi)) * dt;
    v.y = v.y + (v.u * sin(v.psi) + v.v * cos(v.psi)) * dt;
    v.psi = v.psi + v.wz * dt;
    v.phi = v.phi + v.wx * dt;
}

// Canonical record (15 per TMeasy tire / 11 per Nr tire, then 15 vehicle fields).
inline int record_len(bool nr) { return nr ? 4 * 11 + 15 : 4 * 15 + 15; }

template <class T>
void to_record(const Vehicle<T>& s, double* o) {
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        const TireState<T>& t = s.tire[i];
        if (!s.nr) {
            o[k++] = val(t.xe); o[k++] = val(t.ye); o[k++] = val(t.xedot); o[k++] = val(t.yedot);
        }
        o[k++] = val(t.omega); o[k++] = val(t.dOmega); o[k++] = val(t.xt); o[k++] = val(t.rStat);
        o[k++] = val(t.fx); o[k++] = val(t.fy); o[k++] = val(t.fz); o[k++] = val(t.vsx); o[k++] = val(t.vsy);
        o[k++] = val(t.My); o[k++] = val(t.engTor);
    }
    const VehState<T>& v = s.vs;
    const T f[] = {v.x, v.y, v.u, v.v, v.psi, v.wz, v.phi, v.wx, v.udot, v.vdot, v.wxdot, v.wzdot, v.crankOmega,
                   v.dOmega_crank};
    for (const T& x : f) o[k++] = val(x);
    o[k++] = (double)v.gear;
}

// Reads a record through src(k), which returns record entry k as a T (a dual scalar seeds its
// tangents there).
template <class T, class Src>
void from_record_src(Vehicle<T>& s, Src&& src) {
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        TireState<T>& t = s.tire[i];
        if (!s.nr) {
            t.xe = src(k++); t.ye = src(k++); t.xedot = src(k++); t.yedot = src(k++);
        } else {
            t.xe = t.ye = t.xedot = t.yedot = T(0.0);
        }
        t.omega = src(k++); t.dOmega = src(k++); t.xt = src(k++); t.rStat = src(k++);
        t.fx = src(k++); t.fy = src(k++); t.fz = src(k++); t.vsx = src(k++); t.vsy = src(k++);
        t.My = src(k++); t.engTor = src(k++);
    }
    VehState<T>& v = s.vs;
    T* f[] = {&v.x, &v.y, &v.u, &v.v, &v.psi, &v.wz, &v.phi, &v.wx, &v.udot, &v.vdot, &v.wxdot, &v.wzdot,
              &v.crankOmega, &v.dOmega_crank};
    for (T* p : f) *p = src(k++);
    v.gear = (int)val(src(k++));
}

template <class T>
void from_record(Vehicle<T>& s, const double* o) {
    from_record_src(s, [&](int k) { return T(o[k]); });
}

}  // namespace d18
}  // namespace orc
