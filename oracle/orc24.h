// ORACLE — test infrastructure only (see orc_common.h).
//
// CPU restatement of the reference 24-DOF model (heave, pitch, roll, four suspension struts,
// tire compression) and its half-implicit step: wheeled_vehicle_models/24dof/dof24.cpp and
// dof24_halfImplicit.cpp, same operation order.  The tire force core is the 18-DOF one.
#pragma once
#include <cmath>
#include <vector>

#include "orc18.h"

namespace orc {
namespace d24 {

using d18::TireParam;

// dof24.h:580-764
template <class T>
struct VehParam {
    T a, b, h, m, jz, jx, jy, cf, cr, muf, mur, hrcf, hrcr, maxSteer, crankInertia, maxBrakeTorque;
    bool nonLinearSteer, tcbool, throttleMod, driveType, whichWheels;
    std::vector<T> gears;
    std::vector<MapE<T>> shift, steer, torque, losses, cfmap, trmap;
    // the 18-DOF drive/split helpers read these names
};

template <class T>
struct SusParam {
    T ks, bs;
};

// dof24.h:161-297: the 18-DOF tire state plus compression rate, ground-frame forces and velocities
template <class T>
struct TireState : d18::TireState<T> {
    T dxt, fxg, fyg, fzg, vgx, vgy, vgz, vsz;
};

// dof24.h:958-1095
template <class T>
struct SusState {
    T xs, xsi, dxs, ls, lsi, us, vs, ws, dus, dvs, dws, uu, vu, wu, duu, dvu, dwu, fxs, fys, fzs, fzd, mx, my, mz;
};

// dof24.h:773-927 (without the CPU-only dx, dy, tor)
template <class T>
struct VehState {
    T x, y, z, u, v, w, psi, wz, phi, wx, theta, wy, udot, vdot, wdot, wxdot, wydot, wzdot, dtheta, dphi, dpsi,
        crankOmega, dOmega_crank;
    int gear;
};

template <class T>
struct Vehicle {
    VehParam<T> vp;
    TireParam<T> tp;
    SusParam<T> sp;
    bool nr = false;
    TireState<T> tire[4];  // LF RF LR RR
    SusState<T> sus[4];
    VehState<T> vs;
};

// Flat 24-DOF vector: a b h m jz jx jy cf cr muf mur hrcf hrcr maxSteer crankInertia maxBrakeTorque,
// 5 flags, gears, six maps, the 35 / 40 tire scalars, then ks bs.
template <class T>
void read_params(const double* flat, bool nr, VehParam<T>& v, TireParam<T>& t, SusParam<T>& sp) {
    Flat f{flat};
    T* vs[] = {&v.a, &v.b, &v.h, &v.m, &v.jz, &v.jx, &v.jy, &v.cf, &v.cr, &v.muf, &v.mur, &v.hrcf, &v.hrcr,
               &v.maxSteer, &v.crankInertia, &v.maxBrakeTorque};
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
    sp.ks = T(f.next());
    sp.bs = T(f.next());
}

// initializeTireSus (dof24.cpp:34-89): static tire loads and compressions, strut preload and
// lengths, unsprung heave velocity.  Called by Initialize (dof24_halfImplicit.cpp:119-121).
template <class T>
void init_tire_sus(Vehicle<T>& s) {
    const auto& p = s.vp;
    T ws = ((p.m * kG * p.b) / (2 * (p.a + p.b)) + p.muf * kG);
    s.tire[0].fz = ws;
    s.tire[1].fz = ws;
    ws = ((p.m * kG * p.a) / (2 * (p.a + p.b)) + p.mur * kG);
    s.tire[2].fz = ws;
    s.tire[3].fz = ws;
    for (int i = 0; i < 4; ++i) s.tire[i].xt = s.tire[i].fz / s.tp.kt;
    s.sus[0].xsi = (p.m * kG * p.b) / (2 * (p.a + p.b) * s.sp.ks);
    s.sus[2].xsi = (p.m * kG * p.a) / (2 * (p.a + p.b) * s.sp.ks);
    s.sus[1].xsi = s.sus[0].xsi;
    s.sus[3].xsi = s.sus[2].xsi;
    for (int i = 0; i < 4; ++i) s.sus[i].xs = s.sus[i].xsi;
    for (int i = 0; i < 4; ++i) s.sus[i].lsi = p.h - (s.tp.r0 - s.tire[i].xt);
    for (int i = 0; i < 4; ++i) s.sus[i].ls = s.sus[i].lsi;
    for (int i = 0; i < 4; ++i) s.sus[i].wu = s.vs.w;
}

// vehToSusTransform (dof24.cpp:148-230)
template <class T>
ORC_REF_FN void veh_to_sus(Vehicle<T>& s) {
    const auto& p = s.vp;
    const auto& v = s.vs;
    SusState<T>* q = s.sus;
    q[0].us = (-p.cf * v.wz / 2.) + v.u;
    q[0].vs = p.a * v.wz + v.v;
    q[0].ws = (p.cf * v.wx / 2.) - (p.a * v.wy) + v.w;
    q[1].us = (p.cf * v.wz / 2.) + v.u;
    q[1].vs = p.a * v.wz + v.v;
    q[1].ws = (-p.cf * v.wx / 2.) - (p.a * v.wy) + v.w;
    q[2].us = (-p.cr * v.wz / 2.) + v.u;
    q[2].vs = -p.b * v.wz + v.v;
    q[2].ws = (p.cr * v.wx / 2.) + p.b * v.wy + v.w;
    q[3].us = (p.cr * v.wz / 2.) + v.u;
    q[3].vs = -p.b * v.wz + v.v;
    q[3].ws = (-p.cr * v.wx / 2.) + p.b * v.wy + v.w;
    for (int i = 0; i < 4; ++i) q[i].ls = q[i].lsi - (q[i].xs - q[i].xsi);
    for (int i = 0; i < 4; ++i) q[i].uu = q[i].us - q[i].ls * v.wy;
    for (int i = 0; i < 4; ++i) q[i].vu = q[i].vs + q[i].ls * v.wx;
    q[0].dus = (-p.cf * v.wzdot) / 2 + v.udot;
    q[0].dvs = p.a * v.wzdot + v.vdot;
    q[1].dus = (p.cf * v.wzdot) / 2 + v.udot;
    q[1].dvs = p.a * v.wzdot + v.vdot;
    q[2].dus = (-p.cr * v.wzdot) / 2 + v.udot;
    q[2].dvs = -p.b * v.wzdot + v.vdot;
    q[3].dus = (p.cr * v.wzdot) / 2 + v.udot;
    q[3].dvs = (-p.b * v.wzdot) + v.vdot;
    for (int i = 0; i < 4; ++i) {
        q[i].duu = q[i].dus - (-q[i].dxs * v.wy + q[i].ls * v.wydot);
        q[i].dvu = q[i].dvs + (-q[i].dxs * v.wx + q[i].ls * v.wxdot);
    }
}

// vehToTireTransform (dof24.cpp:313-385); reads the previous step's rStat
template <class T>
ORC_REF_FN void veh_to_tire(Vehicle<T>& s, const T& steering) {
    using std::cos;
    using std::sin;
    const auto& v = s.vs;
    T delta = d18::steer_angle(s.vp, steering);
    for (int i = 0; i < 4; ++i) {
        TireState<T>& t = s.tire[i];
        const SusState<T>& q = s.sus[i];
        t.vgx = cos(v.theta) * (q.uu - v.wy * t.rStat) +
                sin(v.theta) * (q.wu * cos(v.phi) + sin(v.phi) * (v.wx * t.rStat + q.vu));
        t.vgy = cos(v.phi) * (q.vu + v.wx * t.rStat) - q.wu * sin(v.phi);
    }
    for (int i = 0; i < 2; ++i) {
        TireState<T>& t = s.tire[i];
        t.vsy = t.vgy;
        t.vsx = t.vgx * cos(delta) + t.vgy * sin(delta);
    }
    for (int i = 2; i < 4; ++i) {
        TireState<T>& t = s.tire[i];
        t.vsy = t.vgy;
        t.vsx = t.vgx;
    }
}

// computeTireRHS (dof24.cpp:827-987, Nr :988-1121): fz from the compression state, rStat from the
// attitude, then the 18-DOF force core.
template <class T>
ORC_REF_FN void tire_rhs(Vehicle<T>& s, TireState<T>& ts, const T& steering) {
    using std::cos;
    const auto& v = s.vs;
    const auto& tp = s.tp;
    T delta = d18::steer_angle(s.vp, steering);
    ts.fz = ts.xt * tp.kt;
    T fz = ts.fz;
    ts.rStat = (tp.r0 - ts.xt) / (cos(v.theta) * cos(v.phi));
    if (val(fz) <= 0) {
        if (!s.nr) {
            ts.xe = T(0);
            ts.ye = T(0);
        }
        ts.omega = T(0);
        ts.xt = T(0);
        ts.rStat = tp.r0;
        ts.fz = T(0);
        ts.fx = T(0);
        ts.fy = T(0);
        ts.My = T(0);
        return;
    }
    if (!s.nr)
        d18::tire_core_tmeasy<T>(ts, tp, delta, fz);
    else
        d18::tire_core_tmeasy_nr<T>(ts, tp, delta, fz);
}

// computeTireCompressionVelocity (dof24.cpp:1122-1151)
template <class T>
ORC_REF_FN void tire_compression(Vehicle<T>& s) {
    using std::cos;
    using std::sin;
    const auto& v = s.vs;
    for (int i = 0; i < 4; ++i) {
        const SusState<T>& q = s.sus[i];
        s.tire[i].dxt = s.tire[i].vsz - (cos(v.theta) * (q.wu * cos(v.phi) + q.vu * sin(v.phi)) - q.uu * sin(v.theta));
    }
}

// tireToVehTransform (dof24.cpp:459-528)
template <class T>
ORC_REF_FN void tire_to_veh(Vehicle<T>& s, const T& steering) {
    using std::cos;
    using std::sin;
    const auto& v = s.vs;
    T delta = d18::steer_angle(s.vp, steering);
    for (int i = 0; i < 2; ++i) {
        TireState<T>& t = s.tire[i];
        T fx = t.fx * cos(delta) - t.fy * sin(delta);
        T fy = t.fx * sin(delta) + t.fy * cos(delta);
        t.fx = fx;
        t.fy = fy;
    }
    for (int i = 0; i < 4; ++i) {
        TireState<T>& t = s.tire[i];
        t.fxg = t.fx * cos(v.theta) - t.fz * sin(v.theta);
        t.fyg = t.fx * sin(v.phi) * sin(v.theta) + t.fy * cos(v.phi) + t.fz * sin(v.phi) * cos(v.theta);
        t.fzg = t.fx * sin(v.theta) * cos(v.phi) - t.fy * sin(v.phi) + t.fz * cos(v.phi) * cos(v.theta);
    }
}

// computeSusRHS (dof24.cpp:1182-1229): reads last step's fzd and dxs, then updates dxs
template <class T>
ORC_REF_FN void sus_rhs(Vehicle<T>& s) {
    using std::cos;
    using std::sin;
    const auto& v = s.vs;
    const auto& p = s.vp;
    T theta = v.theta, phi = v.phi;
    for (int i = 0; i < 4; ++i) {
        const TireState<T>& t = s.tire[i];
        SusState<T>& q = s.sus[i];
        T mu = i < 2 ? p.muf : p.mur;
        q.dwu = (1 / mu) * (cos(phi) * (cos(theta) * (t.fz - mu * kG) + sin(theta) * t.fx) - sin(phi) * t.fy - q.fzd -
                            q.xs * s.sp.ks - q.dxs * s.sp.bs - mu * (q.vu * v.wx - q.uu * v.wy));
    }
    for (int i = 0; i < 4; ++i) s.sus[i].dxs = -s.sus[i].ws + s.sus[i].wu;
}

// computeForcesThroughSus (dof24.cpp:599-680), including the rear-right lateral force that the
// reference writes without the cos(theta) factor (:639)
template <class T>
ORC_REF_FN void forces_through_sus(Vehicle<T>& s) {
    using std::cos;
    using std::sin;
    const auto& v = s.vs;
    const auto& p = s.vp;
    for (int i = 0; i < 4; ++i) {
        const TireState<T>& t = s.tire[i];
        SusState<T>& q = s.sus[i];
        T mu = i < 2 ? p.muf : p.mur;
        q.fxs = t.fxg + mu * kG * sin(v.theta) - mu * q.duu + mu * v.wz * q.vu - mu * v.wy * q.wu;
        if (i < 3)
            q.fys = t.fyg - mu * kG * sin(v.phi) * cos(v.theta) - mu * q.dvu + mu * v.wx * q.wu - mu * v.wz * q.uu;
        else
            q.fys = t.fyg - mu * kG * sin(v.phi) - mu * q.dvu + mu * v.wx * q.wu - mu * v.wz * q.uu;
        q.fzs = q.xs * s.sp.ks + q.dxs * s.sp.bs;
    }
    TireState<T>* t = s.tire;
    SusState<T>* q = s.sus;
    q[1].fzd = (t[1].fyg * t[1].rStat + q[1].fys * q[1].ls + t[0].fyg * t[0].rStat + q[0].fys * q[0].ls -
                (q[1].fys + q[0].fys) * p.hrcf) /
               p.cf;
    q[0].fzd = -q[1].fzd;
    q[3].fzd = (t[3].fyg * t[3].rStat + q[3].fys * q[3].ls + t[2].fyg * t[2].rStat + q[2].fys * q[2].ls -
                (q[3].fys + q[2].fys) * p.hrcr) /
               p.cr;
    q[2].fzd = -q[3].fzd;
    q[0].mx = q[0].fys * p.hrcf;
    q[1].mx = q[1].fys * p.hrcf;
    q[2].mx = q[2].fys * p.hrcr;
    q[3].mx = q[3].fys * p.hrcr;
    for (int i = 0; i < 4; ++i) q[i].my = -(t[i].fxg * t[i].rStat + q[i].fxs * q[i].ls);
    for (int i = 0; i < 4; ++i) q[i].mz = T(0);
}

// computeVehicleRHS (dof24.cpp:1682-1733)
template <class T>
ORC_REF_FN void veh_rhs(Vehicle<T>& s) {
    using std::cos;
    using std::sin;
    using std::tan;
    auto& v = s.vs;
    const auto& p = s.vp;
    const SusState<T>* q = s.sus;
    v.udot = v.wz * v.v - v.wy * v.w + (1 / p.m) * (q[0].fxs + q[1].fxs + q[2].fxs + q[3].fxs) + kG * sin(v.theta);
    v.vdot = v.wx * v.w - v.wz * v.u + (1 / p.m) * (q[0].fys + q[1].fys + q[2].fys + q[3].fys) -
             kG * sin(v.phi) * cos(v.theta);
    v.wdot = v.wy * v.u - v.wx * v.v +
             (1 / p.m) * (q[0].fzs + q[1].fzs + q[2].fzs + q[3].fzs + q[0].fzd + q[1].fzd + q[2].fzd + q[3].fzd) -
             kG * cos(v.phi) * cos(v.theta);
    v.wxdot = (1 / p.jx) * ((q[0].mx + q[1].mx + q[2].mx + q[3].mx) + (q[0].fzs - q[1].fzs) * (p.cf / 2.) +
                            (q[2].fzs - q[3].fzs) * (p.cr / 2.));
    v.wydot = (1 / p.jy) *
              ((q[0].my + q[1].my + q[2].my + q[3].my) + (q[2].fzs + q[3].fzs) * p.b - (q[0].fzs + q[1].fzs) * p.a);
    v.wzdot = (1 / p.jz) * ((q[0].mz + q[1].mz + q[2].mz + q[3].mz) + (q[0].fys + q[1].fys) * p.a -
                            (q[2].fys + q[3].fys) * p.b + (-q[0].fxs + q[1].fxs) * (p.cf / 2.) +
                            (-q[2].fxs + q[3].fxs) * (p.cr / 2.));
    v.dtheta = v.wy * cos(v.phi) - v.wz * sin(v.phi);
    v.dpsi = (v.wy * sin(v.phi) / cos(v.theta)) + (v.wz * cos(v.phi) / cos(v.theta));
    v.dphi = v.wx + v.wy * sin(v.phi) * tan(v.theta) + v.wz * cos(v.phi) * tan(v.theta);
}

// rhsFun order (dof24_halfImplicit.cpp:594-706) and the step update (:188-297)
template <class T>
void step(Vehicle<T>& s, const T& steering, const T& throttle, const T& braking, double dt) {
    using std::cos;
    using std::sin;
    using std::tan;
    veh_to_sus(s);
    veh_to_tire(s, steering);
    tire_rhs(s, s.tire[0], steering);
    tire_rhs(s, s.tire[1], steering);
    tire_rhs(s, s.tire[2], T(0.0));
    tire_rhs(s, s.tire[3], T(0.0));
    tire_compression(s);
    d18::powertrain_rhs<T>(s, throttle, braking);  // dof24.cpp:1338-1509 is the 18-DOF code
    tire_to_veh(s, steering);
    sus_rhs(s);
    forces_through_sus(s);
    veh_rhs(s);
    for (int i = 0; i < 4; ++i) {
        TireState<T>& t = s.tire[i];
        if (!s.nr) {
            t.xe = t.xe + t.xedot * dt;
            t.ye = t.ye + t.yedot * dt;
        }
        t.xt = t.xt + t.dxt * dt;
        t.omega = t.omega + t.dOmega * dt;
    }
    for (int i = 0; i < 4; ++i) {
        s.sus[i].wu = s.sus[i].wu + s.sus[i].dwu * dt;
        s.sus[i].xs = s.sus[i].xs + s.sus[i].dxs * dt;
    }
    VehState<T>& v = s.vs;
    if (s.vp.tcbool) v.crankOmega = v.crankOmega + v.dOmega_crank * dt;
    v.u = v.u + v.udot * dt;
    v.v = v.v + v.vdot * dt;
    v.w = v.w + v.wdot * dt;
    v.wx = v.wx + v.wxdot * dt;
    v.wy = v.wy + v.wydot * dt;
    v.wz = v.wz + v.wzdot * dt;
    v.dtheta = v.wy * cos(v.phi) - v.wz * sin(v.phi);
    v.dpsi = (v.wy * sin(v.phi) / cos(v.theta)) + (v.wz * cos(v.phi) / cos(v.theta));
    v.dphi = v.wx + v.wy * sin(v.phi) * tan(v.theta) + v.wz * cos(v.phi) * tan(v.theta);
    v.theta = v.theta + v.dtheta * dt;
    v.psi = v.psi + v.dpsi * dt;
    v.phi = v.phi + v.dphi * dt;
    v.x = v.x + (v.u * cos(v.psi) - v.v * sin(v.psi)) * dt;
    v.y = v.y + (v.u * sin(v.psi) + v.v * cos(v.psi)) * dt;
}

// Record: per tire 23 (TMeasy) / 19 (Nr), per strut 24, vehicle 24.
inline int tire_len(bool nr) { return nr ? 19 : 23; }
inline int record_len(bool nr) { return 4 * tire_len(nr) + 4 * 24 + 24; }

template <class T>
void to_record(const Vehicle<T>& s, double* o) {
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        const TireState<T>& t = s.tire[i];
        if (!s.nr) {
            o[k++] = val(t.xe); o[k++] = val(t.ye); o[k++] = val(t.xedot); o[k++] = val(t.yedot);
        }
        const T f[] = {t.omega, t.dOmega, t.xt, t.dxt, t.rStat, t.fx, t.fy, t.fz, t.fxg, t.fyg,
                       t.fzg,   t.vgx,    t.vgy, t.vgz, t.vsx,  t.vsy, t.vsz, t.My, t.engTor};
        for (const T& x : f) o[k++] = val(x);
    }
    for (int i = 0; i < 4; ++i) {
        const SusState<T>& q = s.sus[i];
        const T f[] = {q.xs, q.xsi, q.dxs, q.ls, q.lsi, q.us, q.vs, q.ws, q.dus, q.dvs, q.dws, q.uu,
                       q.vu, q.wu,  q.duu, q.dvu, q.dwu, q.fxs, q.fys, q.fzs, q.fzd, q.mx, q.my, q.mz};
        for (const T& x : f) o[k++] = val(x);
    }
    const VehState<T>& v = s.vs;
    const T f[] = {v.x,    v.y,     v.z,     v.u,     v.v,      v.w,      v.psi,  v.wz,
                   v.phi,  v.wx,    v.theta, v.wy,    v.udot,   v.vdot,   v.wdot, v.wxdot,
                   v.wydot, v.wzdot, v.dtheta, v.dphi, v.dpsi,   v.crankOmega, v.dOmega_crank};
    for (const T& x : f) o[k++] = val(x);
    o[k++] = (double)v.gear;
}

template <class T>
void from_record(Vehicle<T>& s, const double* o) {
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        TireState<T>& t = s.tire[i];
        if (!s.nr) {
            t.xe = T(o[k++]); t.ye = T(o[k++]); t.xedot = T(o[k++]); t.yedot = T(o[k++]);
        } else {
            t.xe = t.ye = t.xedot = t.yedot = T(0.0);
        }
        T* f[] = {&t.omega, &t.dOmega, &t.xt, &t.dxt, &t.rStat, &t.fx, &t.fy, &t.fz, &t.fxg, &t.fyg,
                  &t.fzg,   &t.vgx,    &t.vgy, &t.vgz, &t.vsx,  &t.vsy, &t.vsz, &t.My, &t.engTor};
        for (T* p : f) *p = T(o[k++]);
    }
    for (int i = 0; i < 4; ++i) {
        SusState<T>& q = s.sus[i];
        T* f[] = {&q.xs, &q.xsi, &q.dxs, &q.ls, &q.lsi, &q.us, &q.vs, &q.ws, &q.dus, &q.dvs, &q.dws, &q.uu,
                  &q.vu, &q.wu,  &q.duu, &q.dvu, &q.dwu, &q.fxs, &q.fys, &q.fzs, &q.fzd, &q.mx, &q.my, &q.mz};
        for (T* p : f) *p = T(o[k++]);
    }
    VehState<T>& v = s.vs;
    T* f[] = {&v.x,     &v.y,     &v.z,      &v.u,    &v.v,    &v.w,    &v.psi,  &v.wz,
              &v.phi,   &v.wx,    &v.theta,  &v.wy,   &v.udot, &v.vdot, &v.wdot, &v.wxdot,
              &v.wydot, &v.wzdot, &v.dtheta, &v.dphi, &v.dpsi, &v.crankOmega, &v.dOmega_crank};
    for (T* p : f) *p = T(o[k++]);
    v.gear = (int)o[k++];
}

}  // namespace d24
}  // namespace orc
