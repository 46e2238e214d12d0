// ORACLE — test infrastructure only (see orc_common.h).
//
// CPU restatement of the reference 11-DOF (single-track) model and its half-implicit step:
// wheeled_vehicle_models/11dof/dof11.cpp and dof11_halfImplicit.cpp, same operation order.  The
// tire right-hand sides are the 18-DOF ones (dof11.cpp:312-606 is a copy of dof18.cpp:380-673).
#pragma once
#include <cmath>
#include <vector>

#include "orc18.h"

namespace orc {
namespace d11 {

using d18::TireParam;
using d18::TireState;

// dof11.h:496-620
template <class T>
struct VehParam {
    T a, b, h, m, jz, muf, mur, maxSteer, crankInertia, maxBrakeTorque;
    bool nonLinearSteer, tcbool, throttleMod, driveType, whichWheels;
    std::vector<T> gears;
    std::vector<MapE<T>> shift, steer, torque, losses, cfmap, trmap;
};

// dof11_gpu.cuh:686-765 (x y u v psi wz udot vdot wzdot crankOmega dOmega_crank gear)
template <class T>
struct VehState {
    T x, y, u, v, psi, wz, udot, vdot, wzdot, crankOmega, dOmega_crank;
    int gear;
};

template <class T>
struct Vehicle {
    VehParam<T> vp;
    TireParam<T> tp;
    bool nr = false;
    TireState<T> tire[2];  // front, rear
    VehState<T> vs;
};

// Flat 11-DOF vector: a b h m jz muf mur maxSteer crankInertia maxBrakeTorque, 5 flags, gears,
// six maps, then the tire scalars (already scaled by the 11-DOF loader and passed through tireInit).
template <class T>
void read_params(const double* flat, bool nr, VehParam<T>& v, TireParam<T>& t) {
    Flat f{flat};
    T* vs[] = {&v.a, &v.b, &v.h, &v.m, &v.jz, &v.muf, &v.mur, &v.maxSteer, &v.crankInertia, &v.maxBrakeTorque};
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
    // tire block: same layout as the 18-DOF vector's
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

// dof11.cpp:15-42 (same as the 18-DOF driveTorque)
template <class T>
T drive_torque(const VehParam<T>& v, const T& throttle, const T& omega) {
    if (v.throttleMod) {
        std::vector<MapE<T>> m = v.torque;
        for (size_t i = 0; i < m.size(); ++i) {
            m[i].x = v.torque[i].x * throttle;
            m[i].y = v.torque[i].y * throttle;
        }
        return map_y(m, omega) + map_y(v.losses, omega);
    }
    T tq = map_y(v.torque, omega);
    T loss = map_y(v.losses, omega);
    return tq * throttle + loss;
}

// dof11.cpp:271-311: split = 4WD; otherwise whichWheels picks the driven axle.
template <class T>
void differential_split(const T& torque, double max_bias, const T& s_rear, const T& s_front, T& t_rear, T& t_front,
                        bool split, bool which) {
    using std::fabs;
    T diff = fabs(s_rear - s_front);
    if (split) {
        T bias = T(1.0);
        if (val(diff) > 0.5)
            bias = T(max_bias);
        else if (val(diff) > 0.25)
            bias = 4 * (max_bias - 1) * diff + (2 - max_bias);
        T alpha = bias / (1 + bias);
        T slow = alpha * torque;
        T fast = torque - slow;
        if (std::fabs(val(s_rear)) < std::fabs(val(s_front))) {
            t_rear = slow;
            t_front = fast;
        } else {
            t_rear = fast;
            t_front = slow;
        }
    } else if (which) {
        t_rear = torque;
        t_front = T(0);
    } else {
        t_rear = T(0);
        t_front = torque;
    }
}

// dof11.cpp:229-247
template <class T>
void tire_loads(const Vehicle<T>& s, T* loads) {
    const auto& p = s.vp;
    const auto& st = s.vs;
    T huf = s.tp.r0, hur = s.tp.r0;
    T Z1 = (p.m * kG * p.b) / (2. * (p.a + p.b)) + (p.muf * kG) / 2.;
    T Z2 = ((p.m * p.h + p.muf * huf + p.mur * hur) * (st.udot - st.wz * st.v)) / (2. * (p.a + p.b));
    T l0 = Z1 - Z2;
    loads[0] = val(l0) > 0. ? l0 : T(0.);
    T Z3 = (p.m * kG * p.a) / (2. * (p.a + p.b)) + (p.mur * kG) / 2.;
    T l1 = Z3 + Z2;
    loads[1] = val(l1) > 0. ? l1 : T(0.);
}

// dof11.cpp:607-749 (Nr overload :750-892 is the same code)
template <class T>
void powertrain_rhs(Vehicle<T>& s, const T& throttle, const T& braking) {
    const auto& v = s.vp;
    auto& st = s.vs;
    TireState<T>* t = s.tire;
    T torque_t = T(0);
    const double max_bias = 2;
    if (v.tcbool) {
        bool reverse = false;
        T omega_t = 0.5 * (t[0].omega + t[1].omega);
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
        T torque_in = -(q * q);
        if (reverse) torque_in = -torque_in;
        T torque_out = reverse ? -torque_in : -tr * torque_in;
        torque_t = torque_out / v.gears[st.gear];
        if (val(st.u) < 1e-9 && val(torque_t) < 0) torque_t = T(0);
        st.dOmega_crank = (1. / v.crankInertia) * (drive_torque(v, throttle, st.crankOmega) + torque_in);
        if (val(omega_out) > val(v.shift[st.gear].y)) {
            if (st.gear < (int)v.gears.size() - 1) st.gear++;
        } else if (val(omega_out) < val(v.shift[st.gear].x)) {
            if (st.gear > 0) st.gear--;
        }
    } else {
        st.crankOmega = 0.5 * (t[0].omega + t[1].omega) / v.gears[st.gear];
        torque_t = drive_torque(v, throttle, st.crankOmega) / v.gears[st.gear];
        if (val(st.u) < 1e-9 && val(torque_t) < 0) torque_t = T(0);
        if (val(st.crankOmega) > val(v.shift[st.gear].y)) {
            if (st.gear < (int)v.gears.size() - 1) st.gear++;
        } else if (val(st.crankOmega) < val(v.shift[st.gear].x)) {
            if (st.gear > 0) st.gear--;
        }
    }
    differential_split(torque_t, max_bias, t[1].omega, t[0].omega, t[1].engTor, t[0].engTor, v.driveType,
                       v.whichWheels);
    T bt = v.maxBrakeTorque * braking;
    for (int i = 0; i < 2; ++i)
        t[i].dOmega = (1 / s.tp.jw) * (t[i].engTor + t[i].My - (double)sgn(t[i].omega) * bt - t[i].fx * t[i].rStat);
}

// dof11.cpp:893-905
template <class T>
void veh_rhs(VehState<T>& st, const VehParam<T>& p, const T* fx, const T* fy) {
    T mt = p.m + 2 * (p.muf + p.mur);
    st.vdot = -st.u * st.wz + (fy[0] + fy[1]) / mt;
    st.udot = st.v * st.wz + (fx[0] + fx[1]) / mt;
    st.wzdot = (p.a * fy[0] - p.b * fy[1]) / p.jz;
}

// rhsFun order (dof11_halfImplicit.cpp:560-605) and the step update (:172-235).
template <class T>
void step(Vehicle<T>& s, const T& steering, const T& throttle, const T& braking, double dt) {
    using std::cos;
    using std::sin;
    T loads[2];
    tire_loads(s, loads);
    // vehToTireTransform (dof11.cpp:45-71)
    T delta = d18::steer_angle(s.vp, steering);
    s.tire[0].fz = loads[0];
    s.tire[0].vsy = s.vs.v + s.vs.wz * s.vp.a;
    s.tire[0].vsx = s.vs.u * cos(delta) + s.tire[0].vsy * sin(delta);
    s.tire[1].fz = loads[1];
    s.tire[1].vsy = s.vs.v - s.vs.wz * s.vp.b;
    s.tire[1].vsx = s.vs.u;
    if (!s.nr) {
        d18::tire_rhs_tmeasy(s.tire[0], s.tp, s.vp, steering);
        d18::tire_rhs_tmeasy(s.tire[1], s.tp, s.vp, T(0.0));
    } else {
        d18::tire_rhs_tmeasy_nr(s.tire[0], s.tp, s.vp, steering);
        d18::tire_rhs_tmeasy_nr(s.tire[1], s.tp, s.vp, T(0.0));
    }
    powertrain_rhs(s, throttle, braking);
    // tireToVehTransform (dof11.cpp:100-124)
    {
        T d2 = d18::steer_angle(s.vp, steering);
        TireState<T>& t = s.tire[0];
        T fx = t.fx * cos(d2) - t.fy * sin(d2);
        T fy = t.fx * sin(d2) + t.fy * cos(d2);
        t.fx = fx;
        t.fy = fy;
    }
    T fx[2] = {s.tire[0].fx, s.tire[1].fx};
    T fy[2] = {s.tire[0].fy, s.tire[1].fy};
    veh_rhs(s.vs, s.vp, fx, fy);
    for (int i = 0; i < 2; ++i) {
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
    v.wz = v.wz + v.wzdot * dt;
    v.x = v.x + (v.u * cos(v.psi) - v.v * sin(v.psi)) * dt;
    v.y = v.y + (v.u * sin(v.psi) + v.v * cos(v.psi)) * dt;
    v.psi = v.psi + v.wz * dt;
}

inline int record_len(bool nr) { return nr ? 2 * 11 + 12 : 2 * 15 + 12; }

template <class T>
void to_record(const Vehicle<T>& s, double* o) {
    int k = 0;
    for (int i = 0; i < 2; ++i) {
        const TireState<T>& t = s.tire[i];
        if (!s.nr) {
            o[k++] = val(t.xe); o[k++] = val(t.ye); o[k++] = val(t.xedot); o[k++] = val(t.yedot);
        }
        o[k++] = val(t.omega); o[k++] = val(t.dOmega); o[k++] = val(t.xt); o[k++] = val(t.rStat);
        o[k++] = val(t.fx); o[k++] = val(t.fy); o[k++] = val(t.fz); o[k++] = val(t.vsx); o[k++] = val(t.vsy);
        o[k++] = val(t.My); o[k++] = val(t.engTor);
    }
    const VehState<T>& v = s.vs;
    const T f[] = {v.x, v.y, v.u, v.v, v.psi, v.wz, v.udot, v.vdot, v.wzdot, v.crankOmega, v.dOmega_crank};
    for (const T& x : f) o[k++] = val(x);
    o[k++] = (double)v.gear;
}

template <class T>
void from_record(Vehicle<T>& s, const double* o) {
    int k = 0;
    for (int i = 0; i < 2; ++i) {
        TireState<T>& t = s.tire[i];
        if (!s.nr) {
            t.xe = T(o[k++]); t.ye = T(o[k++]); t.xedot = T(o[k++]); t.yedot = T(o[k++]);
        } else {
            t.xe = t.ye = t.xedot = t.yedot = T(0.0);
        }
        t.omega = T(o[k++]); t.dOmega = T(o[k++]); t.xt = T(o[k++]); t.rStat = T(o[k++]);
        t.fx = T(o[k++]); t.fy = T(o[k++]); t.fz = T(o[k++]); t.vsx = T(o[k++]); t.vsy = T(o[k++]);
        t.My = T(o[k++]); t.engTor = T(o[k++]);
    }
    VehState<T>& v = s.vs;
    T* f[] = {&v.x, &v.y, &v.u, &v.v, &v.psi, &v.wz, &v.udot, &v.vdot, &v.wzdot, &v.crankOmega, &v.dOmega_crank};
    for (T* p : f) *p = T(o[k++]);
    v.gear = (int)o[k++];
}

}  // namespace d11
}  // namespace orc
