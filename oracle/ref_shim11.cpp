// Test-infrastructure shim: exposes the reference 11-DOF CPU half-implicit solver
// (wheeled_vehicle_models/11dof/dof11_halfImplicit.{h,cpp}) through a flat C ABI for golden
// fixtures (same shape as ref_shim18.cpp).  Built only by oracle/Makefile into oracle/_ref/.
//
// 11-DOF record: per tire (front, rear) the 18-DOF tire fields (15 TMeasy / 11 TMeasyNr), then
// x y u v psi wz udot vdot wzdot crankOmega dOmega_crank gear.
#include <cstring>
#include <string>
#include <vector>

#include "dof11_halfImplicit.h"

using namespace d11;

namespace {

struct RefS {
    d11SolverHalfImplicit solver;
    DriverData driver;
    bool nr = false;
};

template <class TS>
void put_tire(const TS& s, double* o, bool nr) {
    int k = 0;
    if (!nr) {
        const auto& q = reinterpret_cast<const TMeasyState&>(s);
        o[k++] = q._xe; o[k++] = q._ye; o[k++] = q._xedot; o[k++] = q._yedot;
    }
    const double v[11] = {s._omega, s._dOmega, s._xt, s._rStat, s._fx, s._fy, s._fz, s._vsx, s._vsy, s._My, s._engTor};
    std::memcpy(o + k, v, sizeof(v));
}
void put_veh(const VehicleState& s, double* o) {
    const double v[12] = {s._x,    s._y,     s._u,          s._v,            s._psi,
                          s._wz,   s._udot,  s._vdot,       s._wzdot,        s._crankOmega,
                          s._dOmega_crank, (double)s._current_gr};
    std::memcpy(o, v, sizeof(v));
}

}  // namespace

extern "C" {

int ref11_record_len(int nr) { return nr ? 2 * 11 + 12 : 2 * 15 + 12; }

void* ref11_create(const char* veh, const char* tire, int nr, const char* driver_file) {
    RefS* r = new RefS();
    r->nr = nr != 0;
    TireType tt = r->nr ? TireType::TMeasyNr : TireType::TMeasy;
    if (driver_file && driver_file[0]) {
        r->solver.Construct(veh, tire, driver_file, tt);
        LoadDriverData(r->driver, driver_file);
    } else {
        r->solver.Construct(veh, tire, tt);
    }
    r->solver.SetTimeStep(1e-3);
    VehicleState vs;
    if (r->nr) {
        TMeasyNrState a, b;
        r->solver.Initialize(vs, a, b);
    } else {
        TMeasyState a, b;
        r->solver.Initialize(vs, a, b);
    }
    return r;
}

void ref11_get_record(void* h, double* o) {
    RefS* r = static_cast<RefS*>(h);
    auto& s = r->solver;
    const int tl = r->nr ? 11 : 15;
    if (r->nr) {
        put_tire(s.m_tireTMNrf_state, o, true);
        put_tire(s.m_tireTMNrr_state, o + tl, true);
    } else {
        put_tire(s.m_tireTMf_state, o, false);
        put_tire(s.m_tireTMr_state, o + tl, false);
    }
    put_veh(s.m_veh_state, o + 2 * tl);
}

int ref11_dump_params(void* h, double* o) {
    RefS* r = static_cast<RefS*>(h);
    const VehicleParam& p = r->solver.m_veh_param;
    int k = 0;
    const double v[] = {p._a, p._b, p._h, p._m, p._jz, p._muf, p._mur, p._maxSteer, p._crankInertia,
                        p._maxBrakeTorque, (double)p._nonLinearSteer, (double)p._tcbool, (double)p._throttleMod,
                        (double)p._driveType, (double)p._whichWheels};
    for (double x : v) o[k++] = x;
    auto dump_map = [&](const std::vector<MapEntry>& m) {
        o[k++] = (double)m.size();
        for (const auto& e : m) { o[k++] = e._x; o[k++] = e._y; }
    };
    o[k++] = (double)p._gearRatios.size();
    for (double g : p._gearRatios) o[k++] = g;
    dump_map(p._shiftMap);
    dump_map(p._steerMap);
    dump_map(p._powertrainMap);
    dump_map(p._lossesMap);
    dump_map(p._CFmap);
    dump_map(p._TRmap);
    if (!r->nr) {
        const auto& t = r->solver.m_tireTM_param;
        const double w[] = {t._jw,     t._rr,      t._mu,      t._r0,       t._pn,     t._pnmax,  t._cx,
                            t._cy,     t._kt,      t._dx,      t._dy,       t._rdyncoPn, t._rdyncoP2n,
                            t._fzRdynco, t._rdyncoCrit, t._dfx0Pn, t._dfx0P2n, t._fxmPn, t._fxmP2n, t._fxsPn,
                            t._fxsP2n, t._sxmPn,   t._sxmP2n,  t._sxsPn,    t._sxsP2n, t._dfy0Pn, t._dfy0P2n,
                            t._fymPn,  t._fymP2n,  t._fysPn,   t._fysP2n,   t._symPn,  t._symP2n, t._sysPn,
                            t._sysP2n};
        for (double x : w) o[k++] = x;
    } else {
        const auto& t = r->solver.m_tireTMNr_param;
        const double w[] = {t._jw,     t._rr,      t._mu,      t._r0,       t._width,  t._rim_radius, t._pn,
                            t._pnmax,  t._kt,      t._rdyncoPn, t._rdyncoP2n, t._fzRdynco, t._rdyncoCrit,
                            t._dfx0Pn, t._dfx0P2n, t._fxmPn,   t._fxmP2n,   t._fxsPn,  t._fxsP2n, t._sxmPn,
                            t._sxmP2n, t._sxsPn,   t._sxsP2n,  t._dfy0Pn,   t._dfy0P2n, t._fymPn, t._fymP2n,
                            t._fysPn,  t._fysP2n,  t._symPn,   t._symP2n,   t._sysPn,  t._sysP2n, t._vcoulomb,
                            t._frblend_begin, t._frblend_end, t._bearingCapacity, t._li, t._p_li, t._p_use};
        for (double x : w) o[k++] = x;
    }
    return k;
}

double ref11_run(void* h, double t0, int nsteps, const double* ctrl, int every, double* out) {
    RefS* r = static_cast<RefS*>(h);
    const int len = ref11_record_len(r->nr);
    double t = t0;
    int rec = 0;
    for (int k = 0; k < nsteps; ++k) {
        double str, thr, brk;
        if (ctrl) {
            str = ctrl[3 * k + 0];
            thr = ctrl[3 * k + 1];
            brk = ctrl[3 * k + 2];
        } else {
            DriverInput di = GetDriverInput(t, r->driver);
            str = di.m_steering;
            thr = di.m_throttle;
            brk = di.m_braking;
        }
        t = r->solver.IntegrateStep(t, thr, str, brk);
        if (out && every > 0 && (k + 1) % every == 0) {
            ref11_get_record(h, out + (size_t)rec * len);
            ++rec;
        }
    }
    return t;
}

void ref11_solve(void* h) { static_cast<RefS*>(h)->solver.Solve(); }
void ref11_destroy(void* h) { delete static_cast<RefS*>(h); }
void ref11_set_step(void* h, double dt) { static_cast<RefS*>(h)->solver.SetTimeStep(dt); }

int ref11_set_param(void* h, const char* name, double value) {
    RefS* r = static_cast<RefS*>(h);
    auto& vp = r->solver.m_veh_param;
    std::string n(name);
    if (n == "m") vp._m = value;
    else if (n == "jz") vp._jz = value;
    
    else if (n == "maxBrakeTorque") vp._maxBrakeTorque = value;
    else if (n == "torque_scale") {
        for (auto& e : vp._powertrainMap) e._y = e._y * value;
    } else if (n == "dfy0Pn") {
        if (r->nr) r->solver.m_tireTMNr_param._dfy0Pn = value; else r->solver.m_tireTM_param._dfy0Pn = value;
    } else if (n == "dfy0P2n") {
        if (r->nr) r->solver.m_tireTMNr_param._dfy0P2n = value; else r->solver.m_tireTM_param._dfy0P2n = value;
    } else return -1;
    return 0;
}

}  // extern "C"
