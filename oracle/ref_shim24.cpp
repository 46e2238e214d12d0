// Test-infrastructure shim: exposes the reference 24-DOF CPU half-implicit solver
// (wheeled_vehicle_models/24dof/dof24_halfImplicit.{h,cpp}) through a flat C ABI for golden
// fixtures (same shape as ref_shim18.cpp).  Built only by oracle/Makefile into oracle/_ref/.
//
// 24-DOF record: per tire LF RF LR RR (TMeasy: xe ye xedot yedot, then) omega dOmega xt dxt rStat
// fx fy fz fxg fyg fzg vgx vgy vgz vsx vsy vsz My engTor; per strut xs xsi dxs ls lsi us vs ws dus
// dvs dws uu vu wu duu dvu dwu fxs fys fzs fzd mx my mz; vehicle x y z u v w psi wz phi wx theta wy
// udot vdot wdot wxdot wydot wzdot dtheta dphi dpsi crankOmega dOmega_crank gear.
#include <cstring>
#include <string>
#include <vector>

#include "dof24_halfImplicit.h"

using namespace d24;

namespace {

struct RefS {
    d24SolverHalfImplicit solver;
    DriverData driver;
    bool nr = false;
};

template <class TS>
int put_tire(const TS& s, double* o, bool nr) {
    int k = 0;
    if (!nr) {
        const auto& q = reinterpret_cast<const TMeasyState&>(s);
        o[k++] = q._xe; o[k++] = q._ye; o[k++] = q._xedot; o[k++] = q._yedot;
    }
    const double v[19] = {s._omega, s._dOmega, s._xt,  s._dxt, s._rStat, s._fx,  s._fy,  s._fz, s._fxg,   s._fyg,
                          s._fzg,   s._vgx,    s._vgy, s._vgz, s._vsx,   s._vsy, s._vsz, s._My, s._engTor};
    std::memcpy(o + k, v, sizeof(v));
    return k + 19;
}
int put_sus(const SuspensionState& q, double* o) {
    const double v[24] = {q._xs, q._xsi, q._dxs, q._ls, q._lsi, q._us, q._vs, q._ws, q._dus, q._dvs, q._dws, q._uu,
                          q._vu, q._wu,  q._duu, q._dvu, q._dwu, q._fxs, q._fys, q._fzs, q._fzd, q._mx, q._my, q._mz};
    std::memcpy(o, v, sizeof(v));
    return 24;
}
int put_veh(const VehicleState& s, double* o) {
    const double v[24] = {s._x,     s._y,     s._z,      s._u,     s._v,     s._w,     s._psi,   s._wz,
                          s._phi,   s._wx,    s._theta,  s._wy,    s._udot,  s._vdot,  s._wdot,  s._wxdot,
                          s._wydot, s._wzdot, s._dtheta, s._dphi,  s._dpsi,  s._crankOmega, s._dOmega_crank,
                          (double)s._current_gr};
    std::memcpy(o, v, sizeof(v));
    return 24;
}

}  // namespace

extern "C" {

int ref24_record_len(int nr) { return 4 * (nr ? 19 : 23) + 4 * 24 + 24; }

void* ref24_create(const char* veh, const char* tire, const char* sus, int nr, const char* driver_file) {
    RefS* r = new RefS();
    r->nr = nr != 0;
    TireType tt = r->nr ? TireType::TMeasyNr : TireType::TMeasy;
    if (driver_file && driver_file[0]) {
        r->solver.Construct(veh, tire, sus, driver_file, tt);
        LoadDriverData(r->driver, driver_file);
    } else {
        r->solver.Construct(veh, tire, sus, tt);
    }
    r->solver.SetTimeStep(1e-3);
    return r;
}

// Initialize with default states plus an initial forward speed (runs initializeTireSus).
void ref24_initialize(void* h, double u0) {
    RefS* r = static_cast<RefS*>(h);
    VehicleState vs;
    vs._u = u0;
    SuspensionState s0, s1, s2, s3;
    if (r->nr) {
        TMeasyNrState a, b, c, d;
        r->solver.Initialize(vs, a, b, c, d, s0, s1, s2, s3);
    } else {
        TMeasyState a, b, c, d;
        r->solver.Initialize(vs, a, b, c, d, s0, s1, s2, s3);
    }
}

void ref24_get_record(void* h, double* o) {
    RefS* r = static_cast<RefS*>(h);
    auto& s = r->solver;
    int k = 0;
    if (r->nr) {
        k += put_tire(s.m_tireTMNrlf_state, o + k, true);
        k += put_tire(s.m_tireTMNrrf_state, o + k, true);
        k += put_tire(s.m_tireTMNrlr_state, o + k, true);
        k += put_tire(s.m_tireTMNrrr_state, o + k, true);
    } else {
        k += put_tire(s.m_tireTMlf_state, o + k, false);
        k += put_tire(s.m_tireTMrf_state, o + k, false);
        k += put_tire(s.m_tireTMlr_state, o + k, false);
        k += put_tire(s.m_tireTMrr_state, o + k, false);
    }
    k += put_sus(s.m_suslf_state, o + k);
    k += put_sus(s.m_susrf_state, o + k);
    k += put_sus(s.m_suslr_state, o + k);
    k += put_sus(s.m_susrr_state, o + k);
    put_veh(s.m_veh_state, o + k);
}

int ref24_dump_params(void* h, double* o) {
    RefS* r = static_cast<RefS*>(h);
    const VehicleParam& p = r->solver.m_veh_param;
    int k = 0;
    const double v[] = {p._a,   p._b,   p._h,    p._m,    p._jz,       p._jx,           p._jy,
                        p._cf,  p._cr,  p._muf,  p._mur,  p._hrcf,     p._hrcr,         p._maxSteer,
                        p._crankInertia, p._maxBrakeTorque, (double)p._nonLinearSteer, (double)p._tcbool,
                        (double)p._throttleMod, (double)p._driveType, (double)p._whichWheels};
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
    o[k++] = r->solver.m_sus_param._ks;
    o[k++] = r->solver.m_sus_param._bs;
    return k;
}

double ref24_run(void* h, double t0, int nsteps, const double* ctrl, int every, double* out) {
    RefS* r = static_cast<RefS*>(h);
    const int len = ref24_record_len(r->nr);
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
            ref24_get_record(h, out + (size_t)rec * len);
            ++rec;
        }
    }
    return t;
}

void ref24_solve(void* h) { static_cast<RefS*>(h)->solver.Solve(); }
void ref24_destroy(void* h) { delete static_cast<RefS*>(h); }
void ref24_set_step(void* h, double dt) { static_cast<RefS*>(h)->solver.SetTimeStep(dt); }

int ref24_set_param(void* h, const char* name, double value) {
    RefS* r = static_cast<RefS*>(h);
    auto& vp = r->solver.m_veh_param;
    std::string n(name);
    if (n == "m") vp._m = value;
    else if (n == "jz") vp._jz = value;
    else if (n == "jx") vp._jx = value;
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
