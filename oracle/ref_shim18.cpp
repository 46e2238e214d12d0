// Test-infrastructure shim: exposes the reference 18-DOF CPU half-implicit solver
// (wheeled_vehicle_models/18dof/dof18_halfImplicit.{h,cpp}) through a flat C ABI so that
// golden fixtures can be written from the reference itself. Built only by oracle/Makefile
// into oracle/_ref/; never linked into the product library.
//
// Canonical 18-DOF state record (shared by the oracle, the fixtures and the HIP solver):
//   per tire (LF, RF, LR, RR), TMeasy : xe ye xedot yedot omega dOmega xt rStat fx fy fz vsx vsy My engTor (15)
//   per tire, TMeasyNr                : omega dOmega xt rStat fx fy fz vsx vsy My engTor              (11)
//   vehicle                           : x y u v psi wz phi wx udot vdot wxdot wzdot crankOmega dOmega_crank gear (15)
#include <cstring>
#include <string>
#include <vector>

#include "dof18_halfImplicit.h"

using namespace d18;

namespace {

struct Ref18 {
    d18SolverHalfImplicit solver;
    DriverData driver;
    bool nr = false;
};

void put_tire(const TMeasyState& s, double* o) {
    const double v[15] = {s._xe, s._ye, s._xedot, s._yedot, s._omega, s._dOmega, s._xt, s._rStat,
                          s._fx, s._fy, s._fz,   s._vsx,   s._vsy,   s._My,    s._engTor};
    std::memcpy(o, v, sizeof(v));
}
void put_tire(const TMeasyNrState& s, double* o) {
    const double v[11] = {s._omega, s._dOmega, s._xt, s._rStat, s._fx, s._fy, s._fz, s._vsx, s._vsy, s._My, s._engTor};
    std::memcpy(o, v, sizeof(v));
}
void put_veh(const VehicleState& s, double* o) {
    const double v[15] = {s._x,    s._y,    s._u,     s._v,     s._psi,       s._wz,          s._phi,
                          s._wx,   s._udot, s._vdot,  s._wxdot, s._wzdot,     s._crankOmega,  s._dOmega_crank,
                          (double)s._current_gr};
    std::memcpy(o, v, sizeof(v));
}
void get_tire(TMeasyState& s, const double* v) {
    s._xe = v[0]; s._ye = v[1]; s._xedot = v[2]; s._yedot = v[3]; s._omega = v[4]; s._dOmega = v[5];
    s._xt = v[6]; s._rStat = v[7]; s._fx = v[8]; s._fy = v[9]; s._fz = v[10]; s._vsx = v[11]; s._vsy = v[12];
    s._My = v[13]; s._engTor = v[14];
}
void get_tire(TMeasyNrState& s, const double* v) {
    s._omega = v[0]; s._dOmega = v[1]; s._xt = v[2]; s._rStat = v[3]; s._fx = v[4]; s._fy = v[5]; s._fz = v[6];
    s._vsx = v[7]; s._vsy = v[8]; s._My = v[9]; s._engTor = v[10];
}
void get_veh(VehicleState& s, const double* v) {
    s._x = v[0]; s._y = v[1]; s._u = v[2]; s._v = v[3]; s._psi = v[4]; s._wz = v[5]; s._phi = v[6]; s._wx = v[7];
    s._udot = v[8]; s._vdot = v[9]; s._wxdot = v[10]; s._wzdot = v[11]; s._crankOmega = v[12];
    s._dOmega_crank = v[13]; s._current_gr = (int)v[14];
}

}  // namespace

extern "C" {

int ref18_record_len(int nr) { return nr ? 4 * 11 + 15 : 4 * 15 + 15; }

void* ref18_create(const char* veh, const char* tire, int nr, const char* driver_file) {
    Ref18* r = new Ref18();
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
        TMeasyNrState a, b, c, d;
        r->solver.Initialize(vs, a, b, c, d);
    } else {
        TMeasyState a, b, c, d;
        r->solver.Initialize(vs, a, b, c, d);
    }
    return r;
}

void ref18_destroy(void* h) { delete static_cast<Ref18*>(h); }

void ref18_set_step(void* h, double dt) { static_cast<Ref18*>(h)->solver.SetTimeStep(dt); }

// Overrides one of the per-vehicle jitter parameters after Construct (tireInit has already run,
// and none of these fields enter it).
int ref18_set_param(void* h, const char* name, double value) {
    Ref18* r = static_cast<Ref18*>(h);
    VehicleParam& vp = r->solver.m_veh_param;
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

// Writes the parsed parameters in the order of the oracle's parameter dump; returns the count.
int ref18_dump_params(void* h, double* o) {
    Ref18* r = static_cast<Ref18*>(h);
    const VehicleParam& p = r->solver.m_veh_param;
    int k = 0;
    const double v[] = {p._a,    p._b,    p._h,    p._m,     p._jz,   p._jx,   p._jxz,  p._cf,
                        p._cr,   p._muf,  p._mur,  p._hrcf,  p._hrcr, p._krof, p._kror, p._brof,
                        p._bror, p._maxSteer, p._crankInertia, p._maxBrakeTorque,
                        (double)p._nonLinearSteer, (double)p._tcbool, (double)p._throttleMod,
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
        const TMeasyParam& t = r->solver.m_tireTM_param;
        const double w[] = {t._jw,     t._rr,      t._mu,      t._r0,       t._pn,     t._pnmax,  t._cx,
                            t._cy,     t._kt,      t._dx,      t._dy,       t._rdyncoPn, t._rdyncoP2n,
                            t._fzRdynco, t._rdyncoCrit, t._dfx0Pn, t._dfx0P2n, t._fxmPn, t._fxmP2n, t._fxsPn,
                            t._fxsP2n, t._sxmPn,   t._sxmP2n,  t._sxsPn,    t._sxsP2n, t._dfy0Pn, t._dfy0P2n,
                            t._fymPn,  t._fymP2n,  t._fysPn,   t._fysP2n,   t._symPn,  t._symP2n, t._sysPn,
                            t._sysP2n};
        for (double x : w) o[k++] = x;
    } else {
        const TMeasyNrParam& t = r->solver.m_tireTMNr_param;
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

void ref18_get_record(void* h, double* o) {
    Ref18* r = static_cast<Ref18*>(h);
    auto& s = r->solver;
    if (r->nr) {
        put_tire(s.m_tireTMNrlf_state, o + 0);
        put_tire(s.m_tireTMNrrf_state, o + 11);
        put_tire(s.m_tireTMNrlr_state, o + 22);
        put_tire(s.m_tireTMNrrr_state, o + 33);
        put_veh(s.m_veh_state, o + 44);
    } else {
        put_tire(s.m_tireTMlf_state, o + 0);
        put_tire(s.m_tireTMrf_state, o + 15);
        put_tire(s.m_tireTMlr_state, o + 30);
        put_tire(s.m_tireTMrr_state, o + 45);
        put_veh(s.m_veh_state, o + 60);
    }
}

void ref18_set_record(void* h, const double* v) {
    Ref18* r = static_cast<Ref18*>(h);
    auto& s = r->solver;
    if (r->nr) {
        get_tire(s.m_tireTMNrlf_state, v + 0);
        get_tire(s.m_tireTMNrrf_state, v + 11);
        get_tire(s.m_tireTMNrlr_state, v + 22);
        get_tire(s.m_tireTMNrrr_state, v + 33);
        get_veh(s.m_veh_state, v + 44);
    } else {
        get_tire(s.m_tireTMlf_state, v + 0);
        get_tire(s.m_tireTMrf_state, v + 15);
        get_tire(s.m_tireTMlr_state, v + 30);
        get_tire(s.m_tireTMrr_state, v + 45);
        get_veh(s.m_veh_state, v + 60);
    }
}

// Advances nsteps steps with IntegrateStep. Controls come from `ctrl` ([nsteps][3] = steering, throttle,
// braking) when non-null, else from the driver file through the reference GetDriverInput at the
// step's start time (the same lookup Integrate() performs). Time accumulates as t += dt from t0.
// Every `every` steps the full record is appended to `out`. Returns the final time.
double ref18_run(void* h, double t0, int nsteps, const double* ctrl, int every, double* out) {
    Ref18* r = static_cast<Ref18*>(h);
    const int len = ref18_record_len(r->nr);
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
            ref18_get_record(h, out + (size_t)rec * len);
            ++rec;
        }
    }
    return t;
}

// Runs the reference's own Solve() (driver file, end time from the last driver row).
void ref18_solve(void* h) { static_cast<Ref18*>(h)->solver.Solve(); }

// One IntegrateStepWithJacobian(t, throttle, steering, braking, true) from the current state: writes the
// reference's central-difference Jacobians of the right-hand side, state [n][n] and controls
// [n][2] (steering, throttle), row-major in its packY order, then advances one step as the reference
// does.  Returns n (21/20 for TMeasy with/without converter, 13/12 for TMeasyNr).
int ref18_step_jacobian(void* h, double t, double throttle, double steering, double braking, double* js,
                        double* jc) {
    Ref18* r = static_cast<Ref18*>(h);
    r->solver.IntegrateStepWithJacobian(t, throttle, steering, braking, true);
    const auto J = r->solver.GetJacobianState();
    const auto C = r->solver.GetJacobianControls();
    const int n = (int)J.size();
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) js[i * n + j] = J[i][j];
        jc[2 * i] = C[i][0];
        jc[2 * i + 1] = C[i][1];
    }
    return n;
}

}  // extern "C"
