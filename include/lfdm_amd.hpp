// C++ drop-in for the reference's GPU solver classes, on top of the C ABI in lfdm_amd.h.
//
// d11SolverHalfImplicitGPU, d18SolverHalfImplicitGPU and d24SolverHalfImplicitGPU keep the
// reference's class names, member names, argument order and the state structs users touch
// (wheeled_vehicle_models/{11,18,24}dof-gpu/dof*_halfImplicit_gpu.cuh, dof*_gpu.cuh), so code written
// against the reference compiles after replacing its includes with this header and linking
// liblfdm_amd.so.  Plain C++17; no HIP compiler is needed (the functor form of SolveStep is in
// lfdm_amd_functor.hpp, which does need one).
//
// Differences from the reference, on purpose: failures throw lfdm_amd::Error (the reference
// asserts, prints or exits); the solver does not call cudaSetDevice(0); the time base of Solve and
// advance is carried continuously (see lfdm_amd.h).
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "lfdm_amd.h"

namespace lfdm_amd {

// TireType (dof18_gpu.cuh, enum TireType)
enum class TireType : int { TMeasy = LFDM_TIRE_TMEASY, TMeasyNr = LFDM_TIRE_TMEASYNR };

// DriverInput (utils_gpu.cuh:29-40)
struct DriverInput {
    double m_time = 0., m_steering = 0., m_throttle = 0., m_braking = 0.;
};

// A failed C-ABI call: status code and message.
class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc) {
    if (rc < 0) throw Error(rc, lfdm_last_error());
}

// MapEntry (utils_gpu.cuh): one point of a piecewise-linear map.
struct MapEntry {
    double _x = 0., _y = 0.;
};

namespace detail {

// Readers and writers of the flat parameter vector (the order of lfdm_params_from_json): a parameter
// struct lists its fields once, in that order, in visit(), which both directions walk.
struct FlatIn {
    const std::vector<double>& v;
    size_t i = 0;
    double next() {
        if (i >= v.size()) throw Error(LFDM_E_PARSE, "parameter vector too short for the parameter structs");
        return v[i++];
    }
    void operator()(double& x) { x = next(); }
    void operator()(bool& b) { b = next() != 0.0; }
    void operator()(std::vector<double>& g) {
        g.resize((size_t)next());
        for (double& x : g) x = next();
    }
    void operator()(std::vector<MapEntry>& m) {
        m.resize((size_t)next());
        for (MapEntry& e : m) {
            e._x = next();
            e._y = next();
        }
    }
};

struct FlatOut {
    std::vector<double> v;
    void operator()(const double& x) { v.push_back(x); }
    void operator()(const bool& b) { v.push_back(b ? 1.0 : 0.0); }
    void operator()(const std::vector<double>& g) {
        v.push_back((double)g.size());
        v.insert(v.end(), g.begin(), g.end());
    }
    void operator()(const std::vector<MapEntry>& m) {
        v.push_back((double)m.size());
        for (const MapEntry& e : m) {
            v.push_back(e._x);
            v.push_back(e._y);
        }
    }
};

// The fields every model's VehicleParam has after its scalars (dof18_gpu.cuh:601-708): the steering
// mode and the powertrain switches, gear ratios and the six maps (shift: x downshift, y upshift speed).
struct PowertrainParam {
    bool _nonLinearSteer = false, _tcbool = false, _throttleMod = false, _driveType = false, _whichWheels = false;
    std::vector<double> _gearRatios;
    std::vector<MapEntry> _shiftMap, _steerMap, _powertrainMap, _lossesMap, _CFmap, _TRmap;
    template <class V>
    void visit_powertrain(V& f) {
        f(_nonLinearSteer); f(_tcbool); f(_throttleMod); f(_driveType); f(_whichWheels); f(_gearRatios);
        f(_shiftMap); f(_steerMap); f(_powertrainMap); f(_lossesMap); f(_CFmap); f(_TRmap);
    }
};

}  // namespace detail

// TMeasyParam (dof18_gpu.cuh:25-148; the same fields in every model), after tireInit.
struct TMeasyParam {
    double _jw = 0., _rr = 0., _mu = 0., _r0 = 0., _pn = 0., _pnmax = 0., _cx = 0., _cy = 0., _kt = 0., _dx = 0.,
           _dy = 0., _rdyncoPn = 0., _rdyncoP2n = 0., _fzRdynco = 0., _rdyncoCrit = 0., _dfx0Pn = 0., _dfx0P2n = 0.,
           _fxmPn = 0., _fxmP2n = 0., _fxsPn = 0., _fxsP2n = 0., _sxmPn = 0., _sxmP2n = 0., _sxsPn = 0., _sxsP2n = 0.,
           _dfy0Pn = 0., _dfy0P2n = 0., _fymPn = 0., _fymP2n = 0., _fysPn = 0., _fysP2n = 0., _symPn = 0.,
           _symP2n = 0., _sysPn = 0., _sysP2n = 0.;
    template <class V>
    void visit(V& f) {
        for (double* p : {&_jw, &_rr, &_mu, &_r0, &_pn, &_pnmax, &_cx, &_cy, &_kt, &_dx, &_dy, &_rdyncoPn,
                          &_rdyncoP2n, &_fzRdynco, &_rdyncoCrit, &_dfx0Pn, &_dfx0P2n, &_fxmPn, &_fxmP2n, &_fxsPn,
                          &_fxsP2n, &_sxmPn, &_sxmP2n, &_sxsPn, &_sxsP2n, &_dfy0Pn, &_dfy0P2n, &_fymPn, &_fymP2n,
                          &_fysPn, &_fysP2n, &_symPn, &_symP2n, &_sysPn, &_sysP2n})
            f(*p);
    }
};

// TMeasyNrParam (dof18_gpu.cuh:255-399), after tireInit.
struct TMeasyNrParam {
    double _jw = 0., _rr = 0., _mu = 0., _r0 = 0., _width = 0., _rim_radius = 0., _pn = 0., _pnmax = 0., _kt = 0.,
           _rdyncoPn = 0., _rdyncoP2n = 0., _fzRdynco = 0., _rdyncoCrit = 0., _dfx0Pn = 0., _dfx0P2n = 0.,
           _fxmPn = 0., _fxmP2n = 0., _fxsPn = 0., _fxsP2n = 0., _sxmPn = 0., _sxmP2n = 0., _sxsPn = 0., _sxsP2n = 0.,
           _dfy0Pn = 0., _dfy0P2n = 0., _fymPn = 0., _fymP2n = 0., _fysPn = 0., _fysP2n = 0., _symPn = 0.,
           _symP2n = 0., _sysPn = 0., _sysP2n = 0., _vcoulomb = 0., _frblend_begin = 0., _frblend_end = 0.,
           _bearingCapacity = 0., _li = 0., _p_li = 0., _p_use = 0.;
    template <class V>
    void visit(V& f) {
        for (double* p : {&_jw, &_rr, &_mu, &_r0, &_width, &_rim_radius, &_pn, &_pnmax, &_kt, &_rdyncoPn,
                          &_rdyncoP2n, &_fzRdynco, &_rdyncoCrit, &_dfx0Pn, &_dfx0P2n, &_fxmPn, &_fxmP2n, &_fxsPn,
                          &_fxsP2n, &_sxmPn, &_sxmP2n, &_sxsPn, &_sxsP2n, &_dfy0Pn, &_dfy0P2n, &_fymPn, &_fymP2n,
                          &_fysPn, &_fysP2n, &_symPn, &_symP2n, &_sysPn, &_sysP2n, &_vcoulomb, &_frblend_begin,
                          &_frblend_end, &_bearingCapacity, &_li, &_p_li, &_p_use})
            f(*p);
    }
};


namespace d18GPU {


// TMeasyState (dof18_gpu.cuh:149-253)
struct TMeasyState {
    double _xe = 0.;
    double _ye = 0.;
    double _xedot = 0.;
    double _yedot = 0.;
    double _omega = 0.;
    double _dOmega = 0.;
    double _xt = 0.;
    double _rStat = 0.;
    double _fx = 0.;
    double _fy = 0.;
    double _fz = 0.;
    double _vsx = 0.;
    double _vsy = 0.;
    double _My = 0.;
    double _engTor = 0.;
    static constexpr int kLen = 15;
    void to_record(double* r) const {
        r[0] = _xe;
        r[1] = _ye;
        r[2] = _xedot;
        r[3] = _yedot;
        r[4] = _omega;
        r[5] = _dOmega;
        r[6] = _xt;
        r[7] = _rStat;
        r[8] = _fx;
        r[9] = _fy;
        r[10] = _fz;
        r[11] = _vsx;
        r[12] = _vsy;
        r[13] = _My;
        r[14] = _engTor;
    }
    void from_record(const double* r) {
        _xe = r[0];
        _ye = r[1];
        _xedot = r[2];
        _yedot = r[3];
        _omega = r[4];
        _dOmega = r[5];
        _xt = r[6];
        _rStat = r[7];
        _fx = r[8];
        _fy = r[9];
        _fz = r[10];
        _vsx = r[11];
        _vsy = r[12];
        _My = r[13];
        _engTor = r[14];
    }
};

// TMeasyNrState (dof18_gpu.cuh:400-500)
struct TMeasyNrState {
    double _omega = 0.;
    double _dOmega = 0.;
    double _xt = 0.;
    double _rStat = 0.;
    double _fx = 0.;
    double _fy = 0.;
    double _fz = 0.;
    double _vsx = 0.;
    double _vsy = 0.;
    double _My = 0.;
    double _engTor = 0.;
    static constexpr int kLen = 11;
    void to_record(double* r) const {
        r[0] = _omega;
        r[1] = _dOmega;
        r[2] = _xt;
        r[3] = _rStat;
        r[4] = _fx;
        r[5] = _fy;
        r[6] = _fz;
        r[7] = _vsx;
        r[8] = _vsy;
        r[9] = _My;
        r[10] = _engTor;
    }
    void from_record(const double* r) {
        _omega = r[0];
        _dOmega = r[1];
        _xt = r[2];
        _rStat = r[3];
        _fx = r[4];
        _fy = r[5];
        _fz = r[6];
        _vsx = r[7];
        _vsy = r[8];
        _My = r[9];
        _engTor = r[10];
    }
};

// VehicleState (dof18_gpu.cuh:719-810)
struct VehicleState {
    double _x = 0.;
    double _y = 0.;
    double _u = 0.;
    double _v = 0.;
    double _psi = 0.;
    double _wz = 0.;
    double _phi = 0.;
    double _wx = 0.;
    double _udot = 0.;
    double _vdot = 0.;
    double _wxdot = 0.;
    double _wzdot = 0.;
    double _crankOmega = 0.;
    double _dOmega_crank = 0.;
    int _current_gr = 0;
    static constexpr int kLen = 15;
    void to_record(double* r) const {
        r[0] = _x;
        r[1] = _y;
        r[2] = _u;
        r[3] = _v;
        r[4] = _psi;
        r[5] = _wz;
        r[6] = _phi;
        r[7] = _wx;
        r[8] = _udot;
        r[9] = _vdot;
        r[10] = _wxdot;
        r[11] = _wzdot;
        r[12] = _crankOmega;
        r[13] = _dOmega_crank;
        r[14] = (double)_current_gr;
    }
    void from_record(const double* r) {
        _x = r[0];
        _y = r[1];
        _u = r[2];
        _v = r[3];
        _psi = r[4];
        _wz = r[5];
        _phi = r[6];
        _wx = r[7];
        _udot = r[8];
        _vdot = r[9];
        _wxdot = r[10];
        _wzdot = r[11];
        _crankOmega = r[12];
        _dOmega_crank = r[13];
        _current_gr = (int)r[14];
    }
};

// SimState / SimStateNr (dof18_gpu.cuh:864-905)

struct SimState {
    VehicleState _veh_state;
    TMeasyState _tirelf_state;
    TMeasyState _tirerf_state;
    TMeasyState _tirelr_state;
    TMeasyState _tirerr_state;
    void from_record(const double* r) {
        int k = 0;
        _tirelf_state.from_record(r + k);
        k += TMeasyState::kLen;
        _tirerf_state.from_record(r + k);
        k += TMeasyState::kLen;
        _tirelr_state.from_record(r + k);
        k += TMeasyState::kLen;
        _tirerr_state.from_record(r + k);
        k += TMeasyState::kLen;
        _veh_state.from_record(r + k);
    }
};

struct SimStateNr {
    VehicleState _veh_state;
    TMeasyNrState _tirelf_state;
    TMeasyNrState _tirerf_state;
    TMeasyNrState _tirelr_state;
    TMeasyNrState _tirerr_state;
    void from_record(const double* r) {
        int k = 0;
        _tirelf_state.from_record(r + k);
        k += TMeasyNrState::kLen;
        _tirerf_state.from_record(r + k);
        k += TMeasyNrState::kLen;
        _tirelr_state.from_record(r + k);
        k += TMeasyNrState::kLen;
        _tirerr_state.from_record(r + k);
        k += TMeasyNrState::kLen;
        _veh_state.from_record(r + k);
    }
};

// VehicleParam (dof18_gpu.cuh:502-810); maps and gear ratios as vectors instead of device pointers and sizes.
struct VehicleParam : detail::PowertrainParam {
    double _a = 0., _b = 0., _h = 0., _m = 0., _jz = 0., _jx = 0., _jxz = 0., _cf = 0., _cr = 0., _muf = 0.,
           _mur = 0., _hrcf = 0., _hrcr = 0., _krof = 0., _kror = 0., _brof = 0., _bror = 0., _maxSteer =
           0., _crankInertia = 0., _maxBrakeTorque = 0.;
    template <class V>
    void visit(V& f) {
        for (double* p : {&_a, &_b, &_h, &_m, &_jz, &_jx, &_jxz, &_cf, &_cr, &_muf, &_mur, &_hrcf,
                          &_hrcr, &_krof, &_kror, &_brof, &_bror, &_maxSteer, &_crankInertia,
                          &_maxBrakeTorque})
            f(*p);
        visit_powertrain(f);
    }
};
using lfdm_amd::TMeasyNrParam;
using lfdm_amd::TMeasyParam;

}  // namespace d18GPU

namespace d11GPU {


using d18GPU::TMeasyState;
using d18GPU::TMeasyNrState;

// VehicleState (dof11_gpu.cuh:686-765)
struct VehicleState {
    double _x = 0.;
    double _y = 0.;
    double _u = 0.;
    double _v = 0.;
    double _psi = 0.;
    double _wz = 0.;
    double _udot = 0.;
    double _vdot = 0.;
    double _wzdot = 0.;
    double _crankOmega = 0.;
    double _dOmega_crank = 0.;
    int _current_gr = 0;
    static constexpr int kLen = 12;
    void to_record(double* r) const {
        r[0] = _x;
        r[1] = _y;
        r[2] = _u;
        r[3] = _v;
        r[4] = _psi;
        r[5] = _wz;
        r[6] = _udot;
        r[7] = _vdot;
        r[8] = _wzdot;
        r[9] = _crankOmega;
        r[10] = _dOmega_crank;
        r[11] = (double)_current_gr;
    }
    void from_record(const double* r) {
        _x = r[0];
        _y = r[1];
        _u = r[2];
        _v = r[3];
        _psi = r[4];
        _wz = r[5];
        _udot = r[6];
        _vdot = r[7];
        _wzdot = r[8];
        _crankOmega = r[9];
        _dOmega_crank = r[10];
        _current_gr = (int)r[11];
    }
};

// SimState / SimStateNr (dof11_gpu.cuh:815-842)

struct SimState {
    VehicleState _veh_state;
    TMeasyState _tiref_state;
    TMeasyState _tirer_state;
    void from_record(const double* r) {
        int k = 0;
        _tiref_state.from_record(r + k);
        k += TMeasyState::kLen;
        _tirer_state.from_record(r + k);
        k += TMeasyState::kLen;
        _veh_state.from_record(r + k);
    }
};

struct SimStateNr {
    VehicleState _veh_state;
    TMeasyNrState _tiref_state;
    TMeasyNrState _tirer_state;
    void from_record(const double* r) {
        int k = 0;
        _tiref_state.from_record(r + k);
        k += TMeasyNrState::kLen;
        _tirer_state.from_record(r + k);
        k += TMeasyNrState::kLen;
        _veh_state.from_record(r + k);
    }
};

// VehicleParam (dof11_gpu.cuh:496-620); maps and gear ratios as vectors instead of device pointers and sizes.
struct VehicleParam : detail::PowertrainParam {
    double _a = 0., _b = 0., _h = 0., _m = 0., _jz = 0., _muf = 0., _mur = 0., _maxSteer = 0., _crankInertia
           = 0., _maxBrakeTorque = 0.;
    template <class V>
    void visit(V& f) {
        for (double* p : {&_a, &_b, &_h, &_m, &_jz, &_muf, &_mur, &_maxSteer, &_crankInertia,
                          &_maxBrakeTorque})
            f(*p);
        visit_powertrain(f);
    }
};
using lfdm_amd::TMeasyNrParam;
using lfdm_amd::TMeasyParam;

}  // namespace d11GPU

namespace d24GPU {


// TMeasyState (dof24_gpu.cuh)
struct TMeasyState {
    double _xe = 0.;
    double _ye = 0.;
    double _xedot = 0.;
    double _yedot = 0.;
    double _omega = 0.;
    double _dOmega = 0.;
    double _xt = 0.;
    double _dxt = 0.;
    double _rStat = 0.;
    double _fx = 0.;
    double _fy = 0.;
    double _fz = 0.;
    double _fxg = 0.;
    double _fyg = 0.;
    double _fzg = 0.;
    double _vgx = 0.;
    double _vgy = 0.;
    double _vgz = 0.;
    double _vsx = 0.;
    double _vsy = 0.;
    double _vsz = 0.;
    double _My = 0.;
    double _engTor = 0.;
    static constexpr int kLen = 23;
    void to_record(double* r) const {
        r[0] = _xe;
        r[1] = _ye;
        r[2] = _xedot;
        r[3] = _yedot;
        r[4] = _omega;
        r[5] = _dOmega;
        r[6] = _xt;
        r[7] = _dxt;
        r[8] = _rStat;
        r[9] = _fx;
        r[10] = _fy;
        r[11] = _fz;
        r[12] = _fxg;
        r[13] = _fyg;
        r[14] = _fzg;
        r[15] = _vgx;
        r[16] = _vgy;
        r[17] = _vgz;
        r[18] = _vsx;
        r[19] = _vsy;
        r[20] = _vsz;
        r[21] = _My;
        r[22] = _engTor;
    }
    void from_record(const double* r) {
        _xe = r[0];
        _ye = r[1];
        _xedot = r[2];
        _yedot = r[3];
        _omega = r[4];
        _dOmega = r[5];
        _xt = r[6];
        _dxt = r[7];
        _rStat = r[8];
        _fx = r[9];
        _fy = r[10];
        _fz = r[11];
        _fxg = r[12];
        _fyg = r[13];
        _fzg = r[14];
        _vgx = r[15];
        _vgy = r[16];
        _vgz = r[17];
        _vsx = r[18];
        _vsy = r[19];
        _vsz = r[20];
        _My = r[21];
        _engTor = r[22];
    }
};

// TMeasyNrState (dof24_gpu.cuh)
struct TMeasyNrState {
    double _omega = 0.;
    double _dOmega = 0.;
    double _xt = 0.;
    double _dxt = 0.;
    double _rStat = 0.;
    double _fx = 0.;
    double _fy = 0.;
    double _fz = 0.;
    double _fxg = 0.;
    double _fyg = 0.;
    double _fzg = 0.;
    double _vgx = 0.;
    double _vgy = 0.;
    double _vgz = 0.;
    double _vsx = 0.;
    double _vsy = 0.;
    double _vsz = 0.;
    double _My = 0.;
    double _engTor = 0.;
    static constexpr int kLen = 19;
    void to_record(double* r) const {
        r[0] = _omega;
        r[1] = _dOmega;
        r[2] = _xt;
        r[3] = _dxt;
        r[4] = _rStat;
        r[5] = _fx;
        r[6] = _fy;
        r[7] = _fz;
        r[8] = _fxg;
        r[9] = _fyg;
        r[10] = _fzg;
        r[11] = _vgx;
        r[12] = _vgy;
        r[13] = _vgz;
        r[14] = _vsx;
        r[15] = _vsy;
        r[16] = _vsz;
        r[17] = _My;
        r[18] = _engTor;
    }
    void from_record(const double* r) {
        _omega = r[0];
        _dOmega = r[1];
        _xt = r[2];
        _dxt = r[3];
        _rStat = r[4];
        _fx = r[5];
        _fy = r[6];
        _fz = r[7];
        _fxg = r[8];
        _fyg = r[9];
        _fzg = r[10];
        _vgx = r[11];
        _vgy = r[12];
        _vgz = r[13];
        _vsx = r[14];
        _vsy = r[15];
        _vsz = r[16];
        _My = r[17];
        _engTor = r[18];
    }
};

// SuspensionState (dof24_gpu.cuh)
struct SuspensionState {
    double _xs = 0.;
    double _xsi = 0.;
    double _dxs = 0.;
    double _ls = 0.;
    double _lsi = 0.;
    double _us = 0.;
    double _vs = 0.;
    double _ws = 0.;
    double _dus = 0.;
    double _dvs = 0.;
    double _dws = 0.;
    double _uu = 0.;
    double _vu = 0.;
    double _wu = 0.;
    double _duu = 0.;
    double _dvu = 0.;
    double _dwu = 0.;
    double _fxs = 0.;
    double _fys = 0.;
    double _fzs = 0.;
    double _fzd = 0.;
    double _mx = 0.;
    double _my = 0.;
    double _mz = 0.;
    static constexpr int kLen = 24;
    void to_record(double* r) const {
        r[0] = _xs;
        r[1] = _xsi;
        r[2] = _dxs;
        r[3] = _ls;
        r[4] = _lsi;
        r[5] = _us;
        r[6] = _vs;
        r[7] = _ws;
        r[8] = _dus;
        r[9] = _dvs;
        r[10] = _dws;
        r[11] = _uu;
        r[12] = _vu;
        r[13] = _wu;
        r[14] = _duu;
        r[15] = _dvu;
        r[16] = _dwu;
        r[17] = _fxs;
        r[18] = _fys;
        r[19] = _fzs;
        r[20] = _fzd;
        r[21] = _mx;
        r[22] = _my;
        r[23] = _mz;
    }
    void from_record(const double* r) {
        _xs = r[0];
        _xsi = r[1];
        _dxs = r[2];
        _ls = r[3];
        _lsi = r[4];
        _us = r[5];
        _vs = r[6];
        _ws = r[7];
        _dus = r[8];
        _dvs = r[9];
        _dws = r[10];
        _uu = r[11];
        _vu = r[12];
        _wu = r[13];
        _duu = r[14];
        _dvu = r[15];
        _dwu = r[16];
        _fxs = r[17];
        _fys = r[18];
        _fzs = r[19];
        _fzd = r[20];
        _mx = r[21];
        _my = r[22];
        _mz = r[23];
    }
};

// VehicleState (dof24_gpu.cuh)
struct VehicleState {
    double _x = 0.;
    double _y = 0.;
    double _z = 0.;
    double _u = 0.;
    double _v = 0.;
    double _w = 0.;
    double _psi = 0.;
    double _wz = 0.;
    double _phi = 0.;
    double _wx = 0.;
    double _theta = 0.;
    double _wy = 0.;
    double _udot = 0.;
    double _vdot = 0.;
    double _wdot = 0.;
    double _wxdot = 0.;
    double _wydot = 0.;
    double _wzdot = 0.;
    double _dtheta = 0.;
    double _dphi = 0.;
    double _dpsi = 0.;
    double _crankOmega = 0.;
    double _dOmega_crank = 0.;
    int _current_gr = 0;
    static constexpr int kLen = 24;
    void to_record(double* r) const {
        r[0] = _x;
        r[1] = _y;
        r[2] = _z;
        r[3] = _u;
        r[4] = _v;
        r[5] = _w;
        r[6] = _psi;
        r[7] = _wz;
        r[8] = _phi;
        r[9] = _wx;
        r[10] = _theta;
        r[11] = _wy;
        r[12] = _udot;
        r[13] = _vdot;
        r[14] = _wdot;
        r[15] = _wxdot;
        r[16] = _wydot;
        r[17] = _wzdot;
        r[18] = _dtheta;
        r[19] = _dphi;
        r[20] = _dpsi;
        r[21] = _crankOmega;
        r[22] = _dOmega_crank;
        r[23] = (double)_current_gr;
    }
    void from_record(const double* r) {
        _x = r[0];
        _y = r[1];
        _z = r[2];
        _u = r[3];
        _v = r[4];
        _w = r[5];
        _psi = r[6];
        _wz = r[7];
        _phi = r[8];
        _wx = r[9];
        _theta = r[10];
        _wy = r[11];
        _udot = r[12];
        _vdot = r[13];
        _wdot = r[14];
        _wxdot = r[15];
        _wydot = r[16];
        _wzdot = r[17];
        _dtheta = r[18];
        _dphi = r[19];
        _dpsi = r[20];
        _crankOmega = r[21];
        _dOmega_crank = r[22];
        _current_gr = (int)r[23];
    }
};

// SimState / SimStateNr (dof24_gpu.cuh:1180-1230)

struct SimState {
    VehicleState _v_states;
    TMeasyState _tirelf_st;
    TMeasyState _tirerf_st;
    TMeasyState _tirelr_st;
    TMeasyState _tirerr_st;
    SuspensionState _suslf_st;
    SuspensionState _susrf_st;
    SuspensionState _suslr_st;
    SuspensionState _susrr_st;
    void from_record(const double* r) {
        int k = 0;
        _tirelf_st.from_record(r + k);
        k += TMeasyState::kLen;
        _tirerf_st.from_record(r + k);
        k += TMeasyState::kLen;
        _tirelr_st.from_record(r + k);
        k += TMeasyState::kLen;
        _tirerr_st.from_record(r + k);
        k += TMeasyState::kLen;
        _suslf_st.from_record(r + k);
        k += SuspensionState::kLen;
        _susrf_st.from_record(r + k);
        k += SuspensionState::kLen;
        _suslr_st.from_record(r + k);
        k += SuspensionState::kLen;
        _susrr_st.from_record(r + k);
        k += SuspensionState::kLen;
        _v_states.from_record(r + k);
    }
};

struct SimStateNr {
    VehicleState _v_states;
    TMeasyNrState _tirelf_st;
    TMeasyNrState _tirerf_st;
    TMeasyNrState _tirelr_st;
    TMeasyNrState _tirerr_st;
    SuspensionState _suslf_st;
    SuspensionState _susrf_st;
    SuspensionState _suslr_st;
    SuspensionState _susrr_st;
    void from_record(const double* r) {
        int k = 0;
        _tirelf_st.from_record(r + k);
        k += TMeasyNrState::kLen;
        _tirerf_st.from_record(r + k);
        k += TMeasyNrState::kLen;
        _tirelr_st.from_record(r + k);
        k += TMeasyNrState::kLen;
        _tirerr_st.from_record(r + k);
        k += TMeasyNrState::kLen;
        _suslf_st.from_record(r + k);
        k += SuspensionState::kLen;
        _susrf_st.from_record(r + k);
        k += SuspensionState::kLen;
        _suslr_st.from_record(r + k);
        k += SuspensionState::kLen;
        _susrr_st.from_record(r + k);
        k += SuspensionState::kLen;
        _v_states.from_record(r + k);
    }
};

// VehicleParam (dof24_gpu.cuh:580-764); maps and gear ratios as vectors instead of device pointers and sizes.
struct VehicleParam : detail::PowertrainParam {
    double _a = 0., _b = 0., _h = 0., _m = 0., _jz = 0., _jx = 0., _jy = 0., _cf = 0., _cr = 0., _muf = 0.,
           _mur = 0., _hrcf = 0., _hrcr = 0., _maxSteer = 0., _crankInertia = 0., _maxBrakeTorque = 0.;
    template <class V>
    void visit(V& f) {
        for (double* p : {&_a, &_b, &_h, &_m, &_jz, &_jx, &_jy, &_cf, &_cr, &_muf, &_mur, &_hrcf,
                          &_hrcr, &_maxSteer, &_crankInertia, &_maxBrakeTorque})
            f(*p);
        visit_powertrain(f);
    }
};
using lfdm_amd::TMeasyNrParam;
using lfdm_amd::TMeasyParam;

// SuspensionParam (dof24_gpu.cuh:950-955)
struct SuspensionParam {
    double _ks = 0., _bs = 0.;
    template <class V>
    void visit(V& f) {
        f(_ks);
        f(_bs);
    }
};

}  // namespace d24GPU


namespace detail {

// The members every model's class shares (dof18_halfImplicit_gpu.cuh:93-165, 204).
class SolverBase {
  public:
    SolverBase(int dof, unsigned total) : dof_(dof) { check(lfdm_create(dof, total, &h_)); }
    ~SolverBase() { lfdm_destroy(h_); }
    SolverBase(const SolverBase&) = delete;
    SolverBase& operator=(const SolverBase&) = delete;

    void SetEndTime(double tend) { check(lfdm_set_end_time(h_, tend)); }
    void SetTimeStep(double step) { check(lfdm_set_time_step(h_, step)); }
    void SetKernelSimTime(double time) { check(lfdm_set_kernel_sim_time(h_, time)); }
    void SetHostDumpTime(double time) { check(lfdm_set_host_dump_time(h_, time)); }
    void SetThreadsPerBlock(unsigned threads) { check(lfdm_set_threads_per_block(h_, threads)); }
    double GetStep() const {
        double v;
        check(lfdm_get_step(h_, &v));
        return v;
    }
    double GetEndTime() const {
        double v;
        check(lfdm_get_end_time(h_, &v));
        return v;
    }
    TireType GetTireType() const {
        int v;
        check(lfdm_get_tire_type(h_, &v));
        return static_cast<TireType>(v);
    }
    void SetOutput(const std::string& output_file, double output_freq, bool store_all = false,
                   unsigned int no_outs = 50) {
        check(lfdm_set_output(h_, output_file.c_str(), output_freq, store_all ? 1 : 0, no_outs));
    }
    void Solve() { check(lfdm_solve(h_)); }
    double SolveStep(double t, double steering, double throttle, double braking) {
        double t_out;
        check(lfdm_solve_step(h_, t, steering, throttle, braking, &t_out));
        return t_out;
    }
    // Full record of one vehicle (layout of lfdm_amd.h).
    std::vector<double> GetRecord(unsigned vehicle) const {
        std::vector<double> r(record_len());
        check(lfdm_get_sim_state(h_, vehicle, r.data()));
        return r;
    }
    int record_len() const {
        return check_len(lfdm_record_len(dof_, static_cast<int>(GetTireType())));
    }
    // The C-ABI handle, for the batched extensions of lfdm_amd.h.
    lfdm_solver* handle() const { return h_; }
    // One vehicle's parameters as the flat vector of lfdm_params_from_json (lfdm_get_vehicle_params).
    std::vector<double> GetParamsFlat(unsigned vehicle) const {
        std::vector<double> v(8192);
        int n = 0;
        check(lfdm_get_vehicle_params(h_, vehicle, v.data(), (int)v.size(), &n));
        v.resize((size_t)n);
        return v;
    }
    void SetParamsFlat(unsigned vehicle, const std::vector<double>& flat) {
        check(lfdm_set_vehicle_params(h_, vehicle, flat.data(), (int)flat.size()));
    }

  protected:
    template <class... P>
    void get_params(unsigned vehicle, P&... parts) const {
        const std::vector<double> flat = GetParamsFlat(vehicle);
        FlatIn in{flat};
        (parts.visit(in), ...);
    }
    template <class... P>
    void set_params(unsigned vehicle, P... parts) {  // copies: visit() walks the fields of a mutable object
        FlatOut out;
        (parts.visit(out), ...);
        SetParamsFlat(vehicle, out.v);
    }
    static int check_len(int n) {
        check(n);
        return n;
    }
    void construct(const std::string& veh, const std::string& tire, const char* sus, unsigned n, const char* driver,
                   TireType type) {
        check(lfdm_construct(h_, veh.c_str(), tire.c_str(), sus, n, driver, static_cast<int>(type)));
    }
    int dof_;
    lfdm_solver* h_ = nullptr;
};

}  // namespace detail

// d18SolverHalfImplicitGPU (dof18_halfImplicit_gpu.cuh:15-343)
class d18SolverHalfImplicitGPU : public detail::SolverBase {
  public:
    explicit d18SolverHalfImplicitGPU(unsigned int total_num_vehicles) : SolverBase(18, total_num_vehicles) {}
    // Construct overloads, .cuh:34-87
    void Construct(const std::string& vehicle_params_file, const std::string& tire_params_file,
                   unsigned int num_vehicles, const std::string& driver_inputs_file, TireType type = TireType::TMeasy) {
        construct(vehicle_params_file, tire_params_file, nullptr, num_vehicles, driver_inputs_file.c_str(), type);
    }
    void Construct(const std::string& vehicle_params_file, const std::string& tire_params_file,
                   unsigned int num_vehicles, TireType type = TireType::TMeasy) {
        construct(vehicle_params_file, tire_params_file, nullptr, num_vehicles, nullptr, type);
    }
    // Initialize overloads, .cuh:172-195
    void Initialize(d18GPU::VehicleState& v, d18GPU::TMeasyState& lf, d18GPU::TMeasyState& rf, d18GPU::TMeasyState& lr,
                    d18GPU::TMeasyState& rr, unsigned int num_vehicles) {
        init4(v, lf, rf, lr, rr, num_vehicles);
    }
    void Initialize(d18GPU::VehicleState& v, d18GPU::TMeasyNrState& lf, d18GPU::TMeasyNrState& rf,
                    d18GPU::TMeasyNrState& lr, d18GPU::TMeasyNrState& rr, unsigned int num_vehicles) {
        init4(v, lf, rf, lr, rr, num_vehicles);
    }
    // GetSimState, .cuh:204 (TMeasy); GetSimStateNr is the TMeasyNr form
    d18GPU::SimState GetSimState(unsigned int vehicle_index) const { return get<d18GPU::SimState>(vehicle_index); }
    d18GPU::SimStateNr GetSimStateNr(unsigned int vehicle_index) const {
        return get<d18GPU::SimStateNr>(vehicle_index);
    }

    // A vehicle's parameters, as the reference's m_sim_data[vehicle] holds them (the per-vehicle
    // parameters applied); SetParams gives one vehicle its own (any field).
    void GetParams(unsigned vehicle, d18GPU::VehicleParam& veh, d18GPU::TMeasyParam& tire) const {
        get_params(vehicle, veh, tire);
    }
    void GetParams(unsigned vehicle, d18GPU::VehicleParam& veh, d18GPU::TMeasyNrParam& tire) const {
        get_params(vehicle, veh, tire);
    }
    void SetParams(unsigned vehicle, const d18GPU::VehicleParam& veh, const d18GPU::TMeasyParam& tire) {
        set_params(vehicle, veh, tire);
    }
    void SetParams(unsigned vehicle, const d18GPU::VehicleParam& veh, const d18GPU::TMeasyNrParam& tire) {
        set_params(vehicle, veh, tire);
    }

  private:
    template <class TS>
    void init4(d18GPU::VehicleState& v, TS& lf, TS& rf, TS& lr, TS& rr, unsigned n) {
        std::vector<double> r(4 * TS::kLen + d18GPU::VehicleState::kLen);
        lf.to_record(&r[0]);
        rf.to_record(&r[TS::kLen]);
        lr.to_record(&r[2 * TS::kLen]);
        rr.to_record(&r[3 * TS::kLen]);
        v.to_record(&r[4 * TS::kLen]);
        check(lfdm_initialize(h_, r.data(), n));
    }
    template <class SS>
    SS get(unsigned idx) const {
        SS s;
        s.from_record(GetRecord(idx).data());
        return s;
    }
};

// d11SolverHalfImplicitGPU (dof11_halfImplicit_gpu.cuh:16-330)
class d11SolverHalfImplicitGPU : public detail::SolverBase {
  public:
    explicit d11SolverHalfImplicitGPU(unsigned int total_num_vehicles) : SolverBase(11, total_num_vehicles) {}
    void Construct(const std::string& vehicle_params_file, const std::string& tire_params_file,
                   unsigned int num_vehicles, const std::string& driver_inputs_file, TireType type = TireType::TMeasy) {
        construct(vehicle_params_file, tire_params_file, nullptr, num_vehicles, driver_inputs_file.c_str(), type);
    }
    void Construct(const std::string& vehicle_params_file, const std::string& tire_params_file,
                   unsigned int num_vehicles, TireType type = TireType::TMeasy) {
        construct(vehicle_params_file, tire_params_file, nullptr, num_vehicles, nullptr, type);
    }
    // Initialize overloads, .cuh:167-190
    void Initialize(d11GPU::VehicleState& v, d11GPU::TMeasyState& f, d11GPU::TMeasyState& r,
                    unsigned int num_vehicles) {
        init2(v, f, r, num_vehicles);
    }
    void Initialize(d11GPU::VehicleState& v, d11GPU::TMeasyNrState& f, d11GPU::TMeasyNrState& r,
                    unsigned int num_vehicles) {
        init2(v, f, r, num_vehicles);
    }
    d11GPU::SimState GetSimState(unsigned int vehicle_index) const { return get<d11GPU::SimState>(vehicle_index); }
    d11GPU::SimStateNr GetSimStateNr(unsigned int vehicle_index) const {
        return get<d11GPU::SimStateNr>(vehicle_index);
    }

    // A vehicle's parameters, as the reference's m_sim_data[vehicle] holds them (the per-vehicle
    // parameters applied); SetParams gives one vehicle its own (any field).
    void GetParams(unsigned vehicle, d11GPU::VehicleParam& veh, d11GPU::TMeasyParam& tire) const {
        get_params(vehicle, veh, tire);
    }
    void GetParams(unsigned vehicle, d11GPU::VehicleParam& veh, d11GPU::TMeasyNrParam& tire) const {
        get_params(vehicle, veh, tire);
    }
    void SetParams(unsigned vehicle, const d11GPU::VehicleParam& veh, const d11GPU::TMeasyParam& tire) {
        set_params(vehicle, veh, tire);
    }
    void SetParams(unsigned vehicle, const d11GPU::VehicleParam& veh, const d11GPU::TMeasyNrParam& tire) {
        set_params(vehicle, veh, tire);
    }

  private:
    template <class TS>
    void init2(d11GPU::VehicleState& v, TS& f, TS& r, unsigned n) {
        std::vector<double> rec(2 * TS::kLen + d11GPU::VehicleState::kLen);
        f.to_record(&rec[0]);
        r.to_record(&rec[TS::kLen]);
        v.to_record(&rec[2 * TS::kLen]);
        check(lfdm_initialize(h_, rec.data(), n));
    }
    template <class SS>
    SS get(unsigned idx) const {
        SS s;
        s.from_record(GetRecord(idx).data());
        return s;
    }
};

// d24SolverHalfImplicitGPU (dof24_halfImplicit_gpu.cuh:15-330)
class d24SolverHalfImplicitGPU : public detail::SolverBase {
  public:
    explicit d24SolverHalfImplicitGPU(unsigned int total_num_vehicles) : SolverBase(24, total_num_vehicles) {}
    // Construct overloads with the suspension file after the tire file, .cuh:35-95
    void Construct(const std::string& vehicle_params_file, const std::string& tire_params_file,
                   const std::string& sus_params_file, unsigned int num_vehicles, const std::string& driver_inputs_file,
                   TireType type = TireType::TMeasy) {
        construct(vehicle_params_file, tire_params_file, sus_params_file.c_str(), num_vehicles,
                  driver_inputs_file.c_str(), type);
    }
    void Construct(const std::string& vehicle_params_file, const std::string& tire_params_file,
                   const std::string& sus_params_file, unsigned int num_vehicles, TireType type = TireType::TMeasy) {
        construct(vehicle_params_file, tire_params_file, sus_params_file.c_str(), num_vehicles, nullptr, type);
    }
    // Initialize overloads, .cuh:184-236; like the reference, the caller's structs receive the
    // result of initializeTireSus (that of the last vehicle appended)
    void Initialize(d24GPU::VehicleState& v, d24GPU::TMeasyState& lf, d24GPU::TMeasyState& rf, d24GPU::TMeasyState& lr,
                    d24GPU::TMeasyState& rr, d24GPU::SuspensionState& slf, d24GPU::SuspensionState& srf,
                    d24GPU::SuspensionState& slr, d24GPU::SuspensionState& srr, unsigned int num_vehicles) {
        init(v, lf, rf, lr, rr, slf, srf, slr, srr, num_vehicles);
    }
    void Initialize(d24GPU::VehicleState& v, d24GPU::TMeasyNrState& lf, d24GPU::TMeasyNrState& rf,
                    d24GPU::TMeasyNrState& lr, d24GPU::TMeasyNrState& rr, d24GPU::SuspensionState& slf,
                    d24GPU::SuspensionState& srf, d24GPU::SuspensionState& slr, d24GPU::SuspensionState& srr,
                    unsigned int num_vehicles) {
        init(v, lf, rf, lr, rr, slf, srf, slr, srr, num_vehicles);
    }
    d24GPU::SimState GetSimState(unsigned int vehicle_index) const { return get<d24GPU::SimState>(vehicle_index); }
    d24GPU::SimStateNr GetSimStateNr(unsigned int vehicle_index) const {
        return get<d24GPU::SimStateNr>(vehicle_index);
    }

    // A vehicle's parameters, as the reference's m_sim_data[vehicle] holds them (the per-vehicle
    // parameters applied); SetParams gives one vehicle its own (any field).
    void GetParams(unsigned vehicle, d24GPU::VehicleParam& veh, d24GPU::TMeasyParam& tire, d24GPU::SuspensionParam& sus) const {
        get_params(vehicle, veh, tire, sus);
    }
    void GetParams(unsigned vehicle, d24GPU::VehicleParam& veh, d24GPU::TMeasyNrParam& tire, d24GPU::SuspensionParam& sus) const {
        get_params(vehicle, veh, tire, sus);
    }
    void SetParams(unsigned vehicle, const d24GPU::VehicleParam& veh, const d24GPU::TMeasyParam& tire, const d24GPU::SuspensionParam& sus) {
        set_params(vehicle, veh, tire, sus);
    }
    void SetParams(unsigned vehicle, const d24GPU::VehicleParam& veh, const d24GPU::TMeasyNrParam& tire, const d24GPU::SuspensionParam& sus) {
        set_params(vehicle, veh, tire, sus);
    }

  private:
    template <class TS>
    void init(d24GPU::VehicleState& v, TS& lf, TS& rf, TS& lr, TS& rr, d24GPU::SuspensionState& slf,
              d24GPU::SuspensionState& srf, d24GPU::SuspensionState& slr, d24GPU::SuspensionState& srr, unsigned n) {
        using SusS = d24GPU::SuspensionState;
        std::vector<double> rec(4 * TS::kLen + 4 * SusS::kLen + d24GPU::VehicleState::kLen);
        TS* t[4] = {&lf, &rf, &lr, &rr};
        SusS* q[4] = {&slf, &srf, &slr, &srr};
        for (int i = 0; i < 4; ++i) t[i]->to_record(&rec[i * TS::kLen]);
        for (int i = 0; i < 4; ++i) q[i]->to_record(&rec[4 * TS::kLen + i * SusS::kLen]);
        v.to_record(&rec[4 * TS::kLen + 4 * SusS::kLen]);
        check(lfdm_initialize(h_, rec.data(), n));
        if (n == 0) return;
        // the appended vehicles are the last n initialized; read the last one back
        unsigned count = 0;
        check(lfdm_initialized_count(h_, &count));
        std::vector<double> out(rec.size());
        check(lfdm_get_sim_state(h_, count - 1, out.data()));
        for (int i = 0; i < 4; ++i) t[i]->from_record(&out[i * TS::kLen]);
        for (int i = 0; i < 4; ++i) q[i]->from_record(&out[4 * TS::kLen + i * SusS::kLen]);
        v.from_record(&out[4 * TS::kLen + 4 * SusS::kLen]);
    }
    template <class SS>
    SS get(unsigned idx) const {
        SS s;
        s.from_record(GetRecord(idx).data());
        return s;
    }
};

}  // namespace lfdm_amd
