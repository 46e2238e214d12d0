// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product library never links or calls it.
//
// Common pieces of the CPU restatement of the reference's half-implicit vehicle solvers:
// scalar helpers, piecewise-linear maps, driver-table lookup and the flat parameter reader.
// Everything is templated on the scalar type T so the same source runs in double (the oracle
// proper), in a counting type (FLOPs per step) and in a dual type (forward tangents).
#pragma once
#include <cmath>
#include <cstddef>
#include <vector>

namespace orc {

// wheeled_vehicle_models/utils/utils.h:11-14
static const double kPi = 3.141592653589793238462643383279;
static const double k2Pi = 6.283185307179586476925286766559;
static const double kG = 9.81;

inline double val(double x) { return x; }

// Marks a restated function that is its own function in the reference's physics sources
// (dof24.cpp).  g++ -O2 and above turns sin(x) and cos(x) of one argument inside one function into a
// single sincos call, and glibc's sincos does not round exactly like its sin and cos for every
// argument; the reference's physics functions are compiled apart from the step that calls them, so
// each pairs only its own calls.  Keeping these functions out of line gives the restatement the same
// pairing (the step's Euler-angle rates, for one, call cos(theta) without sin(theta)).
#define ORC_REF_FN __attribute__((noinline))

template <class T>
struct MapE {
    T x, y;
};

// Clamped piecewise-linear lookup with std::lower_bound interval choice (utils.cpp:109-125).
template <class T>
T map_y(const std::vector<MapE<T>>& m, const T& x) {
    using std::fabs;
    if (val(x) <= val(m[0].x)) return m[0].y;
    if (val(x) >= val(m.back().x)) return m.back().y;
    size_t r = 0;
    while (r < m.size() && val(m[r].x) < val(x)) ++r;  // first entry with !(entry.x < x)
    const size_t l = r - 1;
    T mbar = (x - m[l].x) / (m[r].x - m[l].x);
    return m[l].y + mbar * (m[r].y - m[l].y);
}

// utils.h:93-95
template <class T>
T interp_l(const T& fz, const T& w1, const T& w2, const T& pn) {
    return w1 + (w2 - w1) * (fz / pn - 1.);
}
// utils.h:98-100
template <class T>
T interp_q(const T& fz, const T& w1, const T& w2, const T& pn) {
    return (fz / pn) * (2. * w1 - 0.5 * w2 - (w1 - 0.5 * w2) * (fz / pn));
}
// utils.h:104-106
template <class T>
int sgn(const T& v) {
    return (0.0 < val(v)) - (val(v) < 0.0);
}
// utils.cpp:128-138
template <class T>
T sine_step(const T& x, double x1, double y1, double x2, double y2) {
    using std::sin;
    if (val(x) <= x1) return T(y1);
    if (val(x) >= x2) return T(y2);
    double dx = x2 - x1;
    double dy = y2 - y1;
    return y1 + dy * (x - x1) / dx - (dy / k2Pi) * sin(k2Pi * (x - x1) / dx);
}
// utils.h:110-117
template <class T>
T clamp(const T& v, const T& lo, const T& hi) {
    if (val(v) < val(lo)) return lo;
    if (val(v) > val(hi)) return hi;
    return v;
}

// Driver-table row (utils.h:19-31) and its clamped lerp (utils.cpp:34-56).
struct DriverRow {
    double t, steering, throttle, braking;
};
inline void driver_input(double t, const DriverRow* rows, int n, double& str, double& thr, double& brk) {
    if (t <= rows[0].t) {
        str = rows[0].steering; thr = rows[0].throttle; brk = rows[0].braking;
        return;
    }
    if (t >= rows[n - 1].t) {
        str = rows[n - 1].steering; thr = rows[n - 1].throttle; brk = rows[n - 1].braking;
        return;
    }
    int r = 0;
    while (r < n && rows[r].t < t) ++r;
    const int l = r - 1;
    double tbar = (t - rows[l].t) / (rows[r].t - rows[l].t);
    str = rows[l].steering + tbar * (rows[r].steering - rows[l].steering);
    thr = rows[l].throttle + tbar * (rows[r].throttle - rows[l].throttle);
    brk = rows[l].braking + tbar * (rows[r].braking - rows[l].braking);
}

// Synthetic "knot program" controls: steering lerped on a uniform grid of `ds` seconds
// (ns knots), throttle and braking piecewise constant on a grid of `dp` seconds (np values each).
// knots layout: [ns steering][np throttle][np braking].
inline void knot_input(double t, const double* kn, int ns, double ds, int np, double dp, double& str, double& thr,
                       double& brk) {
    int k = (int)std::floor(t / ds);
    if (k < 0) k = 0;
    if (k > ns - 2) k = ns - 2;
    double tk = k * ds;
    double tbar = (t - tk) / ds;
    if (tbar > 1.0) tbar = 1.0;
    str = kn[k] + tbar * (kn[k + 1] - kn[k]);
    int j = (int)std::floor(t / dp);
    if (j < 0) j = 0;
    if (j > np - 1) j = np - 1;
    thr = kn[ns + j];
    brk = kn[ns + np + j];
}

// Sequential reader over the flat parameter vector.
struct Flat {
    const double* p;
    int i = 0;
    double next() { return p[i++]; }
    template <class T>
    void map(std::vector<MapE<T>>& m) {
        int n = (int)next();
        m.resize(n);
        for (int k = 0; k < n; ++k) {
            m[k].x = T(next());
            m[k].y = T(next());
        }
    }
};

}  // namespace orc
