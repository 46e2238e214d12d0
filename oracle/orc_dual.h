// ORACLE — test infrastructure only (see orc_common.h).
//
// A forward-mode dual scalar: a value and N directional derivatives (tangents), carried through the
// oracle's own source by operator overloading, so the restatement of the reference's step yields the
// exact derivatives of the discrete half-implicit map it computes.  Every branch of the restatement
// compares values (val), so the derivative is the one of the smooth piece the trajectory is on.
// Conventions at the non-smooth points the model reaches: d|x| = sgn(x) dx with sgn(0) = 0, and
// hypot(0, 0) has zero derivative.
//
// Pinned in tests/test_oracle_dual.py against what the mathematics fixes: the derivative rules of
// each operation at random points against complex-step or high-order finite differences of the same
// double function, and whole-step tangents against fourth-order central differences of the double
// oracle.
#pragma once
#include <cmath>

namespace orc {

template <int N>
struct Dual {
    double v;
    double d[N];
    Dual() : v(0) {
        for (int i = 0; i < N; ++i) d[i] = 0;
    }
    Dual(double x) : v(x) {  // NOLINT: constants enter arithmetic implicitly, as doubles do
        for (int i = 0; i < N; ++i) d[i] = 0;
    }
};

template <int N>
inline double val(const Dual<N>& x) {
    return x.v;
}

// a + b, a - b, a * b, a / b with the tangents of both operands
template <int N>
inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N>
inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}

// mixed with plain numbers (constants)
template <int N> inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> inline Dual<N> operator*(const Dual<N>& a, double b) { return a * Dual<N>(b); }
template <int N> inline Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> inline Dual<N> operator*(double a, const Dual<N>& b) { return Dual<N>(a) * b; }
template <int N> inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, int b) { return a * Dual<N>((double)b); }
template <int N> inline Dual<N> operator*(int a, const Dual<N>& b) { return Dual<N>((double)a) * b; }
template <int N> inline Dual<N> operator/(int a, const Dual<N>& b) { return Dual<N>((double)a) / b; }
template <int N> inline Dual<N> operator+(int a, const Dual<N>& b) { return Dual<N>((double)a) + b; }
template <int N> inline Dual<N> operator-(int a, const Dual<N>& b) { return Dual<N>((double)a) - b; }

// f(a) with derivative fp: the chain rule
template <int N>
inline Dual<N> chain(const Dual<N>& a, double f, double fp) {
    Dual<N> r(f);
    for (int i = 0; i < N; ++i) r.d[i] = fp * a.d[i];
    return r;
}

template <int N>
inline Dual<N> fabs(const Dual<N>& a) {
    return chain(a, std::fabs(a.v), (double)((0.0 < a.v) - (a.v < 0.0)));
}
template <int N>
inline Dual<N> sqrt(const Dual<N>& a) {
    const double s = std::sqrt(a.v);
    return chain(a, s, s > 0 ? 0.5 / s : 0.0);
}
template <int N>
inline Dual<N> sin(const Dual<N>& a) {
    return chain(a, std::sin(a.v), std::cos(a.v));
}
template <int N>
inline Dual<N> cos(const Dual<N>& a) {
    return chain(a, std::cos(a.v), -std::sin(a.v));
}
template <int N>
inline Dual<N> tan(const Dual<N>& a) {
    const double t = std::tan(a.v);
    return chain(a, t, 1.0 + t * t);
}
template <int N>
inline Dual<N> tanh(const Dual<N>& a) {
    const double t = std::tanh(a.v);
    return chain(a, t, 1.0 - t * t);
}
template <int N>
inline Dual<N> atan2(const Dual<N>& y, const Dual<N>& x) {
    // d atan2(y, x) = (x dy - y dx) / (x^2 + y^2)
    Dual<N> r(std::atan2(y.v, x.v));
    const double q = x.v * x.v + y.v * y.v;
    for (int i = 0; i < N; ++i) r.d[i] = q > 0 ? (x.v * y.d[i] - y.v * x.d[i]) / q : 0.0;
    return r;
}
template <int N>
inline Dual<N> hypot(const Dual<N>& a, const Dual<N>& b) {
    // d hypot(a, b) = (a da + b db) / hypot(a, b)
    Dual<N> r(std::hypot(a.v, b.v));
    for (int i = 0; i < N; ++i) r.d[i] = r.v > 0 ? (a.v * a.d[i] + b.v * b.d[i]) / r.v : 0.0;
    return r;
}

}  // namespace orc
