// ORACLE — test infrastructure only (see orc_common.h).
//
// A counting scalar: behaves as double and tallies every arithmetic operation and libm call, so
// the oracle's own source yields the algorithmic FLOPs of one vehicle-step.  Counting rule:
// + - * / sqrt and fabs-free comparisons: + - * / sqrt count 1 each; each libm call is tallied
// separately by name and weighted afterwards.
#pragma once
#include <cmath>

namespace orc {

struct OpCount {
    long long add = 0, mul = 0, div = 0, sqrt = 0, sin = 0, cos = 0, tan = 0, atan2 = 0, hypot = 0, tanh = 0,
              floor_ = 0;
};
extern thread_local OpCount g_ops;

struct CD {
    double v;
    CD() : v(0) {}
    CD(double x) : v(x) {}
};
inline double val(const CD& x) { return x.v; }

inline CD operator+(CD a, CD b) { ++g_ops.add; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { ++g_ops.add; return CD(a.v - b.v); }
inline CD operator*(CD a, CD b) { ++g_ops.mul; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { ++g_ops.div; return CD(a.v / b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
inline CD operator+(CD a, double b) { return a + CD(b); }
inline CD operator-(CD a, double b) { return a - CD(b); }
inline CD operator*(CD a, double b) { return a * CD(b); }
inline CD operator/(CD a, double b) { return a / CD(b); }
inline CD operator+(double a, CD b) { return CD(a) + b; }
inline CD operator-(double a, CD b) { return CD(a) - b; }
inline CD operator*(double a, CD b) { return CD(a) * b; }
inline CD operator/(double a, CD b) { return CD(a) / b; }
inline CD operator*(CD a, int b) { return a * CD((double)b); }
inline CD operator*(int a, CD b) { return CD((double)a) * b; }
inline CD operator/(int a, CD b) { return CD((double)a) / b; }
inline CD operator+(int a, CD b) { return CD((double)a) + b; }
inline CD operator-(int a, CD b) { return CD((double)a) - b; }
inline CD fabs(CD a) { return CD(std::fabs(a.v)); }
inline CD sqrt(CD a) { ++g_ops.sqrt; return CD(std::sqrt(a.v)); }
inline CD sin(CD a) { ++g_ops.sin; return CD(std::sin(a.v)); }
inline CD cos(CD a) { ++g_ops.cos; return CD(std::cos(a.v)); }
inline CD tan(CD a) { ++g_ops.tan; return CD(std::tan(a.v)); }
inline CD tanh(CD a) { ++g_ops.tanh; return CD(std::tanh(a.v)); }
inline CD atan2(CD a, CD b) { ++g_ops.atan2; return CD(std::atan2(a.v, b.v)); }
inline CD hypot(CD a, CD b) { ++g_ops.hypot; return CD(std::hypot(a.v, b.v)); }

}  // namespace orc
