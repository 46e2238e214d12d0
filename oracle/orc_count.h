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
    // libm calls with distinct arguments within one step (a call repeating an earlier call of the
    // same function on the same argument in that step is not counted here): the transcendental work a
    // step that evaluates each value once performs
    long long dsin = 0, dcos = 0, dtan = 0, datan2 = 0, dhypot = 0, dtanh = 0;
};
extern thread_local OpCount g_ops;

// Calls of the current step, for the distinct tallies (cleared by the runner at every step).
struct SeenCall {
    int fn;
    double a, b;
};
extern thread_local SeenCall g_seen[512];
extern thread_local int g_nseen;

inline bool first_call(int fn, double a, double b = 0.0) {
    for (int i = 0; i < g_nseen; ++i)
        if (g_seen[i].fn == fn && g_seen[i].a == a && g_seen[i].b == b) return false;
    if (g_nseen < 512) g_seen[g_nseen++] = SeenCall{fn, a, b};
    return true;
}
inline void new_step() { g_nseen = 0; }

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
inline CD sin(CD a) { ++g_ops.sin; g_ops.dsin += first_call(0, a.v); return CD(std::sin(a.v)); }
inline CD cos(CD a) { ++g_ops.cos; g_ops.dcos += first_call(1, a.v); return CD(std::cos(a.v)); }
inline CD tan(CD a) { ++g_ops.tan; g_ops.dtan += first_call(2, a.v); return CD(std::tan(a.v)); }
inline CD tanh(CD a) { ++g_ops.tanh; g_ops.dtanh += first_call(3, a.v); return CD(std::tanh(a.v)); }
inline CD atan2(CD a, CD b) { ++g_ops.atan2; g_ops.datan2 += first_call(4, a.v, b.v); return CD(std::atan2(a.v, b.v)); }
inline CD hypot(CD a, CD b) { ++g_ops.hypot; g_ops.dhypot += first_call(5, a.v, b.v); return CD(std::hypot(a.v, b.v)); }

}  // namespace orc
