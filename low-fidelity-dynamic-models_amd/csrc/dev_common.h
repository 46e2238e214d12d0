// Device helpers shared by the vehicle-model kernels (gfx950, one vehicle per lane).
//
// Every helper keeps the reference's IEEE operation order (the library is compiled with
// -ffp-contract=off), so results are bit-identical to the reference CPU solver wherever the
// device libm agrees with glibc.
#pragma once
#include <hip/hip_runtime.h>

#include "lfdm_types.h"

namespace lfdm {
namespace dev {

constexpr double kG = 9.81;                                      // utils.h:13
constexpr double k2Pi = 6.283185307179586476925286766559;       // utils.h:12
constexpr double kHalfSqrt2 = 0.70710678118654757;              // sqrt(2.) / 2.

// Host-and-device helpers: the model code below also compiles for the host, where the adjoint's
// finite-difference checks run it.
#define LFDM_HD __host__ __device__ __forceinline__

// Keeps the compiler from carrying values read from memory across this point in registers.  Step
// loops that keep per-vehicle constants in LDS put it at the top of every step, so each step reads
// the constants where it uses them instead of holding them in registers for the whole loop.
#define LFDM_RELOAD_MEMORY() asm volatile("" ::: "memory")

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Quotient and square root of the step code.  On the device: Newton-Raphson refinement of the
// hardware reciprocal v_rcp_f64 (two iterations, then the quotient with one remainder correction,
// Markstein's method) and of the reciprocal square root v_rsq_f64 (Goldschmidt-style coupled
// iteration for sqrt(x) and 1 / (2 sqrt(x)), then two remainder corrections).  These are exactly the
// operations of the compiler's correctly rounded expansion of `/` and sqrt without its exponent
// range scaling (v_div_scale / v_div_fmas / v_div_fixup, v_ldexp) and special-operand fixups, so
// they return the same bits as `/` and sqrt whenever no scaling is needed: operands and quotient of
// qdiv finite, nonzero divisor, exponents within +-900 (a zero numerator can come back as a zero
// of the other sign, which no later operation of the step can tell apart), and qsqrt's x finite and
// >= 2^-767 or zero.  The step's operands are physical quantities far inside that
// range, and every divisor it passes is bounded away from zero.  Without the scaling and fixups a
// division costs 8 VALU instructions instead of 11 plus a VCC dependency, a square root 12 instead
// of 17.  Host builds (the adjoint's finite-difference checks) use `/` and sqrt.
LFDM_HD double qdiv(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = n * r;
    const double rem = __builtin_fma(-d, q, n);
    return __builtin_fma(rem, r, q);
#else
    return n / d;
#endif
}

LFDM_HD double qsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x == 0.0 ? x : g;  // v_rsq_f64(0) is +inf
#else
    return sqrt(x);
#endif
}

// Quotient, reciprocal and square root of the fast arithmetic form: the same refinements stopped one
// step earlier.  fdiv: one Newton-Raphson iteration on v_rcp_f64, then the quotient with one
// remainder correction (6 VALU instructions); frcp: two iterations (5); fsqrt: one coupled iteration
// from v_rsq_f64 and one remainder correction (9; the operand of v_rsq_f64 is raised to 2^-1000 so
// that x = 0 runs through to 0 without a select).  Each result is within one ulp of the correctly
// rounded one, which the refinement's error bound (the square of the first iteration's, itself the
// square of the hardware estimate's) makes the correctly rounded value for all but operands within a
// minute fraction of an ulp of a rounding boundary.  Same operand ranges as qdiv and qsqrt.  Host
// builds use `/` and sqrt.
LFDM_HD double fdiv(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = n * r;
    const double rem = __builtin_fma(-d, q, n);
    return __builtin_fma(rem, r, q);
#else
    return n / d;
#endif
}

LFDM_HD double frcp(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
#else
    return 1.0 / d;
#endif
}

LFDM_HD double fsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(__builtin_fmax(x, 0x1p-1000));
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
#else
    return sqrt(x);
#endif
}

// The quotient of the arithmetic form: correctly rounded (the reference's `/`) in the strict form,
// faithful in the fast one.
template <bool STRICT>
LFDM_HD double xdiv(double n, double d) {
    return STRICT ? qdiv(n, d) : fdiv(n, d);
}

LFDM_HD double sgn(double v) { return (double)((0.0 < v) - (v < 0.0)); }

LFDM_HD double clampd(double v, double lo, double hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// sineStep (utils.cpp:128-138) with x1, y1, x2, y2 = 1, 1, 1.5, 0: the tire-force weighting.
LFDM_HD double sine_step_weight(double x) {
    if (x <= 1.0) return 1.0;
    if (x >= 1.5) return 0.0;
    const double dx = 0.5, dy = -1.0;
    return 1.0 + dy * (x - 1.0) / dx - (dy / k2Pi) * sin(k2Pi * (x - 1.0) / dx);
}

// General sineStep (utils.cpp:128-138).
template <bool STRICT>
LFDM_HD double sine_step(double x, double x1, double y1, double x2, double y2) {
    if (x <= x1) return y1;
    if (x >= x2) return y2;
    const double dx = x2 - x1, dy = y2 - y1;
    return y1 + xdiv<STRICT>(dy * (x - x1), dx) - xdiv<STRICT>(dy, k2Pi) * sin(xdiv<STRICT>(k2Pi * (x - x1), dx));
}

// Clamped piecewise-linear lookup (utils.cpp:109-125). `m` is wave-uniform, so the knot scan
// runs on scalar loads; the interval index counts knots strictly below x, which is exactly
// std::lower_bound's answer on a sorted map.
template <bool STRICT>
LFDM_HD double map_y(const Map& m, double x) {
    const int n = m.n;
    if (x <= m.e[0].x) return m.e[0].y;
    if (x >= m.e[n - 1].x) return m.e[n - 1].y;
    int r = 0;
    for (int k = 1; k < n - 1; ++k) r += (m.e[k].x < x) ? 1 : 0;
    r += 1;
    const MapXY L = m.e[r - 1];
    const MapXY R = m.e[r];
    const double mbar = xdiv<STRICT>(x - L.x, R.x - L.x);
    return L.y + mbar * (R.y - L.y);
}

// Same lookup on a map whose abscissae are scaled by sx and ordinates by sy1 then sy2
// (multiplying by 1.0 is exact, so unit factors reproduce the unscaled map bit for bit).
template <bool STRICT>
LFDM_HD double map_y_scaled(const Map& m, double x, double sx, double sy1, double sy2) {
    const int n = m.n;
    if (x <= m.e[0].x * sx) return m.e[0].y * sy1 * sy2;
    if (x >= m.e[n - 1].x * sx) return m.e[n - 1].y * sy1 * sy2;
    int r = 0;
    for (int k = 1; k < n - 1; ++k) r += (m.e[k].x * sx < x) ? 1 : 0;
    r += 1;
    const MapXY L = m.e[r - 1];
    const MapXY R = m.e[r];
    const double lx = L.x * sx, rx = R.x * sx, ly = L.y * sy1 * sy2, ry = R.y * sy1 * sy2;
    const double mbar = xdiv<STRICT>(x - lx, rx - lx);
    return ly + mbar * (ry - ly);
}

// Driver-table lookup at a wave-uniform time (utils.cpp:34-56).
LFDM_HD void driver_input(double t, const DrvRow* __restrict__ rows, int n, double& s, double& th,
                                             double& b) {
    if (t <= rows[0].t) {
        s = rows[0].steering; th = rows[0].throttle; b = rows[0].braking;
        return;
    }
    if (t >= rows[n - 1].t) {
        s = rows[n - 1].steering; th = rows[n - 1].throttle; b = rows[n - 1].braking;
        return;
    }
    int r = 1;
    while (r < n - 1 && rows[r].t < t) ++r;
    const DrvRow L = rows[r - 1];
    const DrvRow R = rows[r];
    const double tbar = qdiv(t - L.t, R.t - L.t);
    s = L.steering + tbar * (R.steering - L.steering);
    th = L.throttle + tbar * (R.throttle - L.throttle);
    b = L.braking + tbar * (R.braking - L.braking);
}

// Gear-indexed reads by a per-lane gear over a wave-uniform table (select chain, no memory).
LFDM_HD double pick(const double* tab, int n, int i) {
    double v = tab[0];
    for (int k = 1; k < n; ++k) v = (i == k) ? tab[k] : v;
    return v;
}

}  // namespace dev
}  // namespace lfdm
