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
LFDM_HD double sine_step(double x, double x1, double y1, double x2, double y2) {
    if (x <= x1) return y1;
    if (x >= x2) return y2;
    const double dx = x2 - x1, dy = y2 - y1;
    return y1 + dy * (x - x1) / dx - (dy / k2Pi) * sin(k2Pi * (x - x1) / dx);
}

// Clamped piecewise-linear lookup (utils.cpp:109-125). `m` is wave-uniform, so the knot scan
// runs on scalar loads; the interval index counts knots strictly below x, which is exactly
// std::lower_bound's answer on a sorted map.
LFDM_HD double map_y(const Map& m, double x) {
    const int n = m.n;
    if (x <= m.e[0].x) return m.e[0].y;
    if (x >= m.e[n - 1].x) return m.e[n - 1].y;
    int r = 0;
    for (int k = 1; k < n - 1; ++k) r += (m.e[k].x < x) ? 1 : 0;
    r += 1;
    const MapXY L = m.e[r - 1];
    const MapXY R = m.e[r];
    const double mbar = (x - L.x) / (R.x - L.x);
    return L.y + mbar * (R.y - L.y);
}

// Same lookup on a map whose abscissae are scaled by sx and ordinates by sy1 then sy2
// (multiplying by 1.0 is exact, so unit factors reproduce the unscaled map bit for bit).
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
    const double mbar = (x - lx) / (rx - lx);
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
    const double tbar = (t - L.t) / (R.t - L.t);
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
