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

// Makes v opaque to the compiler (no instruction is emitted; the value is unchanged): expressions
// computed from it afterwards are not merged with the same expressions computed from it before, so
// code that recomputes a cheap value near a late use does not get the early result held live in
// registers in between instead.
LFDM_HD void opaque(double& v) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(v));
#else
    (void)v;
#endif
}

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

// sqrt(x) and 1 / sqrt(x) from one v_rsq_f64 (x > 0 for the reciprocal; x = 0 gives g = 0): the
// coupled iteration of fsqrt and its remainder correction for g, then one Newton-Raphson step for
// h = 1 / (2 g) against the corrected g, h' = h (2 - 2 h g).  Both within one ulp of the correctly
// rounded values, for the cost of fsqrt plus three instructions instead of a further v_rcp_f64
// refinement.
LFDM_HD void fsqrt_rcp(double x, double& g, double& rg) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(__builtin_fmax(x, 0x1p-1000));
    g = x * y;
    double h = y * 0.5;
    double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    r = __builtin_fma(-h, g, 0.5);  // (1 - 2 h g) / 2
    h = __builtin_fma(h + h, r, h);
    rg = h + h;
#else
    g = sqrt(x);
    rg = 1.0 / g;
#endif
}

// The quotient of the arithmetic form: correctly rounded (the reference's `/`) in the strict form,
// faithful in the fast one.
template <bool STRICT>
LFDM_HD double xdiv(double n, double d) {
    return STRICT ? qdiv(n, d) : fdiv(n, d);
}

// Sine and cosine of the fast arithmetic form, from the Taylor polynomials of sin and cos on
// [-pi/4, pi/4] (to the r^17 and r^18 terms; truncation below 1e-19 there), evaluated by Horner's rule
// in r^2.  Coefficients: (-1)^k / (2k+1)! and (-1)^k / (2k)!, rounded to double.
LFDM_HD double sin_poly(double r, double z) {
    double p = 0x1.952c77030ad4ap-49;
    p = __builtin_fma(p, z, -0x1.ae7f3e733b81fp-41);
    p = __builtin_fma(p, z, 0x1.6124613a86d09p-33);
    p = __builtin_fma(p, z, -0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, z, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, z, -0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, z, 0x1.1111111111111p-7);
    p = __builtin_fma(p, z, -0x1.5555555555555p-3);
    return __builtin_fma(r * z, p, r);
}

LFDM_HD double cos_poly(double z) {
    double p = -0x1.6827863b97d97p-53;
    p = __builtin_fma(p, z, 0x1.ae7f3e733b81fp-45);
    p = __builtin_fma(p, z, -0x1.93974a8c07c9dp-37);
    p = __builtin_fma(p, z, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, z, -0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, z, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, z, -0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, z, 0x1.5555555555555p-5);
    p = __builtin_fma(p, z, -0.5);
    return __builtin_fma(z, p, 1.0);
}

// Word k of the bits of 2/pi, 32 per word, most significant first: word k >= 2 holds bits
// 32(k-2)+1 .. 32(k-1) after the binary point, words 0 and 1 (the bits at and above it) are zero.
// 37 words of 2/pi, computed exactly in integer arithmetic (Machin's formula for pi; checked by
// tests/test_strict_host.py).
LFDM_HD unsigned int two_over_pi_word(int k) {
    static constexpr unsigned int w[39] = {
        0x00000000u, 0x00000000u, 0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u,
        0xdb629599u, 0x3c439041u, 0xfe5163abu, 0xdebbc561u, 0xb7246e3au, 0x424dd2e0u,
        0x06492eeau, 0x09d1921cu, 0xfe1deb1cu, 0xb129a73eu, 0xe88235f5u, 0x2ebb4484u,
        0xe99c7026u, 0xb45f7e41u, 0x3991d639u, 0x835339f4u, 0x9c845f8bu, 0xbdf9283bu,
        0x1ff897ffu, 0xde05980fu, 0xef2f118bu, 0x5a0a6d1fu, 0x6d367ecfu, 0x27cb09b7u,
        0x4f463f66u, 0x9e5fea2du, 0x7527bac7u, 0xebe5f17bu, 0x3d0739f7u, 0x8a5292eau,
        0x6bfb5fb1u, 0x1f8d5d08u, 0x56033046u,
    };
    return w[k];
}

// Payne-Hanek reduction (M. Payne, R. Hanek, "Radian reduction for trigonometric functions",
// SIGNUM Newsletter 18(1), 1983) of a finite x >= 2^20: x 2/pi = n + f with f in [-1/2, 1/2);
// returns n mod 4 and r = f pi/2 as rh + rl.  With x = M 2^e (M the 53-bit integer significand,
// e >= -32), the words of 2/pi whose products with M are multiples of 4 are skipped; the next seven
// (224 bits) are multiplied by M exactly in 32-bit limbs, and the product's 128 bits from two above
// the binary point down give n mod 4 and f to 126 bits (the words left out weigh below 2^-170 in
// x 2/pi, far below the 2^-62 of the closest double to a multiple of pi/2).  f goes to
// double-double and is multiplied by pi/2 in two parts.  Integer arithmetic only up to there.
LFDM_HD int reduce_pio2_large(double x, double& rh, double& rl) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int e = (int)((b >> 52) & 0x7ff) - 1075;
    const unsigned long long m = (b & 0xfffffffffffffull) | 0x10000000000000ull;
    const int k0 = (e + 62) >> 5;       // first word whose product with M is not a multiple of 4
    const int rho = e + 62 - 32 * k0;   // the binary point of the product is at bit 222 - rho
    const unsigned int m0 = (unsigned int)m, m1 = (unsigned int)(m >> 32);
    unsigned int L[7];
    unsigned long long c = 0;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const unsigned long long t = (unsigned long long)two_over_pi_word(k0 + 6 - j) * m0 + c;
        L[j] = (unsigned int)t;
        c = t >> 32;
    }
    c = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const unsigned long long t = (unsigned long long)two_over_pi_word(k0 + 6 - j) * m1 + L[j + 1] + c;
        L[j + 1] = (unsigned int)t;
        c = t >> 32;
    }
    // bits 96 - rho .. 223 - rho of the product: two integer bits, then the fraction
    const unsigned long long p2 = (unsigned long long)L[6] << 32 | L[5];
    const unsigned long long p1 = (unsigned long long)L[4] << 32 | L[3];
    const unsigned long long p0 = (unsigned long long)L[2] << 32;
    const unsigned long long qh = (p2 << rho) | ((p1 >> 1) >> (63 - rho));
    const unsigned long long ql = (p1 << rho) | ((p0 >> 1) >> (63 - rho));
    int n = (int)(qh >> 62);
    unsigned long long fh = (qh << 2) | (ql >> 62), fl = ql << 2;  // f in [0, 1) as fh:fl / 2^128
    const bool up = (fh >> 63) != 0;                                 // f >= 1/2: n + 1, f - 1
    if (up) {
        n += 1;
        fl = ~fl + 1;
        fh = ~fh + (fl == 0 ? 1 : 0);
    }
    // |f| = fh 2^-64 + fl 2^-128 (fh < 2^63) to double-double, then times pi/2
    const double dh = (double)fh;
    const long long eh = (long long)(fh - (unsigned long long)dh);
    const double f0 = dh * 0x1p-64;
    const double f1 = ((double)eh + (double)fl * 0x1p-64) * 0x1p-64;
    const double s = f0 + f1;
    const double t = f1 - (s - f0);
    const double ph = s * 0x1.921fb54442d18p+0;
    const double pl = __builtin_fma(s, 0x1.921fb54442d18p+0, -ph) + (s * 0x1.1a62633145c07p-54 + t * 0x1.921fb54442d18p+0);
    rh = ph + pl;
    rl = pl - (rh - ph);
    if (up) {
        rh = -rh;
        rl = -rl;
    }
    return n & 3;
}

// sin(a) and cos(a) of the fast form: Cody-Waite reduction a = k pi/2 + r with pi/2 in two parts
// (hi, the double nearest pi/2, and mid, the double nearest pi/2 - hi; the first fused step is exact
// and the neglected third part contributes below |k| 1e-32), the polynomials above, and
// the quadrant's swap and signs.  Within about one ulp of the correctly rounded results for |a| below
// 2^20 pi/2 (~1.6e6); larger finite arguments are reduced by reduce_pio2_large and the reduced
// argument's low part enters to first order (sin(r + rl) = sin r + rl cos r).  Infinities and NaN
// give NaN.  About half the instructions of the libm's sincos, which reduces in double-double
// arithmetic, and no call into it: the libm's reduction of huge arguments, inlined into every step,
// held more registers than the step's own code.
LFDM_HD void sincos_fast(double a, double& s, double& c) {
    double sr, cr;
    int q;
    if (fabs(a) < 0x1.921fb54442d18p+20) {
        const double k = __builtin_rint(a * 0x1.45f306dc9c883p-1);  // a * 2/pi
        double r = __builtin_fma(-k, 0x1.921fb54442d18p+0, a);
        r = __builtin_fma(-k, 0x1.1a62633145c07p-54, r);
        const double z = r * r;
        sr = sin_poly(r, z);
        cr = cos_poly(z);
        q = (int)k;
    } else {
        double rh, rl;
        q = reduce_pio2_large(fabs(a), rh, rl);
        const double z = rh * rh;
        const double s1 = sin_poly(rh, z), c1 = cos_poly(z);
        sr = __builtin_fma(rl, c1, s1);
        cr = __builtin_fma(-rl, s1, c1);
        if (a < 0) {  // -a = q pi/2 + r: a = (-q) pi/2 - r
            sr = -sr;
            q = -q;
        }
        if (!(fabs(a) <= 0x1.fffffffffffffp+1023)) sr = cr = a - a;  // inf, NaN
    }
    const bool swap = (q & 1) != 0;
    const double s0 = swap ? cr : sr, c0 = swap ? sr : cr;
    s = (q & 2) ? -s0 : s0;
    c = ((q + 1) & 2) ? -c0 : c0;
}

// sin(2 pi v) of the fast form, for the sine-step weights, whose angle is 2 pi times a number v in
// [0, 1]: v - rint(v), then +-1/2 - r where |r| > 1/4 (periodicity and sin(pi - a) = sin(a); both
// subtractions are exact), leave t in [-1/4, 1/4], and the Taylor polynomial of sin(2 pi t) to its
// t^21 term (coefficients (-1)^k (2 pi)^(2k+1) / (2k+1)!, rounded to double; truncation below 2e-18
// for |2 pi t| <= pi/2) is evaluated by Horner's rule in t^2, its leading coefficient 2 pi in two parts
// (hi, the double nearest 2 pi, enters the last fused step exactly; lo = 2 pi - hi joins the rest).
// The reference's angle is the rounded product k2Pi * v, so the two differ by that rounding (below
// 1e-15); about a third of the libm's instructions, with no reduction branch.
LFDM_HD double sin_2pi_unit(double v) {
    const double r = v - __builtin_rint(v);
    const double t = fabs(r) > 0.25 ? (r < 0 ? -0.5 : 0.5) - r : r;
    const double z = t * t;
    double p = 0x1.2877020d52cf0p-10;
    p = __builtin_fma(p, z, -0x1.8a404211f9547p-7);
    p = __builtin_fma(p, z, 0x1.aaec32af93359p-4);
    p = __builtin_fma(p, z, -0x1.6fadb9f155744p-1);
    p = __builtin_fma(p, z, 0x1.e8f434d018d63p+1);
    p = __builtin_fma(p, z, -0x1.e3074fde8871fp+3);
    p = __builtin_fma(p, z, 0x1.50783487ee782p+5);
    p = __builtin_fma(p, z, -0x1.32d2cce62bd86p+6);
    p = __builtin_fma(p, z, 0x1.466bc6775aae2p+6);
    p = __builtin_fma(p, z, -0x1.4abbce625be53p+5);
    p = __builtin_fma(p, z, 0x1.1a62633145c07p-52);
    return __builtin_fma(t, 0x1.921fb54442d18p+2, t * p);
}

LFDM_HD double sgn(double v) { return (double)((0.0 < v) - (v < 0.0)); }

LFDM_HD double clampd(double v, double lo, double hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// sineStep (utils.cpp:128-138) with x1, y1, x2, y2 = 1, 1, 1.5, 0: the tire-force weighting.  The
// fast form takes the sine of k2Pi v, v = (x - 1) / dx (exact: x - 1 by Sterbenz's lemma, then a
// doubling), as sin_2pi_unit(v).
template <bool STRICT>
LFDM_HD double sine_step_weight(double x) {
    if (x <= 1.0) return 1.0;
    if (x >= 1.5) return 0.0;
    const double dx = 0.5, dy = -1.0;
    if (STRICT) return 1.0 + dy * (x - 1.0) / dx - (dy / k2Pi) * sin(k2Pi * (x - 1.0) / dx);
    const double v = (x - 1.0) / dx;
    return 1.0 + dy * v - (dy / k2Pi) * sin_2pi_unit(v);
}

// General sineStep (utils.cpp:128-138).
template <bool STRICT>
LFDM_HD double sine_step(double x, double x1, double y1, double x2, double y2) {
    if (x <= x1) return y1;
    if (x >= x2) return y2;
    const double dx = x2 - x1, dy = y2 - y1;
    if (STRICT) return y1 + qdiv(dy * (x - x1), dx) - qdiv(dy, k2Pi) * sin(qdiv(k2Pi * (x - x1), dx));
    // the fast form's sine of 2 pi (x - x1) / dx, with (x - x1) / dx in [0, 1]
    return y1 + fdiv(dy * (x - x1), dx) - fdiv(dy, k2Pi) * sin_2pi_unit(fdiv(x - x1, dx));
}

// Clamped piecewise-linear lookup (utils.cpp:109-125). `m` is wave-uniform, so the knot scan
// runs on scalar loads; the interval index counts knots strictly below x, which is exactly
// std::lower_bound's answer on a sorted map.  Branch-free: the segment the scan picks is
// interpolated for every x (the first segment at or below the first knot, the last at or above the
// last knot; a one-knot map reads the zero entry after it, whose result is never selected) and the
// end values are selected outside the knots, so a wavefront takes no divergent branch and its
// scalar loads are not held behind one.
template <bool STRICT>
LFDM_HD double map_y(const Map& m, double x) {
    const int n = m.n;
    int r = 0;
    for (int k = 1; k < n - 1; ++k) r += (m.e[k].x < x) ? 1 : 0;
    r += 1;
    const MapXY L = m.e[r - 1];
    const MapXY R = m.e[r];
    // a division in both forms: a map through the origin gives exactly 0 at 0 (the powertrain's losses at
    // rest), which a reciprocal width would not
    const double mbar = xdiv<STRICT>(x - L.x, R.x - L.x);
    const double y = L.y + mbar * (R.y - L.y);
    const MapXY F = m.e[0], B = m.e[n - 1];
    return x <= F.x ? F.y : (x >= B.x ? B.y : y);
}

// Same lookup on a map whose abscissae are scaled by sx and ordinates by sy1 then sy2
// (multiplying by 1.0 is exact, so unit factors reproduce the unscaled map bit for bit).
template <bool STRICT>
LFDM_HD double map_y_scaled(const Map& m, double x, double sx, double sy1, double sy2) {
    const int n = m.n;
    int r = 0;
    for (int k = 1; k < n - 1; ++k) r += (m.e[k].x * sx < x) ? 1 : 0;
    r += 1;
    const MapXY L = m.e[r - 1];
    const MapXY R = m.e[r];
    const double lx = L.x * sx, rx = R.x * sx, ly = L.y * sy1 * sy2, ry = R.y * sy1 * sy2;
    const double mbar = xdiv<STRICT>(x - lx, rx - lx);
    const double y = ly + mbar * (ry - ly);
    const MapXY F = m.e[0], B = m.e[n - 1];
    return x <= F.x * sx ? F.y * sy1 * sy2 : (x >= B.x * sx ? B.y * sy1 * sy2 : y);
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
