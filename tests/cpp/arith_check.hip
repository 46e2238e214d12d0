// GPU check of the step code's quotient and square root (qdiv, qsqrt in csrc/dev_common.h) against
// the compiler's IEEE `/` and sqrt on the device and against the host's correctly rounded results, and
// of the fast form's faithful quotient, reciprocal and square root (fdiv, frcp, fsqrt): each within one
// ulp of the correctly rounded result; and of the fast form's sine and cosine (sincos_fast over angles
// up to 1e5 rad and over huge ones, 2^20.7 to 2^1024 rad, which take its Payne-Hanek reduction;
// sin_2pi_unit over [0, 1]): within two ulps of the correctly rounded results (host long double
// sinl / cosl, rounded to double), NaN for infinite and NaN angles; and of fsqrt_rcp (square root and reciprocal square root from
// one v_rsq_f64): each within one ulp.
//
// Operands: random signs, mantissas and exponents over the range the helpers are specified for
// (quotients of operands with exponents in [-450, 450], square roots of x in [2^-700, 2^700]), plus
// operands shaped like the step's (ratios of numbers in [1e-6, 1e6], zero numerators, squares of
// zero).  Prints the mismatch counts; exits non-zero if qdiv or qsqrt differs from the device's `/` or
// sqrt or if fdiv, frcp or fsqrt is more than one ulp from the correctly rounded result (the
// device-versus-host counts, the number of faithful results that are not the correctly rounded one and
// the largest relative error of the hardware estimates v_rcp_f64 and v_rsq_f64 are information).
// Usage: arith_check [n]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../low-fidelity-dynamic-models_amd/csrc/dev_common.h"

using lfdm::dev::fdiv;
using lfdm::dev::frcp;
using lfdm::dev::fsqrt;
using lfdm::dev::qdiv;
using lfdm::dev::qsqrt;

__global__ void arith(const double* a, const double* b, const double* x, int n, double* q_ieee, double* q_fast,
                      double* s_ieee, double* s_fast) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    q_ieee[i] = a[i] / b[i];
    q_fast[i] = qdiv(a[i], b[i]);
    s_ieee[i] = sqrt(x[i]);
    s_fast[i] = qsqrt(x[i]);
}

// out[5 n]: fdiv(a, b), frcp(b), fsqrt(x), v_rcp_f64(b), v_rsq_f64(x)
__global__ void faithful(const double* a, const double* b, const double* x, int n, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = fdiv(a[i], b[i]);
    out[n + i] = frcp(b[i]);
    out[2 * n + i] = fsqrt(x[i]);
    out[3 * n + i] = __builtin_amdgcn_rcp(b[i]);
    out[4 * n + i] = __builtin_amdgcn_rsq(x[i]);
}

// out[5 n]: sincos_fast(a) (sine, cosine), sin_2pi_unit(v), fsqrt_rcp(x) (square root, its reciprocal)
__global__ void trig(const double* a, const double* v, const double* x, int n, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s, c;
    lfdm::dev::sincos_fast(a[i], s, c);
    out[i] = s;
    out[n + i] = c;
    out[2 * n + i] = lfdm::dev::sin_2pi_unit(v[i]);
    double g, rg;
    lfdm::dev::fsqrt_rcp(x[i], g, rg);
    out[3 * n + i] = g;
    out[4 * n + i] = rg;
}

// Distance in ulps between two finite doubles of the same sign (0 and -0 are 0 apart).
static long long ulps(double p, double q) {
    if (p == q) return 0;
    if ((p < 0) != (q < 0)) return 1LL << 62;
    long long ip, iq;
    std::memcpy(&ip, &p, 8);
    std::memcpy(&iq, &q, 8);
    ip &= 0x7fffffffffffffffLL;
    iq &= 0x7fffffffffffffffLL;
    return ip > iq ? ip - iq : iq - ip;
}

static bool same(double p, double q) { return std::memcmp(&p, &q, sizeof p) == 0; }

#define CHECK(call)                                                                  \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));          \
            return 2;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : (1 << 22);
    if (n <= 0) return 2;
    std::mt19937_64 rng(2024);
    std::uniform_real_distribution<double> mant(1.0, 2.0);
    std::uniform_int_distribution<int> e450(-450, 450), e700(-700, 700), sign(0, 1);
    std::uniform_real_distribution<double> phys(-6.0, 6.0);
    std::vector<double> a(n), b(n), x(n);
    for (int i = 0; i < n; ++i) {
        if (i % 4 == 0) {  // step-shaped operands
            a[i] = (sign(rng) ? -1 : 1) * std::pow(10.0, phys(rng));
            b[i] = (sign(rng) ? -1 : 1) * std::pow(10.0, phys(rng));
            const double r = std::pow(10.0, phys(rng)), s = std::pow(10.0, phys(rng));
            x[i] = r * r + s * s;
            if (i % 64 == 0) a[i] = (i % 128 == 0) ? 0.0 : -0.0;
            if (i % 256 == 4) x[i] = 0.0;
        } else {
            a[i] = (sign(rng) ? -1 : 1) * std::ldexp(mant(rng), e450(rng));
            b[i] = (sign(rng) ? -1 : 1) * std::ldexp(mant(rng), e450(rng));
            x[i] = std::ldexp(mant(rng), e700(rng));
        }
    }
    const size_t bytes = sizeof(double) * (size_t)n;
    double *da, *db, *dx, *q1, *q2, *s1, *s2;
    CHECK(hipMalloc(&da, bytes));
    CHECK(hipMalloc(&db, bytes));
    CHECK(hipMalloc(&dx, bytes));
    CHECK(hipMalloc(&q1, bytes));
    CHECK(hipMalloc(&q2, bytes));
    CHECK(hipMalloc(&s1, bytes));
    CHECK(hipMalloc(&s2, bytes));
    CHECK(hipMemcpy(da, a.data(), bytes, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(db, b.data(), bytes, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dx, x.data(), bytes, hipMemcpyHostToDevice));
    const int block = 256, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(arith, dim3(grid), dim3(block), 0, 0, da, db, dx, n, q1, q2, s1, s2);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    double* fo;
    CHECK(hipMalloc(&fo, 5 * bytes));
    hipLaunchKernelGGL(faithful, dim3(grid), dim3(block), 0, 0, da, db, dx, n, fo);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    std::vector<double> hf(5 * (size_t)n);
    CHECK(hipMemcpy(hf.data(), fo, 5 * bytes, hipMemcpyDeviceToHost));
    CHECK(hipFree(fo));
    // angles: steering-sized (|a| <= 0.6), a few turns (|a| <= 20), yaw after long runs (|a| <= 1e5),
    // huge (|a| = 2^U(20.7, 1023.99), the large-argument reduction); the last three are inf, -inf, NaN;
    // sine-step arguments v in [0, 1]
    std::vector<double> ang(n), unit(n);
    std::uniform_real_distribution<double> u06(-0.6, 0.6), u20(-20.0, 20.0), u1e5(-1e5, 1e5), u01(0.0, 1.0),
        uexp(20.7, 1023.99);
    for (int i = 0; i < n; ++i) {
        const int c = i % 4;
        ang[i] = c == 0 ? u06(rng) : c == 1 ? u20(rng) : c == 2 ? u1e5(rng) : (i & 4 ? -1 : 1) * std::exp2(uexp(rng));
        unit[i] = i % 1024 == 0 ? (double)(i / 1024 % 5) * 0.25 : u01(rng);  // include 0, 1/4, 1/2, 3/4, 1
    }
    ang[n - 3] = HUGE_VAL;
    ang[n - 2] = -HUGE_VAL;
    ang[n - 1] = NAN;
    double *dang, *dunit, *to;
    CHECK(hipMalloc(&dang, bytes));
    CHECK(hipMalloc(&dunit, bytes));
    CHECK(hipMalloc(&to, 5 * bytes));
    CHECK(hipMemcpy(dang, ang.data(), bytes, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dunit, unit.data(), bytes, hipMemcpyHostToDevice));
    double* dxs;
    CHECK(hipMalloc(&dxs, bytes));
    CHECK(hipMemcpy(dxs, x.data(), bytes, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(trig, dim3(grid), dim3(block), 0, 0, dang, dunit, dxs, n, to);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    std::vector<double> ht(5 * (size_t)n);
    CHECK(hipMemcpy(ht.data(), to, 5 * bytes, hipMemcpyDeviceToHost));
    for (double* p : {dang, dunit, to, dxs}) CHECK(hipFree(p));
    std::vector<double> hq1(n), hq2(n), hs1(n), hs2(n);
    CHECK(hipMemcpy(hq1.data(), q1, bytes, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(hq2.data(), q2, bytes, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(hs1.data(), s1, bytes, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(hs2.data(), s2, bytes, hipMemcpyDeviceToHost));
    for (double* p : {da, db, dx, q1, q2, s1, s2}) CHECK(hipFree(p));

    long qd = 0, qh = 0, sd = 0, sh = 0, zsign = 0;
    for (int i = 0; i < n; ++i) {
        const double q = a[i] / b[i], s = std::sqrt(x[i]);
        if (!same(hq2[i], hq1[i])) {
            // a zero numerator: the fast path returns +0 where IEEE division keeps the sign
            if (a[i] == 0.0 && hq2[i] == 0.0 && hq1[i] == 0.0) ++zsign;
            else if (qd++ < 5) std::printf("qdiv(%a, %a) = %a, '/' gives %a\n", a[i], b[i], hq2[i], hq1[i]);
        }
        if (!same(hq1[i], q)) ++qh;
        if (!same(hs2[i], hs1[i]) && sd++ < 5) std::printf("qsqrt(%a) = %a, sqrt gives %a\n", x[i], hs2[i], hs1[i]);
        if (!same(hs1[i], s)) ++sh;
    }
    std::printf("n = %d\n", n);
    std::printf("qdiv vs device '/': %ld mismatches (+ %ld signed-zero quotients of zero numerators)\n", qd, zsign);
    std::printf("device '/' vs host: %ld mismatches\n", qh);
    std::printf("qsqrt vs device sqrt: %ld mismatches\n", sd);
    std::printf("device sqrt vs host: %ld mismatches\n", sh);
    // faithful forms: within one ulp of the correctly rounded host result
    long long fmax_[3] = {0, 0, 0}, fnotcr[3] = {0, 0, 0};
    double rcp_err = 0, rsq_err = 0;
    for (int i = 0; i < n; ++i) {
        const double ref[3] = {a[i] / b[i], 1.0 / b[i], std::sqrt(x[i])};
        for (int k = 0; k < 3; ++k) {
            const long long u = ulps(hf[(size_t)k * n + i], ref[k]);
            if (u > fmax_[k]) {
                if (u > 1 && fmax_[k] <= 1)
                    std::printf("%s(%a, %a) = %a, correctly rounded %a\n", k == 0 ? "fdiv" : k == 1 ? "frcp" : "fsqrt",
                                k == 2 ? x[i] : a[i], b[i], hf[(size_t)k * n + i], ref[k]);
                fmax_[k] = u;
            }
            if (u) ++fnotcr[k];
        }
        rcp_err = std::max(rcp_err, std::fabs(hf[3 * (size_t)n + i] * b[i] - 1.0));
        if (x[i] > 0) {
            const double y = hf[4 * (size_t)n + i];
            rsq_err = std::max(rsq_err, std::fabs(y * std::sqrt(x[i]) - 1.0));
        }
    }
    std::printf("fdiv: at most %lld ulp from the correctly rounded quotient; %lld not correctly rounded\n", fmax_[0],
                fnotcr[0]);
    std::printf("frcp: at most %lld ulp from the correctly rounded reciprocal; %lld not correctly rounded\n", fmax_[1],
                fnotcr[1]);
    std::printf("fsqrt: at most %lld ulp from the correctly rounded square root; %lld not correctly rounded\n",
                fmax_[2], fnotcr[2]);
    std::printf("hardware estimates: v_rcp_f64 relative error <= %.3g (2^%.1f), v_rsq_f64 <= %.3g (2^%.1f)\n", rcp_err,
                std::log2(rcp_err), rsq_err, std::log2(rsq_err));
    // fast sine and cosine against the host's long double results rounded to double
    const long double two_pi = 6.283185307179586476925286766559005768L;
    long long tmax[3] = {0, 0, 0};
    long long nonfinite_bad = 0;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(ang[i])) {
            nonfinite_bad += !std::isnan(ht[i]) || !std::isnan(ht[(size_t)n + i]);
            continue;
        }
        // sin(2 pi v) = sin(2 pi t) for the exactly reduced t of sin_2pi_unit (v - rint(v), then +-1/2 - r),
        // whose long double product with 2 pi has no cancellation near the zeros
        const double r = unit[i] - std::rint(unit[i]);
        const double t = std::fabs(r) > 0.25 ? (r < 0 ? -0.5 : 0.5) - r : r;
        const double ref[3] = {(double)sinl((long double)ang[i]), (double)cosl((long double)ang[i]),
                               (double)sinl(two_pi * (long double)t)};
        for (int k = 0; k < 3; ++k) {
            const long long u = ulps(ht[(size_t)k * n + i], ref[k]);
            if (u > tmax[k]) {
                if (u > 2 && tmax[k] <= 2)
                    std::printf("%s(%a) = %a, correctly rounded %a\n", k == 0 ? "sin" : k == 1 ? "cos" : "sin_2pi_unit",
                                k == 2 ? unit[i] : ang[i], ht[(size_t)k * n + i], ref[k]);
                tmax[k] = u;
            }
        }
    }
    std::printf("sincos_fast: sine at most %lld ulp, cosine at most %lld ulp from the correctly rounded values "
                "(angles up to 2^1024); %lld of inf, -inf, NaN not NaN\n",
                tmax[0], tmax[1], nonfinite_bad);
    std::printf("sin_2pi_unit: at most %lld ulp from the correctly rounded value\n", tmax[2]);
    // sqrt and reciprocal square root of fsqrt_rcp (the reciprocal for x > 0; long double reference)
    long long smax[2] = {0, 0};
    for (int i = 0; i < n; ++i) {
        const double g = ht[3 * (size_t)n + i], rg = ht[4 * (size_t)n + i];
        const long long us = ulps(g, std::sqrt(x[i]));
        const long long ur = x[i] > 0 ? ulps(rg, (double)(1.0L / sqrtl((long double)x[i]))) : 0;
        if ((us > 1 && smax[0] <= 1) || (ur > 1 && smax[1] <= 1))
            std::printf("fsqrt_rcp(%a) = %a, %a; correctly rounded %a, %a\n", x[i], g, rg, std::sqrt(x[i]),
                        (double)(1.0L / sqrtl((long double)x[i])));
        smax[0] = std::max(smax[0], us);
        smax[1] = std::max(smax[1], ur);
    }
    std::printf("fsqrt_rcp: square root at most %lld ulp, reciprocal at most %lld ulp from the correctly rounded values\n",
                smax[0], smax[1]);
    const bool ok = qd == 0 && sd == 0 && nonfinite_bad == 0 && fmax_[0] <= 1 && fmax_[1] <= 1 && fmax_[2] <= 1 && tmax[0] <= 2 &&
                    tmax[1] <= 2 && tmax[2] <= 2 && smax[0] <= 1 && smax[1] <= 1;
    std::printf("%s\n", ok ? "ALL OK" : "MISMATCH");
    return ok ? 0 : 1;
}
