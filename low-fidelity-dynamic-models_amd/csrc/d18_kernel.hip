// 18-DOF half-implicit integration kernels (gfx950).
//
// One vehicle per lane, 64 vehicles per wavefront, all belonging to one parameter set.  A launch
// advances every vehicle `nsteps` steps with its carried state in registers; state is read from
// and written to the SoA record array [field][slot] once per launch.  Output rows are written
// decimated as [sample][13][vehicle] like the reference's device response.
#include <hip/hip_runtime.h>

#include "d18_device.h"
#include "kernels.h"

namespace lfdm {

using namespace dev;

namespace {

// Record field offsets (the canonical 18-DOF record, see kernels.h).
template <bool NR>
struct Rec18 {
    static constexpr int kTire = NR ? 11 : 15;
    static constexpr int kVeh = 4 * kTire;
    static constexpr int kLen = kVeh + 15;
    // per-tire field offsets
    static constexpr int xe = 0, ye = 1, xedot = 2, yedot = 3;
    static constexpr int omega = NR ? 0 : 4;
    static constexpr int dOmega = omega + 1, xt = omega + 2, rStat = omega + 3, fx = omega + 4, fy = omega + 5,
                         fz = omega + 6, vsx = omega + 7, vsy = omega + 8, My = omega + 9, engTor = omega + 10;
    static __device__ __forceinline__ int tire(int i, int f) { return i * kTire + f; }
};

struct Ctl {
    double s, th, b;
};

template <int MODE>
struct CtrlState {
    int kcur = -1;
    double s0 = 0, s1 = 0, thr = 0, brk = 0;
};

// Per-vehicle control buffers are indexed by user vehicle index `uv` with stride n_user.
template <int MODE>
__device__ __forceinline__ Ctl controls_at(const Launch18& L, const P18& P, CtrlState<MODE>& cs, int uv, int k,
                                           double t) {
    Ctl c;
    if (MODE == kCtrlConst) {
        c.s = L.c_s; c.th = L.c_t; c.b = L.c_b;
    } else if (MODE == kCtrlTable) {
        driver_input(t, L.drv + P.drv_off, P.drv_rows, c.s, c.th, c.b);
    } else if (MODE == kCtrlSteps) {
        const size_t N = (size_t)L.n_user;
        const double* p = L.steps + (size_t)k * 3 * N + uv;
        c.s = p[0];
        c.th = p[N];
        c.b = p[2 * N];
    } else {  // kCtrlKnots
        const size_t N = (size_t)L.n_user;
        int ks = (int)floor(t / L.ds);
        ks = ks < 0 ? 0 : (ks > L.ns - 2 ? L.ns - 2 : ks);
        int kp = (int)floor(t / L.dp);
        kp = kp < 0 ? 0 : (kp > L.np - 1 ? L.np - 1 : kp);
        const int key = ks * 4096 + kp;  // wave-uniform: t is the same on every lane
        if (key != cs.kcur) {
            cs.kcur = key;
            cs.s0 = L.knots[(size_t)ks * N + uv];
            cs.s1 = L.knots[(size_t)(ks + 1) * N + uv];
            cs.thr = L.knots[(size_t)(L.ns + kp) * N + uv];
            cs.brk = L.knots[(size_t)(L.ns + L.np + kp) * N + uv];
        }
        const double tk = ks * L.ds;
        double tbar = (t - tk) / L.ds;
        if (tbar > 1.0) tbar = 1.0;
        c.s = cs.s0 + tbar * (cs.s1 - cs.s0);
        c.th = cs.thr;
        c.b = cs.brk;
    }
    return c;
}

// Occupancy target of the integration kernels (waves per SIMD); 0 leaves it to the compiler.
// The step needs more registers than two waves per SIMD leave (the compiler spills some to
// scratch), yet two resident waves hide more FP64 latency than the spills cost: 5.0e9 versus
// 3.8e9 vehicle-steps/s at one wave per SIMD on the 1M-vehicle fleet.
#ifndef LFDM_WAVES_PER_EU
#define LFDM_WAVES_PER_EU 2
#endif
#if LFDM_WAVES_PER_EU > 0
#define LFDM_OCCUPANCY __attribute__((amdgpu_waves_per_eu(LFDM_WAVES_PER_EU, LFDM_WAVES_PER_EU)))
#else
#define LFDM_OCCUPANCY
#endif

template <bool NR, int MODE>
__global__ __launch_bounds__(256) LFDM_OCCUPANCY void d18_integrate(Launch18 L) {
    using R = Rec18<NR>;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = uniform(slot >> 6);
    if (wave * kWave >= L.S) return;
    const int set = uniform(L.wave_set[wave]);
    const P18& P = L.sets[set];
    const size_t S = (size_t)L.S;
    const double* st = L.state;

    S18 s;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!NR) {
            s.xe[i] = st[R::tire(i, R::xe) * S + slot];
            s.ye[i] = st[R::tire(i, R::ye) * S + slot];
            s.xedot[i] = st[R::tire(i, R::xedot) * S + slot];
            s.yedot[i] = st[R::tire(i, R::yedot) * S + slot];
        } else {
            s.xe[i] = s.ye[i] = s.xedot[i] = s.yedot[i] = 0.0;
        }
        s.omega[i] = st[R::tire(i, R::omega) * S + slot];
    }
    const double* v = st + (size_t)R::kVeh * S + slot;
    s.x = v[0 * S]; s.y = v[1 * S]; s.u = v[2 * S]; s.v = v[3 * S]; s.psi = v[4 * S]; s.wz = v[5 * S];
    s.phi = v[6 * S]; s.wx = v[7 * S]; s.udot = v[8 * S]; s.vdot = v[9 * S]; s.crank = v[12 * S];
    s.gear = (int)v[14 * S];

    const D18& u = P.d;
    K18 k;
    {
        const double* o = L.ovr + slot;
        k18_init(P, o[0], o[S], o[2 * S], o[3 * S], o[4 * S], o[5 * S], o[6 * S], k);
    }
    const int user = L.slot_user[slot];
    const int uv = user >= 0 ? user : 0;  // padding lanes read vehicle 0's controls
    CtrlState<MODE> cs;
    double t = L.t0;
    const double dt = L.dt;
    int oi = 0;
    int next_out = L.n_out > 0 ? L.out_steps[0] : -1;
    auto emit = [&](int done) {
        if (done != next_out) return;
        if (user >= 0) {
            const size_t nu = (size_t)L.n_user;
            double* o = L.out + (size_t)oi * 13 * nu + user;
            o[0 * nu] = t; o[1 * nu] = s.x; o[2 * nu] = s.y; o[3 * nu] = s.u; o[4 * nu] = s.v;
            o[5 * nu] = s.phi; o[6 * nu] = s.psi; o[7 * nu] = s.wx; o[8 * nu] = s.wz;
            o[9 * nu] = s.omega[0]; o[10 * nu] = s.omega[1]; o[11 * nu] = s.omega[2]; o[12 * nu] = s.omega[3];
        }
        ++oi;
        next_out = oi < L.n_out ? L.out_steps[oi] : -1;
    };
    // all steps but the last keep only the carried state ...
    const int last = L.nsteps - 1;
    for (int kk = 0; kk < last; ++kk) {
        const Ctl c = controls_at<MODE>(L, P, cs, uv, kk, t);
        step18<NR, false>(P, u, k, s, c.s, c.th, c.b, dt, nullptr);
        t = t + dt;
        emit(kk + 1);
    }
    // ... the last one also returns the right-hand-side quantities the full record holds
    X18 x;
    {
        const Ctl c = controls_at<MODE>(L, P, cs, uv, last, t);
        step18<NR, true>(P, u, k, s, c.s, c.th, c.b, dt, &x);
        t = t + dt;
        emit(last + 1);
    }

    double* w = L.state;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!NR) {
            w[R::tire(i, R::xe) * S + slot] = s.xe[i];
            w[R::tire(i, R::ye) * S + slot] = s.ye[i];
            w[R::tire(i, R::xedot) * S + slot] = s.xedot[i];
            w[R::tire(i, R::yedot) * S + slot] = s.yedot[i];
        }
        w[R::tire(i, R::omega) * S + slot] = s.omega[i];
        {
            w[R::tire(i, R::dOmega) * S + slot] = x.dOmega[i];
            w[R::tire(i, R::xt) * S + slot] = x.xt[i];
            w[R::tire(i, R::rStat) * S + slot] = x.rStat[i];
            w[R::tire(i, R::fx) * S + slot] = x.fx[i];
            w[R::tire(i, R::fy) * S + slot] = x.fy[i];
            w[R::tire(i, R::fz) * S + slot] = x.fz[i];
            w[R::tire(i, R::vsx) * S + slot] = x.vsx[i];
            w[R::tire(i, R::vsy) * S + slot] = x.vsy[i];
            w[R::tire(i, R::My) * S + slot] = x.My[i];
            w[R::tire(i, R::engTor) * S + slot] = x.engTor[i];
        }
    }
    double* vw = w + (size_t)R::kVeh * S + slot;
    vw[0 * S] = s.x; vw[1 * S] = s.y; vw[2 * S] = s.u; vw[3 * S] = s.v; vw[4 * S] = s.psi; vw[5 * S] = s.wz;
    vw[6 * S] = s.phi; vw[7 * S] = s.wx; vw[8 * S] = s.udot; vw[9 * S] = s.vdot; vw[12 * S] = s.crank;
    vw[14 * S] = (double)s.gear;
    vw[10 * S] = x.wxdot;
    vw[11 * S] = x.wzdot;
    if (P.tcbool) vw[13 * S] = x.dcrank;
}

template <bool NR>
hipError_t launch_mode(const Launch18& L, int block, hipStream_t stream) {
    const int grid = (L.S + block - 1) / block;
    switch (L.mode) {
        case kCtrlTable: hipLaunchKernelGGL((d18_integrate<NR, kCtrlTable>), grid, block, 0, stream, L); break;
        case kCtrlConst: hipLaunchKernelGGL((d18_integrate<NR, kCtrlConst>), grid, block, 0, stream, L); break;
        case kCtrlKnots: hipLaunchKernelGGL((d18_integrate<NR, kCtrlKnots>), grid, block, 0, stream, L); break;
        case kCtrlSteps: hipLaunchKernelGGL((d18_integrate<NR, kCtrlSteps>), grid, block, 0, stream, L); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_d18(const Launch18& L, int block, hipStream_t stream) {
    if (L.S <= 0 || L.nsteps <= 0) return hipSuccess;
    return L.nr ? launch_mode<true>(L, block, stream) : launch_mode<false>(L, block, stream);
}

}  // namespace lfdm
