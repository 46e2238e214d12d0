// Half-implicit integration kernels of the 11-, 18- and 24-DOF models (gfx950), shared by the
// per-model kernel files d11_kernel.hip, d18_kernel.hip and d24_kernel.hip.
//
// One vehicle per lane, 64 vehicles per wavefront, all belonging to one parameter set.  A launch
// advances every vehicle `nsteps` steps with its carried state in registers; state is read from
// and written to the SoA record array [field][slot] once per launch.  Output rows are written
// decimated as [sample][channel][vehicle] like the reference's device response (9 channels for
// 11-DOF, 13 for 18-DOF, 15 for 24-DOF).
#pragma once
#include <hip/hip_runtime.h>

#include "d11_device.h"
#include "d18_device.h"
#include "d24_device.h"
#include "kernels.h"

namespace lfdm {

using namespace dev;

namespace kern {

struct Ctl {
    double s, th, b;
};

// Controls of kCtrlFunctor launches come from a device functor F with
//   __device__ void eval(double t, int vehicle, double& steering, double& throttle, double& braking);
// the library's own launches (every other control source) carry this empty one.
struct NoControls {
    __device__ void eval(double, int, double&, double&, double&) {}
};

template <int MODE>
struct CtrlState {
    int kcur = -1;
    double s0 = 0, s1 = 0, thr = 0, brk = 0;
};

// Per-vehicle control buffers are indexed by user vehicle index `uv` with stride n_user.
template <int MODE, class F>
__device__ __forceinline__ Ctl controls_at(const Launch& L, const PSet& P, CtrlState<MODE>& cs, int uv, int k,
                                           double t, F& f) {
    Ctl c;
    if constexpr (MODE == kCtrlFunctor) {
        f.eval(t, uv, c.s, c.th, c.b);
    } else if (MODE == kCtrlConst) {
        c.s = L.c_s; c.th = L.c_t; c.b = L.c_b;
    } else if (MODE == kCtrlTable) {
        driver_input(t, L.drv + P.drv_off, P.drv_rows, c.s, c.th, c.b);
    } else if (MODE == kCtrlSteps) {
        if (L.steps_shared) {  // wave-uniform: one row per step for the whole fleet
            const double* p = L.steps + (size_t)k * 3;
            c.s = p[0];
            c.th = p[1];
            c.b = p[2];
        } else {
            const size_t N = (size_t)L.n_user;
            const double* p = L.steps + (size_t)k * 3 * N + uv;
            c.s = p[0];
            c.th = p[N];
            c.b = p[2 * N];
        }
    } else {  // kCtrlKnots
        const size_t N = (size_t)L.n_user;
        int ks = (int)floor(qdiv(t, L.ds));
        ks = ks < 0 ? 0 : (ks > L.ns - 2 ? L.ns - 2 : ks);
        int kp = (int)floor(qdiv(t, L.dp));
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
        double tbar = qdiv(t - tk, L.ds);
        if (tbar > 1.0) tbar = 1.0;
        c.s = cs.s0 + tbar * (cs.s1 - cs.s0);
        c.th = cs.thr;
        c.b = cs.brk;
    }
    return c;
}

// ---- model traits: record layout, state load/store, output channels ----
// (load, kinit, step and store also compile for the host, where the host checks run the step code)

// 18-DOF record: per tire LF RF LR RR (TMeasy: xe ye xedot yedot) omega dOmega xt rStat fx fy fz vsx
// vsy My engTor; vehicle x y u v psi wz phi wx udot vdot wxdot wzdot crankOmega dOmega_crank gear.
struct M18 {
    using S = S18;
    using X = X18;
    using K = K18;
    static constexpr int kCh = 13;
    static constexpr int kWaves = kMaxWaves18;
    static constexpr bool kKInLds = true;
    template <bool NR>
    struct R {
        static constexpr int kTire = NR ? 11 : 15, kVeh = 4 * kTire, om = NR ? 0 : 4;
        static LFDM_HD int f(int i, int o) { return i * kTire + om + o; }  // o: omega = 0 ...
    };
    template <bool NR>
    static LFDM_HD void load(const double* st, size_t S, int slot, S18& s, K18&) {
        using Rr = R<NR>;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!NR) {
                s.xe[i] = st[(i * Rr::kTire + 0) * S + slot];
                s.ye[i] = st[(i * Rr::kTire + 1) * S + slot];
                s.xedot[i] = st[(i * Rr::kTire + 2) * S + slot];
                s.yedot[i] = st[(i * Rr::kTire + 3) * S + slot];
            } else {
                s.xe[i] = s.ye[i] = s.xedot[i] = s.yedot[i] = 0.0;
            }
            s.omega[i] = st[Rr::f(i, 0) * S + slot];
        }
        const double* v = st + (size_t)Rr::kVeh * S + slot;
        s.x = v[0 * S]; s.y = v[1 * S]; s.u = v[2 * S]; s.v = v[3 * S]; s.psi = v[4 * S]; s.wz = v[5 * S];
        s.phi = v[6 * S]; s.wx = v[7 * S]; s.udot = v[8 * S]; s.vdot = v[9 * S]; s.crank = v[12 * S];
        s.gear = (int)v[14 * S];
    }
    static LFDM_HD void kinit(const PSet& P, const double* o, size_t S, K18& k) {
        k18_init(P, o[0], o[S], o[2 * S], o[3 * S], o[4 * S], o[5 * S], o[6 * S], k);
    }
    template <bool NR, bool FULL, bool STRICT, bool POS = true>
    static LFDM_HD void step(const PSet& P, const Derived& u, const K18& k, S18& s, const Ctl& c,
                                                double dt, X18* x) {
        step18<NR, FULL, false, STRICT, POS>(P, u, k, s, c.s, c.th, c.b, dt, x);
    }
    template <bool STRICT>
    static LFDM_HD void position(S18& s, double dt) { position18<STRICT>(s, dt); }
    static __device__ __forceinline__ void emit(double* o, size_t nu, double t, const S18& s) {
        o[0 * nu] = t; o[1 * nu] = s.x; o[2 * nu] = s.y; o[3 * nu] = s.u; o[4 * nu] = s.v;
        o[5 * nu] = s.phi; o[6 * nu] = s.psi; o[7 * nu] = s.wx; o[8 * nu] = s.wz;
        o[9 * nu] = s.omega[0]; o[10 * nu] = s.omega[1]; o[11 * nu] = s.omega[2]; o[12 * nu] = s.omega[3];
    }
    template <bool NR>
    static LFDM_HD void store(double* w, size_t S, int slot, const PSet& P, const S18& s, const K18&,
                                                 const X18& x) {
        using Rr = R<NR>;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!NR) {
                w[(i * Rr::kTire + 0) * S + slot] = s.xe[i];
                w[(i * Rr::kTire + 1) * S + slot] = s.ye[i];
                w[(i * Rr::kTire + 2) * S + slot] = s.xedot[i];
                w[(i * Rr::kTire + 3) * S + slot] = s.yedot[i];
            }
            const double f[11] = {s.omega[i], x.dOmega[i], x.xt[i], x.rStat[i], x.fx[i],    x.fy[i],
                                  x.fz[i],    x.vsx[i],    x.vsy[i], x.My[i],   x.engTor[i]};
#pragma unroll
            for (int j = 0; j < 11; ++j) w[Rr::f(i, j) * S + slot] = f[j];
        }
        double* v = w + (size_t)Rr::kVeh * S + slot;
        v[0 * S] = s.x; v[1 * S] = s.y; v[2 * S] = s.u; v[3 * S] = s.v; v[4 * S] = s.psi; v[5 * S] = s.wz;
        v[6 * S] = s.phi; v[7 * S] = s.wx; v[8 * S] = s.udot; v[9 * S] = s.vdot; v[10 * S] = x.wxdot;
        v[11 * S] = x.wzdot; v[12 * S] = s.crank;
        if (P.tcbool) v[13 * S] = x.dcrank;
        v[14 * S] = (double)s.gear;
    }
};

// 11-DOF record: per tire F R (TMeasy: xe ye xedot yedot) omega dOmega xt rStat fx fy fz vsx vsy My
// engTor; vehicle x y u v psi wz udot vdot wzdot crankOmega dOmega_crank gear.
struct M11 {
    using S = S11;
    using X = X11;
    using K = K11;
    static constexpr int kCh = 9;
    static constexpr int kWaves = kMaxWaves11;
    static constexpr bool kKInLds = true;
    template <bool NR>
    struct R {
        static constexpr int kTire = NR ? 11 : 15, kVeh = 2 * kTire, om = NR ? 0 : 4;
        static LFDM_HD int f(int i, int o) { return i * kTire + om + o; }
    };
    template <bool NR>
    static LFDM_HD void load(const double* st, size_t S, int slot, S11& s, K11&) {
        using Rr = R<NR>;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (!NR) {
                s.xe[i] = st[(i * Rr::kTire + 0) * S + slot];
                s.ye[i] = st[(i * Rr::kTire + 1) * S + slot];
                s.xedot[i] = st[(i * Rr::kTire + 2) * S + slot];
                s.yedot[i] = st[(i * Rr::kTire + 3) * S + slot];
            } else {
                s.xe[i] = s.ye[i] = s.xedot[i] = s.yedot[i] = 0.0;
            }
            s.omega[i] = st[Rr::f(i, 0) * S + slot];
        }
        const double* v = st + (size_t)Rr::kVeh * S + slot;
        s.x = v[0 * S]; s.y = v[1 * S]; s.u = v[2 * S]; s.v = v[3 * S]; s.psi = v[4 * S]; s.wz = v[5 * S];
        s.udot = v[6 * S]; s.crank = v[9 * S];
        s.gear = (int)v[11 * S];
    }
    static LFDM_HD void kinit(const PSet& P, const double* o, size_t S, K11& k) {
        k11_init(P, o[0], o[S], o[3 * S], o[4 * S], o[5 * S], o[6 * S], k);
    }
    template <bool NR, bool FULL, bool STRICT, bool POS = true>
    static LFDM_HD void step(const PSet& P, const Derived& u, const K11& k, S11& s, const Ctl& c,
                                                double dt, X11* x) {
        step11<NR, FULL, STRICT, POS>(P, u, k, s, c.s, c.th, c.b, dt, x);
    }
    template <bool STRICT>
    static LFDM_HD void position(S11& s, double dt) { position11<STRICT>(s, dt); }
    static __device__ __forceinline__ void emit(double* o, size_t nu, double t, const S11& s) {
        o[0 * nu] = t; o[1 * nu] = s.x; o[2 * nu] = s.y; o[3 * nu] = s.u; o[4 * nu] = s.v;
        o[5 * nu] = s.psi; o[6 * nu] = s.wz; o[7 * nu] = s.omega[0]; o[8 * nu] = s.omega[1];
    }
    template <bool NR>
    static LFDM_HD void store(double* w, size_t S, int slot, const PSet& P, const S11& s, const K11&,
                                                 const X11& x) {
        using Rr = R<NR>;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (!NR) {
                w[(i * Rr::kTire + 0) * S + slot] = s.xe[i];
                w[(i * Rr::kTire + 1) * S + slot] = s.ye[i];
                w[(i * Rr::kTire + 2) * S + slot] = s.xedot[i];
                w[(i * Rr::kTire + 3) * S + slot] = s.yedot[i];
            }
            const double f[11] = {s.omega[i], x.dOmega[i], x.xt[i], x.rStat[i], x.fx[i],    x.fy[i],
                                  x.fz[i],    x.vsx[i],    x.vsy[i], x.My[i],   x.engTor[i]};
#pragma unroll
            for (int j = 0; j < 11; ++j) w[Rr::f(i, j) * S + slot] = f[j];
        }
        double* v = w + (size_t)Rr::kVeh * S + slot;
        v[0 * S] = s.x; v[1 * S] = s.y; v[2 * S] = s.u; v[3 * S] = s.v; v[4 * S] = s.psi; v[5 * S] = s.wz;
        v[6 * S] = s.udot; v[7 * S] = x.vdot; v[8 * S] = x.wzdot; v[9 * S] = s.crank;
        if (P.tcbool) v[10 * S] = x.dcrank;
        v[11 * S] = (double)s.gear;
    }
};

// 24-DOF record: per tire LF RF LR RR (TMeasy: xe ye xedot yedot) omega dOmega xt dxt rStat fx fy fz
// fxg fyg fzg vgx vgy vgz vsx vsy vsz My engTor; per strut xs xsi dxs ls lsi us vs ws dus dvs dws uu
// vu wu duu dvu dwu fxs fys fzs fzd mx my mz; vehicle x y z u v w psi wz phi wx theta wy udot vdot
// wdot wxdot wydot wzdot dtheta dphi dpsi crankOmega dOmega_crank gear.
struct M24 {
    using S = S24;
    using X = X24;
    using K = K24;
    static constexpr int kCh = 15;
    static constexpr int kWaves = kMaxWaves24;
    static constexpr bool kKInLds = true;
    template <bool NR>
    struct R {
        static constexpr int kTire = NR ? 19 : 23, kSus = 4 * kTire, kVeh = kSus + 4 * 24, om = NR ? 0 : 4;
        static LFDM_HD int f(int i, int o) { return i * kTire + om + o; }
        static LFDM_HD int q(int i, int o) { return kSus + i * 24 + o; }
    };
    // tire fields after omega: dOmega 1, xt 2, dxt 3, rStat 4, fx 5, fy 6, fz 7, fxg 8, fyg 9, fzg 10,
    // vgx 11, vgy 12, vgz 13, vsx 14, vsy 15, vsz 16, My 17, engTor 18
    template <bool NR>
    static LFDM_HD void load(const double* st, size_t S, int slot, S24& s, K24& k) {
        using Rr = R<NR>;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!NR) {
                s.xe[i] = st[(i * Rr::kTire + 0) * S + slot];
                s.ye[i] = st[(i * Rr::kTire + 1) * S + slot];
                k.xedot[i] = st[(i * Rr::kTire + 2) * S + slot];
                k.yedot[i] = st[(i * Rr::kTire + 3) * S + slot];
            } else {
                s.xe[i] = s.ye[i] = k.xedot[i] = k.yedot[i] = 0.0;
            }
            s.omega[i] = st[Rr::f(i, 0) * S + slot];
            s.xt[i] = st[Rr::f(i, 2) * S + slot];
            s.rStat[i] = st[Rr::f(i, 4) * S + slot];
            k.vsz[i] = st[Rr::f(i, 16) * S + slot];
            s.xs[i] = st[Rr::q(i, 0) * S + slot];
            k.xsi[i] = st[Rr::q(i, 1) * S + slot];
            s.dxs[i] = st[Rr::q(i, 2) * S + slot];
            k.lsi[i] = st[Rr::q(i, 4) * S + slot];
            s.wu[i] = st[Rr::q(i, 13) * S + slot];
            s.fzd[i] = st[Rr::q(i, 20) * S + slot];
        }
        const double* v = st + (size_t)Rr::kVeh * S + slot;
        s.x = v[0 * S]; s.y = v[1 * S]; s.u = v[3 * S]; s.v = v[4 * S]; s.w = v[5 * S]; s.psi = v[6 * S];
        s.wz = v[7 * S]; s.phi = v[8 * S]; s.wx = v[9 * S]; s.theta = v[10 * S]; s.wy = v[11 * S];
        s.udot = v[12 * S]; s.vdot = v[13 * S]; s.wdot = v[14 * S]; s.wxdot = v[15 * S]; s.wydot = v[16 * S];
        s.wzdot = v[17 * S]; s.crank = v[21 * S];
        s.gear = (int)v[23 * S];
    }
    static LFDM_HD void kinit(const PSet& P, const double* o, size_t S, K24& k) {
        k24_init(P, o[0], o[S], o[2 * S], o[3 * S], o[4 * S], o[5 * S], o[6 * S], k);
    }
    template <bool NR, bool FULL, bool STRICT, bool POS = true>
    static LFDM_HD void step(const PSet& P, const Derived& u, K24& k, S24& s, const Ctl& c,
                                                double dt, X24* x) {
        step24<NR, FULL, STRICT, POS>(P, u, k, s, c.s, c.th, c.b, dt, x);
    }
    template <bool STRICT>
    static LFDM_HD void position(S24& s, double dt) { position24<STRICT>(s, dt); }
    static __device__ __forceinline__ void emit(double* o, size_t nu, double t, const S24& s) {
        o[0 * nu] = t; o[1 * nu] = s.x; o[2 * nu] = s.y; o[3 * nu] = s.u; o[4 * nu] = s.v;
        o[5 * nu] = s.phi; o[6 * nu] = s.psi; o[7 * nu] = s.theta; o[8 * nu] = s.wx; o[9 * nu] = s.wz;
        o[10 * nu] = s.wy; o[11 * nu] = s.omega[0]; o[12 * nu] = s.omega[1]; o[13 * nu] = s.omega[2];
        o[14 * nu] = s.omega[3];
    }
    template <bool NR>
    static LFDM_HD void store(double* w, size_t S, int slot, const PSet& P, const S24& s, const K24& k,
                                                 const X24& x) {
        using Rr = R<NR>;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!NR) {
                w[(i * Rr::kTire + 0) * S + slot] = s.xe[i];
                w[(i * Rr::kTire + 1) * S + slot] = s.ye[i];
                w[(i * Rr::kTire + 2) * S + slot] = k.xedot[i];
                w[(i * Rr::kTire + 3) * S + slot] = k.yedot[i];
            }
            // omega .. vgy (vgz, vsz are never written by the model), vsx vsy My engTor
            const double f[13] = {s.omega[i], x.dOmega[i], s.xt[i],   x.dxt[i],  s.rStat[i], x.fx[i], x.fy[i],
                                  x.fz[i],    x.fxg[i],    x.fyg[i],  x.fzg[i],  x.vgx[i],   x.vgy[i]};
#pragma unroll
            for (int j = 0; j < 13; ++j) w[Rr::f(i, j) * S + slot] = f[j];
            w[Rr::f(i, 14) * S + slot] = x.vsx[i];
            w[Rr::f(i, 15) * S + slot] = x.vsy[i];
            w[Rr::f(i, 17) * S + slot] = x.My[i];
            w[Rr::f(i, 18) * S + slot] = x.engTor[i];
            // strut: xs . dxs ls . us vs ws dus dvs . uu vu wu duu dvu dwu fxs fys fzs fzd mx my mz
            w[Rr::q(i, 0) * S + slot] = s.xs[i];
            w[Rr::q(i, 2) * S + slot] = s.dxs[i];
            w[Rr::q(i, 3) * S + slot] = x.ls[i];
            const double g[5] = {x.us[i], x.vs[i], x.ws[i], x.dus[i], x.dvs[i]};
#pragma unroll
            for (int j = 0; j < 5; ++j) w[Rr::q(i, 5 + j) * S + slot] = g[j];
            const double h[13] = {x.uu[i],  x.vu[i],  s.wu[i],  x.duu[i], x.dvu[i], x.dwu[i], x.fxs[i],
                                  x.fys[i], x.fzs[i], s.fzd[i], x.mx[i],  x.my[i],  0.0};
#pragma unroll
            for (int j = 0; j < 13; ++j) w[Rr::q(i, 11 + j) * S + slot] = h[j];
        }
        double* v = w + (size_t)Rr::kVeh * S + slot;
        v[0 * S] = s.x; v[1 * S] = s.y; v[3 * S] = s.u; v[4 * S] = s.v; v[5 * S] = s.w; v[6 * S] = s.psi;
        v[7 * S] = s.wz; v[8 * S] = s.phi; v[9 * S] = s.wx; v[10 * S] = s.theta; v[11 * S] = s.wy;
        v[12 * S] = s.udot; v[13 * S] = s.vdot; v[14 * S] = s.wdot; v[15 * S] = s.wxdot; v[16 * S] = s.wydot;
        v[17 * S] = s.wzdot; v[18 * S] = x.dtheta; v[19 * S] = x.dphi; v[20 * S] = x.dpsi; v[21 * S] = s.crank;
        if (P.tcbool) v[22 * S] = x.dcrank;
        v[23 * S] = (double)s.gear;
    }
};

// Occupancy of the integration kernels (waves per SIMD).  Each model has a default, M::kWaves, the
// most waves per SIMD it is built for: on large fleets more resident waves hide more FP64 latency
// than the register spills of a smaller budget cost, up to four for the 11-DOF step (five run 2 %
// slower), three for the 18-DOF step (four spill twice as much and run 12 % slower) and two for the
// 24-DOF step (1.37x faster than one; three would spill over a kilobyte per lane).  The spills are
// counted for the step loop alone: the last step, which also returns the full record, holds more
// registers, but runs once.  A fleet that fills at most one wavefront per SIMD gains nothing
// from the other waves' register budget, so every kernel also exists at every lower occupancy;
// Launch::waves picks (0: the model's default).  STRICT instantiations evaluate the step with the
// reference's own quotients and transcendentals (Launch::strict; built at the model's default
// occupancy only).
template <class M, bool NR, int MODE, int W, bool STRICT, class F = NoControls>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void integrate(Launch L, F f) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = uniform(slot >> 6);
    if (wave * kWave >= L.S) return;
    // steps of this lane and of its wavefront in this launch: all of them, or up to the vehicle's
    // horizon (the wavefront runs until its last lane's horizon, finished lanes idle)
    int my = L.nsteps, wsteps = L.nsteps;
    if (L.horizon) {
        const int h = L.horizon[slot] - L.step_base;
        my = h < 0 ? 0 : (h > L.nsteps ? L.nsteps : h);
        const int hw = uniform(L.wave_hmax[wave]) - L.step_base;
        wsteps = hw < 0 ? 0 : (hw > L.nsteps ? L.nsteps : hw);
        if (wsteps == 0 && L.n_out == 0) return;
    }
    const int set = uniform(L.wave_set[wave]);
    const PSet& P = pset_at(L.sets, set);
    const size_t S = (size_t)L.S;

    // the per-vehicle constants: in this lane's LDS slot (M::kKInLds; the launch sizes the dynamic LDS
    // with k_lds_bytes), read where the step uses them, or in registers
    extern __shared__ double lfdm_k_lds[];
    typename M::K k_reg;
    typename M::K& k = M::kKInLds ? reinterpret_cast<typename M::K*>(lfdm_k_lds)[threadIdx.x] : k_reg;
    typename M::S s;
    M::template load<NR>(L.state, S, slot, s, k);
    const Derived& u = P.d;
    M::kinit(P, L.ovr + slot, S, k);
    const int user = L.slot_user[slot];
    const int uv = user >= 0 ? user : 0;  // padding lanes read vehicle 0's controls
    CtrlState<MODE> cs;
    double t = L.t0;  // this vehicle's time
    const double dt = L.dt;
    // output rows after fewer steps than the vehicle's last are written as the wavefront passes
    // them; the rest after the last step
    int oi = 0;
    int next_out = L.n_out > 0 ? L.out_steps[0] : -1;
    // all steps but the lane's last keep only the carried state ...  Each step's position update
    // (x, y and, in the 11- and 18-DOF models, psi; nothing else in a step reads them) is applied at
    // the start of the next step, or before an output row or the last step reads the state: the
    // same operations on the same values, without the yaw angle's sine and cosine, which have a
    // branch for huge angles, splitting the step's code where it holds the most registers
    bool pend = false;  // this lane's last step left its position update
    int clean = 0;      // a lane stepping at kk != clean has stepped at kk - 1 with no output row since
    for (int kk = 0; kk < wsteps - 1; ++kk) {
        if constexpr (M::kKInLds) LFDM_RELOAD_MEMORY();
        if (kk < my - 1) {
            if (kk != clean) M::template position<STRICT>(s, dt);  // the same branch on every lane
            const Ctl c = controls_at<MODE>(L, P, cs, uv, kk, t, f);
            M::template step<NR, false, STRICT, false>(P, u, k, s, c, dt, nullptr);
            pend = true;
            t = t + dt;
        }
        if (kk + 1 == next_out) {
            if (pend) M::template position<STRICT>(s, dt);
            pend = false;
            clean = kk + 1;
            if (kk + 1 < my && user >= 0)
                M::emit(L.out + (size_t)oi * M::kCh * L.n_user + user, (size_t)L.n_user, t, s);
            ++oi;
            next_out = oi < L.n_out ? L.out_steps[oi] : -1;
        }
    }
    if (pend) M::template position<STRICT>(s, dt);
    // ... the last one also returns the right-hand-side quantities the full record holds (a lane
    // whose horizon comes earlier than its wavefront's takes it at its own step and time)
    if (my > 0) {
        typename M::X x;
        CtrlState<MODE> cl;
        const Ctl c = controls_at<MODE>(L, P, my == wsteps ? cs : cl, uv, my - 1, t, f);
        M::template step<NR, true, STRICT, false>(P, u, k, s, c, dt, &x);
        M::template position<STRICT>(s, dt);
        t = t + dt;
        M::template store<NR>(L.state, S, slot, P, s, k, x);
    }
    // rows at or after the vehicle's last step: its final state, at the row's time
    if (user >= 0) {
        double tr = t;
        int at = my;
        for (int r = 0; r < L.n_out; ++r) {
            const int st = L.out_steps[r];
            if (st < my) continue;
            for (; at < st; ++at) tr = tr + dt;
            M::emit(L.out + (size_t)r * M::kCh * L.n_user + user, (size_t)L.n_user, tr, s);
        }
    }
}

// Dynamic LDS of an integration launch of `block` threads: one K per lane where the model keeps its
// per-vehicle constants in LDS.
template <class M>
constexpr size_t k_lds_bytes(int block) {
    return M::kKInLds ? (size_t)block * sizeof(typename M::K) : 0;
}

template <class M, bool NR, int W, bool STRICT>
hipError_t launch_mode(const Launch& L, int block, hipStream_t stream) {
    const int grid = (L.S + block - 1) / block;
    const size_t lds = k_lds_bytes<M>(block);
    switch (L.mode) {
        case kCtrlTable: hipLaunchKernelGGL((integrate<M, NR, kCtrlTable, W, STRICT>), grid, block, lds, stream, L, NoControls{}); break;
        case kCtrlConst: hipLaunchKernelGGL((integrate<M, NR, kCtrlConst, W, STRICT>), grid, block, lds, stream, L, NoControls{}); break;
        case kCtrlKnots: hipLaunchKernelGGL((integrate<M, NR, kCtrlKnots, W, STRICT>), grid, block, lds, stream, L, NoControls{}); break;
        case kCtrlSteps: hipLaunchKernelGGL((integrate<M, NR, kCtrlSteps, W, STRICT>), grid, block, lds, stream, L, NoControls{}); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class M, int W, bool STRICT = false>
hipError_t launch_tire(const Launch& L, int block, hipStream_t stream) {
    return L.nr ? launch_mode<M, true, W, STRICT>(L, block, stream) : launch_mode<M, false, W, STRICT>(L, block, stream);
}

template <class M>
hipError_t launch_model(const Launch& L, int block, hipStream_t stream) {
    if (L.strict) return launch_tire<M, M::kWaves, true>(L, block, stream);
    const int w = L.waves == 0 ? M::kWaves : L.waves;
    if constexpr (M::kWaves >= 4) {
        if (w == 4) return launch_tire<M, 4>(L, block, stream);
    }
    if constexpr (M::kWaves >= 3) {
        if (w == 3) return launch_tire<M, 3>(L, block, stream);
    }
    if constexpr (M::kWaves >= 2) {
        if (w == 2) return launch_tire<M, 2>(L, block, stream);
    }
    return launch_tire<M, 1>(L, block, stream);
}

}  // namespace kern
}  // namespace lfdm
