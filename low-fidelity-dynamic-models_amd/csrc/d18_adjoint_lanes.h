// Per-lane bodies of the differentiation-form kernels (d18_adjoint.hip): one vehicle slot each.
// Host-and-device code, so the kernels' exact arithmetic and indexing can also be run on the CPU.
#pragma once
#include "d18_adjoint.h"
#include "kernels.h"

namespace lfdm {
namespace dev {

constexpr int kCarried = 32;  // packed S18: xe ye xedot yedot omega (x4), 11 chassis fields, gear
constexpr int kTireOut = 24;  // packed TireOut18: fx fy My rStat xedot yedot (x4)
constexpr int kSegRow = kCarried + kTireOut;  // one step of the segment buffer

LFDM_HD void pack_to(const TireOut18& t, double* p, size_t S) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[(0 + i) * S] = t.fx[i];
        p[(4 + i) * S] = t.fy[i];
        p[(8 + i) * S] = t.My[i];
        p[(12 + i) * S] = t.rStat[i];
        p[(16 + i) * S] = t.xedot[i];
        p[(20 + i) * S] = t.yedot[i];
    }
}

LFDM_HD void unpack_to(const double* p, size_t S, TireOut18& t) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        t.fx[i] = p[(0 + i) * S];
        t.fy[i] = p[(4 + i) * S];
        t.My[i] = p[(8 + i) * S];
        t.rStat[i] = p[(12 + i) * S];
        t.xedot[i] = p[(16 + i) * S];
        t.yedot[i] = p[(20 + i) * S];
    }
}

LFDM_HD void pack(const S18& s, double* p, size_t S) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[(0 + i) * S] = s.xe[i];
        p[(4 + i) * S] = s.ye[i];
        p[(8 + i) * S] = s.xedot[i];
        p[(12 + i) * S] = s.yedot[i];
        p[(16 + i) * S] = s.omega[i];
    }
    const double v[11] = {s.x, s.y, s.u, s.v, s.psi, s.wz, s.phi, s.wx, s.udot, s.vdot, s.crank};
#pragma unroll
    for (int j = 0; j < 11; ++j) p[(20 + j) * S] = v[j];
    p[31 * S] = (double)s.gear;
}

LFDM_HD void unpack(const double* p, size_t S, S18& s) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        s.xe[i] = p[(0 + i) * S];
        s.ye[i] = p[(4 + i) * S];
        s.xedot[i] = p[(8 + i) * S];
        s.yedot[i] = p[(12 + i) * S];
        s.omega[i] = p[(16 + i) * S];
    }
    s.x = p[20 * S]; s.y = p[21 * S]; s.u = p[22 * S]; s.v = p[23 * S]; s.psi = p[24 * S]; s.wz = p[25 * S];
    s.phi = p[26 * S]; s.wx = p[27 * S]; s.udot = p[28 * S]; s.vdot = p[29 * S]; s.crank = p[30 * S];
    s.gear = (int)p[31 * S];
}

// Carried state of one slot from the SoA record array (TMeasy layout).
LFDM_HD void load_record(const double* st, size_t S, int slot, S18& s) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        s.xe[i] = st[(i * 15 + 0) * S + slot];
        s.ye[i] = st[(i * 15 + 1) * S + slot];
        s.xedot[i] = st[(i * 15 + 2) * S + slot];
        s.yedot[i] = st[(i * 15 + 3) * S + slot];
        s.omega[i] = st[(i * 15 + 4) * S + slot];
    }
    const double* v = st + (size_t)60 * S + slot;
    s.x = v[0 * S]; s.y = v[1 * S]; s.u = v[2 * S]; s.v = v[3 * S]; s.psi = v[4 * S]; s.wz = v[5 * S];
    s.phi = v[6 * S]; s.wx = v[7 * S]; s.udot = v[8 * S]; s.vdot = v[9 * S]; s.crank = v[12 * S];
    s.gear = (int)v[14 * S];
}

LFDM_HD void load_k(const PSet& P, const double* o, size_t S, K18& k) {
    k18_init(P, o[0], o[S], o[2 * S], o[3 * S], o[4 * S], o[5 * S], o[6 * S], k);
}

LFDM_HD void adj_forward_lane(const AdjLaunch& L, int slot, const PSet& P) {
    const size_t S = (size_t)L.S, N = (size_t)L.n_user;
    S18 s;
    load_record(L.state, S, slot, s);
    K18 k;
    load_k(P, L.ovr + slot, S, k);
    const int user = L.slot_user[slot];
    const int uv = user >= 0 ? user : 0;
    for (int j = 0; j < L.nsteps; ++j) {
        if (j % L.K == 0) pack(s, L.ckpt + (size_t)(j / L.K) * kCarried * S + slot, S);
        const double* c = L.ctrl + (size_t)j * 3 * N + uv;
        step18<false, false>(P, P.d, k, s, c[0], c[N], c[2 * N], L.dt, nullptr);
        if (user >= 0) {
            L.xy[(size_t)j * 2 * N + user] = s.x;
            L.xy[((size_t)j * 2 + 1) * N + user] = s.y;
        }
    }
}

LFDM_HD void adj_backward_lane(const AdjLaunch& L, int slot, const PSet& P) {
    const size_t S = (size_t)L.S, N = (size_t)L.n_user;
    K18 k;
    load_k(P, L.ovr + slot, S, k);
    const int user = L.slot_user[slot];
    const int uv = user >= 0 ? user : 0;
    S18b lam;
    s18b_zero(lam);
    K18b kb;
    k18b_zero(kb);
    const int nseg = (L.nsteps + L.K - 1) / L.K;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int j0 = seg * L.K;
        const int j1 = (L.nsteps < j0 + L.K ? L.nsteps : j0 + L.K);
        S18 s;
        unpack(L.ckpt + (size_t)seg * kCarried * S + slot, S, s);
        for (int j = j0; j < j1; ++j) {
            double* row = L.seg + (size_t)(j - j0) * kSegRow * S + slot;
            pack(s, row, S);
            const double* c = L.ctrl + (size_t)j * 3 * N + uv;
            TireOut18 to;
            step18<false, false, true>(P, P.d, k, s, c[0], c[N], c[2 * N], L.dt, nullptr, &to);
            pack_to(to, row + (size_t)kCarried * S, S);
        }
        for (int j = j1 - 1; j >= j0; --j) {
            const double* row = L.seg + (size_t)(j - j0) * kSegRow * S + slot;
            unpack(row, S, s);
            TireOut18 to;
            unpack_to(row + (size_t)kCarried * S, S, to);
            lam.x += L.xy_bar[(size_t)j * 2 * N + uv];
            lam.y += L.xy_bar[((size_t)j * 2 + 1) * N + uv];
            const double* c = L.ctrl + (size_t)j * 3 * N + uv;
            double sb = 0, tb = 0, bb = 0;
            step18_adj(P, P.d, k, s, c[0], c[N], c[2 * N], L.dt, lam, sb, tb, bb, kb, &to);
            if (user >= 0) {
                double* cb = L.ctrl_bar + (size_t)j * 3 * N + user;
                cb[0] = sb;
                cb[N] = tb;
                cb[2 * N] = bb;
            }
        }
    }
    if (user < 0) return;
    // adjoint of the initial record (TMeasy layout: 15 fields per tire, vehicle block at 60)
    double* g = L.state_bar + user;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        g[(15 * i + 0) * N] = lam.xe[i];
        g[(15 * i + 1) * N] = lam.ye[i];
        g[(15 * i + 2) * N] = lam.xedot[i];
        g[(15 * i + 3) * N] = lam.yedot[i];
        g[(15 * i + 4) * N] = lam.omega[i];
    }
    const double v[10] = {lam.x, lam.y, lam.u, lam.v, lam.psi, lam.wz, lam.phi, lam.wx, lam.udot, lam.vdot};
#pragma unroll
    for (int j = 0; j < 10; ++j) g[(60 + j) * N] = v[j];
    g[72 * N] = lam.crank;
    double pb[7] = {0, 0, 0, 0, 0, 0, 0};
    k18_init_adj(P, k, kb, pb);
#pragma unroll
    for (int i = 0; i < 7; ++i) L.param_bar[i * N + user] = pb[i];
}

}  // namespace dev
}  // namespace lfdm
