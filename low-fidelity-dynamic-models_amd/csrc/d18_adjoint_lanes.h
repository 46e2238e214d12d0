// Per-lane bodies of the differentiation-form kernels (d18_adjoint.hip): one vehicle slot each.
// Host-and-device code, so the kernels' exact arithmetic and indexing can also be run on the CPU.
#pragma once
#include "d18_adjoint.h"
#include "kernels.h"

namespace lfdm {
namespace dev {

constexpr int kCarried = 32;  // checkpoint: packed S18, xe ye xedot yedot omega (x4), 11 chassis fields, gear
// One step of the tape: what the step adjoint reads when it is given the step's tire outputs.  It
// never reads x, y or the tire deflection rates (they enter the step only through updates whose
// adjoints need no values), so the tape leaves them out.
constexpr int kTapeState = 22;  // xe ye omega (x4), u v psi wz phi wx udot vdot crank, gear
constexpr int kTapeTires = 16;  // fx fy My rStat (x4)
constexpr int kSegRow = kTapeState + kTapeTires;

LFDM_HD void tape_put_state(const S18& s, double* p, size_t st) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[(0 + i) * st] = s.xe[i];
        p[(4 + i) * st] = s.ye[i];
        p[(8 + i) * st] = s.omega[i];
    }
    const double v[9] = {s.u, s.v, s.psi, s.wz, s.phi, s.wx, s.udot, s.vdot, s.crank};
#pragma unroll
    for (int j = 0; j < 9; ++j) p[(12 + j) * st] = v[j];
    p[21 * st] = (double)s.gear;
}

LFDM_HD void tape_put_tires(const TireOut18& t, double* p, size_t st) {
    p += (size_t)kTapeState * st;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[(0 + i) * st] = t.fx[i];
        p[(4 + i) * st] = t.fy[i];
        p[(8 + i) * st] = t.My[i];
        p[(12 + i) * st] = t.rStat[i];
    }
}

// Reads one tape step; the fields the tape leaves out are set to zero.
LFDM_HD void tape_get(const double* p, size_t st, S18& s, TireOut18& t) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        s.xe[i] = p[(0 + i) * st];
        s.ye[i] = p[(4 + i) * st];
        s.omega[i] = p[(8 + i) * st];
        s.xedot[i] = s.yedot[i] = 0.0;
    }
    s.x = s.y = 0.0;
    s.u = p[12 * st]; s.v = p[13 * st]; s.psi = p[14 * st]; s.wz = p[15 * st]; s.phi = p[16 * st];
    s.wx = p[17 * st]; s.udot = p[18 * st]; s.vdot = p[19 * st]; s.crank = p[20 * st];
    s.gear = (int)p[21 * st];
    const double* q = p + (size_t)kTapeState * st;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        t.fx[i] = q[(0 + i) * st];
        t.fy[i] = q[(4 + i) * st];
        t.My[i] = q[(8 + i) * st];
        t.rStat[i] = q[(12 + i) * st];
        t.xedot[i] = t.yedot[i] = 0.0;
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

// Tape of a backward group, blocked by wavefront: [q][wave][K][kSegRow][64], so the rows a
// wavefront writes (and later reads) for one step are one contiguous run of kSegRow x 512 bytes.
// Returns the lane's field-0 entry of step jj of tape block q; fields are 64 doubles apart.
LFDM_HD double* tape_row(const AdjLaunch& L, int q, int jj, int slot) {
    const size_t wave = (size_t)q * (size_t)(L.S / kWave) + (size_t)(slot / kWave);
    return L.tape + ((wave * (size_t)L.K + (size_t)jj) * kSegRow) * kWave + (size_t)(slot % kWave);
}

// Forward pass of one slot: checkpoints every K steps and (x, y) after every step.  The steps of
// segments fwd_tape_lo and later (the group the backward walks first) also go to the tape, so the
// backward needs no recompute for that group.
LFDM_HD void adj_forward_lane(const AdjLaunch& L, int slot, const PSet& P) {
    const size_t S = (size_t)L.S, N = (size_t)L.n_user;
    S18 s;
    load_record(L.state, S, slot, s);
    K18 k;
    load_k(P, L.ovr + slot, S, k);
    const int user = L.slot_user[slot];
    const int uv = user >= 0 ? user : 0;
    const int jt = L.fwd_tape_lo * L.K < L.nsteps ? L.fwd_tape_lo * L.K : L.nsteps;
    int seg = 0, jj = 0;  // segment of step j and the step's index in it
    for (int j = 0; j < jt; ++j) {
        if (jj == 0) pack(s, L.ckpt + (size_t)seg * kCarried * S + slot, S);
        const double* c = L.ctrl + (size_t)j * 3 * N + uv;
        step18<false, false>(P, P.d, k, s, c[0], c[N], c[2 * N], L.dt, nullptr);
        if (user >= 0) {
            L.xy[(size_t)j * 2 * N + user] = s.x;
            L.xy[((size_t)j * 2 + 1) * N + user] = s.y;
        }
        if (++jj == L.K) {
            jj = 0;
            ++seg;
        }
    }
    for (int j = jt; j < L.nsteps; ++j) {
        if (jj == 0) pack(s, L.ckpt + (size_t)seg * kCarried * S + slot, S);
        double* row = tape_row(L, seg - L.fwd_tape_lo, jj, slot);
        tape_put_state(s, row, kWave);
        const double* c = L.ctrl + (size_t)j * 3 * N + uv;
        TireOut18 to;
        step18<false, false, true>(P, P.d, k, s, c[0], c[N], c[2 * N], L.dt, nullptr, &to);
        tape_put_tires(to, row, kWave);
        if (user >= 0) {
            L.xy[(size_t)j * 2 * N + user] = s.x;
            L.xy[((size_t)j * 2 + 1) * N + user] = s.y;
        }
        if (++jj == L.K) {
            jj = 0;
            ++seg;
        }
    }
}

// Adjoint of the carried state and of the per-vehicle constants, kept between the reverse launches
// of successive segment groups: carry[kCarry][S].
constexpr int kCarry = 31 + 21;
static_assert(kCarried == kAdjCkptRow && kSegRow == kAdjTapeRow && kCarry == kAdjCarry, "adjoint buffer rows");

LFDM_HD void store_carry(const S18b& l, const K18b& kb, double* p, size_t S) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[(0 + i) * S] = l.xe[i];
        p[(4 + i) * S] = l.ye[i];
        p[(8 + i) * S] = l.xedot[i];
        p[(12 + i) * S] = l.yedot[i];
        p[(16 + i) * S] = l.omega[i];
    }
    const double v[32] = {l.x,      l.y,       l.u,     l.v,         l.psi,    l.wz,    l.phi,    l.wx,
                          l.udot,   l.vdot,    l.crank, kb.Z1f,      kb.Z2fc,  kb.Z1r,  kb.Z2rc,  kb.Z4n,
                          kb.neg_mt, kb.E3c,   kb.hm,   kb.A2,       kb.A3,    kb.inv_mt, kb.two_hrc_m, kb.denom,
                          kb.A1A2,  kb.A1A3,   kb.A3sq, kb.qy1,      kb.qy2,   kb.jz,   kb.tqs,   kb.maxBrake};
#pragma unroll
    for (int j = 0; j < 32; ++j) p[(20 + j) * S] = v[j];
}

LFDM_HD void load_carry(const double* p, size_t S, S18b& l, K18b& kb) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        l.xe[i] = p[(0 + i) * S];
        l.ye[i] = p[(4 + i) * S];
        l.xedot[i] = p[(8 + i) * S];
        l.yedot[i] = p[(12 + i) * S];
        l.omega[i] = p[(16 + i) * S];
    }
    double v[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) v[j] = p[(20 + j) * S];
    l.x = v[0]; l.y = v[1]; l.u = v[2]; l.v = v[3]; l.psi = v[4]; l.wz = v[5]; l.phi = v[6]; l.wx = v[7];
    l.udot = v[8]; l.vdot = v[9]; l.crank = v[10];
    kb.Z1f = v[11]; kb.Z2fc = v[12]; kb.Z1r = v[13]; kb.Z2rc = v[14]; kb.Z4n = v[15]; kb.neg_mt = v[16];
    kb.E3c = v[17]; kb.hm = v[18]; kb.A2 = v[19]; kb.A3 = v[20]; kb.inv_mt = v[21]; kb.two_hrc_m = v[22];
    kb.denom = v[23]; kb.A1A2 = v[24]; kb.A1A3 = v[25]; kb.A3sq = v[26]; kb.qy1 = v[27]; kb.qy2 = v[28];
    kb.jz = v[29]; kb.tqs = v[30]; kb.maxBrake = v[31];
}


// Segment recompute of the backward pass, one lane per (segment seg_lo + q, slot): from the
// segment's checkpoint, reruns its steps and writes each step's carried state and tire outputs to
// tape block q.  Segments are independent, so a launch covers a whole group.  k is the lane's
// storage for its per-vehicle constants (LDS on the GPU, read where each step uses them).
LFDM_HD void adj_recompute_lane(const AdjLaunch& L, int q, int slot, const PSet& P, K18& k) {
    const int seg = L.seg_lo + q;
    if (seg >= L.seg_hi) return;
    const size_t S = (size_t)L.S, N = (size_t)L.n_user;
    load_k(P, L.ovr + slot, S, k);
    S18 s;
    unpack(L.ckpt + (size_t)seg * kCarried * S + slot, S, s);
    const int user = L.slot_user[slot];
    const int uv = user >= 0 ? user : 0;
    const int j0 = seg * L.K;
    const int j1 = (L.nsteps < j0 + L.K ? L.nsteps : j0 + L.K);
    for (int j = j0; j < j1; ++j) {
        LFDM_RELOAD_MEMORY();
        double* row = tape_row(L, q, j - j0, slot);
        tape_put_state(s, row, kWave);
        const double* c = L.ctrl + (size_t)j * 3 * N + uv;
        TireOut18 to;
        step18<false, false, true>(P, P.d, k, s, c[0], c[N], c[2 * N], L.dt, nullptr, &to);
        tape_put_tires(to, row, kWave);
    }
}

// Reverse sweep of one slot over the segments [seg_lo, seg_hi), last step first, from the tape the
// recompute launch wrote.  k and kb are the lane's storage for the constants and their adjoints.
// The adjoint enters from carry (zero for the last group) and leaves to carry, or, for the first
// group, to ctrl_bar / state_bar / param_bar.
LFDM_HD void adj_reverse_lane(const AdjLaunch& L, int slot, const PSet& P, K18& k, K18b& kb) {
    const size_t S = (size_t)L.S, N = (size_t)L.n_user;
    const int nseg = (L.nsteps + L.K - 1) / L.K;
    load_k(P, L.ovr + slot, S, k);
    const int user = L.slot_user[slot];
    const int uv = user >= 0 ? user : 0;
    S18b lam;
    if (L.seg_hi >= nseg) {
        s18b_zero(lam);
        k18b_zero(kb);
    } else {
        load_carry(L.carry + slot, S, lam, kb);
    }
    for (int seg = L.seg_hi - 1; seg >= L.seg_lo; --seg) {
        const int j0 = seg * L.K;
        const int j1 = (L.nsteps < j0 + L.K ? L.nsteps : j0 + L.K);
        for (int j = j1 - 1; j >= j0; --j) {
            LFDM_RELOAD_MEMORY();
            S18 s;
            TireOut18 to;
            tape_get(tape_row(L, seg - L.seg_lo, j - j0, slot), kWave, s, to);
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
    if (L.seg_lo > 0) {
        store_carry(lam, kb, L.carry + slot, S);
        return;
    }
    if (user < 0) return;
    // adjoint of the initial record (TMeasy layout: 15 fields per tire, vehicle block at 60); the
    // fields the dynamics never read are zero
    double* g = L.state_bar + user;
    for (int f = 0; f < 75; ++f) g[(size_t)f * N] = 0.0;
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

// Step Jacobian of one slot at its current record: one adjoint sweep of the step per stepped packed
// state (rows), giving its derivatives with respect to the packed states and the three controls
// (columns).  The packed order is the reference's packY order of a vehicle with a torque converter
// (dof18_halfImplicit.cpp:320-341); without a converter crankOmega is no state of the step (the
// reference leaves it out): its column is zero and its row the derivative of the slaved crank speed.
// The lagged accelerations, deflection rates and gear of the record are held at their values.
LFDM_HD void jac_lane(const AdjLaunch& L, int slot, const PSet& P) {
    const int user = L.slot_user[slot];
    if (user < 0) return;
    const size_t S = (size_t)L.S, N = (size_t)L.n_user;
    S18 s;
    load_record(L.state, S, slot, s);
    K18 k;
    load_k(P, L.ovr + slot, S, k);
    const double steer = L.ctrl[user], thr = L.ctrl[N + user], brk = L.ctrl[2 * N + user];
    for (int r = 0; r < kJacRows; ++r) {
        S18b sb;
        s18b_zero(sb);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            sb.xe[i] = r == 2 * i ? 1.0 : 0.0;
            sb.ye[i] = r == 2 * i + 1 ? 1.0 : 0.0;
            sb.omega[i] = r == 8 + i ? 1.0 : 0.0;
        }
        sb.crank = r == 12 ? 1.0 : 0.0;
        sb.x = r == 13 ? 1.0 : 0.0;
        sb.y = r == 14 ? 1.0 : 0.0;
        sb.u = r == 15 ? 1.0 : 0.0;
        sb.v = r == 16 ? 1.0 : 0.0;
        sb.psi = r == 17 ? 1.0 : 0.0;
        sb.wz = r == 18 ? 1.0 : 0.0;
        sb.phi = r == 19 ? 1.0 : 0.0;
        sb.wx = r == 20 ? 1.0 : 0.0;
        K18b kb;
        k18b_zero(kb);
        double gs = 0, gt = 0, gb = 0;
        step18_adj(P, P.d, k, s, steer, thr, brk, L.dt, sb, gs, gt, gb, kb);
        const double col[kJacCols] = {sb.xe[0], sb.ye[0], sb.xe[1], sb.ye[1], sb.xe[2], sb.ye[2], sb.xe[3], sb.ye[3],
                                      sb.omega[0], sb.omega[1], sb.omega[2], sb.omega[3],
                                      P.tcbool ? sb.crank : 0.0, sb.x, sb.y, sb.u, sb.v, sb.psi, sb.wz, sb.phi, sb.wx,
                                      gs, gt, gb};
        double* row = L.jac + (size_t)r * kJacCols * N + user;
#pragma unroll
        for (int j = 0; j < kJacCols; ++j) row[(size_t)j * N] = col[j];
    }
}

}  // namespace dev
}  // namespace lfdm
