// Host run of the differentiation-form kernels' lane code (csrc/d18_adjoint_lanes.h): the exact
// per-vehicle arithmetic and buffer indexing of adj_forward / adj_backward, over host buffers laid
// out as the library lays them out on the device (slots padded to 64, SoA records, checkpoints
// every K steps with a partial last segment).  Each base vehicle has copies whose initial state,
// controls or parameters are moved by -2h, -h, h, 2h; the adjoint of J = sum_k w_k . xy_k is
// compared with fourth-order central differences of the same lane code.  Exits non-zero if any
// group's normwise relative error exceeds 1e-6 (1e-5 for the tire deflections xe, ye, whose
// finite differences stop converging below ~1e-6: steps small enough to stay off the slip-force
// kinks are within a decade of the forward map's rounding noise).
// Usage: adjoint_lanes_check <vehicle.json> <tmeasy.json>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../low-fidelity-dynamic-models_amd/csrc/d18_adjoint_lanes.h"
#include "../../low-fidelity-dynamic-models_amd/csrc/params.h"

using namespace lfdm;
using namespace lfdm::dev;

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::vector<double> flat;
    std::string err;
    PSet P;
    if (!params_from_json(18, argv[1], argv[2], "", 0, flat, err) ||
        !pset_from_flat(18, flat.data(), (int)flat.size(), 0, P, err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 2;
    }
    const int nb = 2, T = 600, rec = 75;
    const int K = argc > 3 ? std::atoi(argv[3]) : 64;
    // finite-difference steps: tire deflections are ~1e-5 m and stiff, roll states ~1e-3, so each
    // direction gets a step well inside its own scale; controls and the rest use larger steps
    // (the forward map's rounding noise, not truncation, limits the smaller ones)
    const double hx = argc > 4 ? std::atof(argv[4]) : 1e-8;
    // directions: (kind, index): kind 0 = record field, 1 = control (step, channel), 2 = parameter
    struct Dir {
        int kind, a, b;
        double h;
    };
    std::vector<Dir> dirs;
    const int fields[] = {0, 1, 4, 19, 45, 49, 60, 61, 62, 63, 64, 65, 66, 67, 68, 69, 72};
    for (int f : fields) dirs.push_back({0, f, 0, 0});
    std::mt19937_64 rng(5);
    const double hc = argc > 5 ? std::atof(argv[5]) : 1e-4;
    const double hp = argc > 6 ? std::atof(argv[6]) : 3e-6;
    for (int i = 0; i < 10; ++i) dirs.push_back({1, (int)(rng() % T), (int)(rng() % 3), hc});
    for (int i = 0; i < 7; ++i) dirs.push_back({2, i, 0, 0});
    const int per = 1 + 4 * (int)dirs.size();
    const int n = nb * per;
    const int S = (n + 63) / 64 * 64;
    std::vector<double> state((size_t)rec * S, 0.0), ovr((size_t)7 * S, 0.0), ctrl((size_t)T * 3 * n, 0.0);
    std::vector<double> base_state(rec * nb, 0.0), base_ovr(7 * nb), w((size_t)T * 2 * nb);
    std::uniform_real_distribution<double> U(-1, 1);
    for (int b = 0; b < nb; ++b) {
        double* r = &base_state[rec * b];
        const double u0 = 12.0 + 2.0 * b, om = u0 / P.tire.r0;
        for (int i = 0; i < 4; ++i) r[15 * i + 4] = om;
        r[62] = u0;
        r[63] = 0.05;
        r[65] = 0.02;
        r[72] = 1.05 * om / P.gears[0];  // engine 5 % ahead of the wheels: converter speed ratio below 1
        const double o[7] = {P.m * (1 + 0.05 * b), P.jz * (1 - 0.04 * b), P.jx, P.tire.dfy0Pn * (1 + 0.03 * b),
                             P.tire.dfy0P2n * (1 + 0.03 * b), 1.0, P.maxBrakeTorque};
        for (int i = 0; i < 7; ++i) base_ovr[7 * b + i] = o[i];
    }
    for (double& x : w) x = U(rng);
    auto cidx = [&](int j, int q, int v) { return ((size_t)j * 3 + q) * n + v; };
    int v = 0;
    std::vector<double> hs;
    for (int b = 0; b < nb; ++b) {
        for (int c = 0; c < per; ++c, ++v) {
            for (int f = 0; f < rec; ++f) state[(size_t)f * S + v] = base_state[rec * b + f];
            for (int i = 0; i < 7; ++i) ovr[(size_t)i * S + v] = base_ovr[7 * b + i];
            for (int j = 0; j < T; ++j) {
                const double t = j * 1e-3;
                ctrl[cidx(j, 0, v)] = 0.2 * std::sin(2.0 * t + b) + 0.05;
                ctrl[cidx(j, 1, v)] = 0.35 + 0.15 * std::sin(3.0 * t + b);
                ctrl[cidx(j, 2, v)] = t > 0.4 ? 0.08 : 0.0;
            }
            if (c == 0) continue;
            const Dir& d = dirs[(c - 1) / 4];
            const double off[4] = {-2, -1, 1, 2};
            const double o = off[(c - 1) % 4];
            double h;
            if (d.kind == 0) {
                const double x0 = base_state[rec * b + d.a];
                const bool defl = d.a < 60 && d.a % 15 < 2;
                const bool roll = d.a == 66 || d.a == 67;
                h = (defl ? hx : (roll ? 1e-7 : 1e-5)) * std::max(1.0, std::fabs(x0));
                state[(size_t)d.a * S + v] += o * h;
            } else if (d.kind == 1) {
                h = d.h;
                ctrl[cidx(d.a, d.b, v)] += o * h;
            } else {
                h = hp * std::fabs(base_ovr[7 * b + d.a]);
                ovr[(size_t)d.a * S + v] += o * h;
            }
            if ((c - 1) % 4 == 0) hs.push_back(h);
        }
    }
    for (int s = n; s < S; ++s) {  // padding slots copy the last vehicle, like the library
        for (int f = 0; f < rec; ++f) state[(size_t)f * S + s] = state[(size_t)f * S + n - 1];
        for (int i = 0; i < 7; ++i) ovr[(size_t)i * S + s] = ovr[(size_t)i * S + n - 1];
    }
    std::vector<int> slot_user(S, -1);
    for (int s = 0; s < n; ++s) slot_user[s] = s;
    const int nseg = (T + K - 1) / K;
    const int G = 3;  // segments per backward group: 10 segments give groups of 3, 3, 3 and 1
    std::vector<double> ckpt((size_t)nseg * kCarried * S), tape((size_t)G * K * kSegRow * S), xy((size_t)T * 2 * n);
    std::vector<double> carry((size_t)kCarry * S);
    std::vector<double> xyb((size_t)T * 2 * n), cb((size_t)T * 3 * n), sb((size_t)rec * n, 0.0), pb((size_t)7 * n);
    AdjLaunch L{};
    L.slot_user = slot_user.data();
    L.state = state.data();
    L.ovr = ovr.data();
    L.ctrl = ctrl.data();
    L.xy = xy.data();
    L.ckpt = ckpt.data();
    L.tape = tape.data();
    L.carry = carry.data();
    L.dt = 1e-3;
    L.nsteps = T;
    L.K = K;
    L.S = S;
    L.n_user = n;
    L.fwd_tape_lo = nseg - G;  // the forward tapes the last group, as the library does
    for (int s = 0; s < S; ++s) adj_forward_lane(L, s, P);
    std::vector<double> J(n, 0.0);
    v = 0;
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < per; ++c, ++v)
            for (int j = 0; j < T; ++j)
                for (int q = 0; q < 2; ++q) {
                    const double wt = w[((size_t)j * 2 + q) * nb + b];
                    J[v] += wt * xy[((size_t)j * 2 + q) * n + v];
                    xyb[((size_t)j * 2 + q) * n + v] = wt;
                }
    L.xy_bar = xyb.data();
    L.ctrl_bar = cb.data();
    L.state_bar = sb.data();
    L.param_bar = pb.data();
    // the backward launches' schedule: groups of G segments, last group first; recompute lanes
    // (q, slot), then one reverse lane per slot with its own constants storage
    for (int hi = nseg; hi > 0; hi -= G) {
        L.seg_hi = hi;
        L.seg_lo = std::max(hi - G, 0);
        for (int q = 0; q < (hi == nseg ? 0 : hi - L.seg_lo); ++q)  // the last group is on the tape
            for (int s = 0; s < S; ++s) {
                K18 k;
                adj_recompute_lane(L, q, s, P, k);
            }
        for (int s = 0; s < S; ++s) {
            K18 k;
            K18b kb;
            adj_reverse_lane(L, s, P, k, kb);
        }
    }
    double worst[4] = {0, 0, 0, 0};
    int hi = 0;
    for (int b = 0; b < nb; ++b) {
        const int v0 = b * per;
        double num[4] = {0, 0, 0, 0}, den[4] = {0, 0, 0, 0};
        for (size_t di = 0; di < dirs.size(); ++di) {
            const Dir& d = dirs[di];
            const double h = hs[hi++];
            const double* Jd = &J[v0 + 1 + 4 * di];
            const double fd = (Jd[0] - 8 * Jd[1] + 8 * Jd[2] - Jd[3]) / (12 * h);
            double ad, unit = 1.0;
            if (d.kind == 0)
                ad = sb[(size_t)d.a * n + v0];
            else if (d.kind == 1)
                ad = cb[cidx(d.a, d.b, v0)];
            else {
                ad = pb[(size_t)d.a * n + v0];
                unit = std::fabs(base_ovr[7 * b + d.a]);
            }
            std::printf("base %d kind %d (%d,%d): adjoint % .10e fd % .10e\n", b, d.kind, d.a, d.b, ad, fd);
            const int g = (d.kind == 0 && d.a < 60 && d.a % 15 < 2) ? 3 : d.kind;
            num[g] = std::max(num[g], std::fabs(ad - fd) * unit);
            den[g] = std::max(den[g], std::fabs(fd) * unit);
        }
        for (int k = 0; k < 4; ++k) worst[k] = std::max(worst[k], num[k] / std::max(den[k], 1e-300));
    }
    const char* names[4] = {"initial state", "controls", "parameters", "tire deflections"};
    const double tol[4] = {1e-6, 1e-6, 1e-6, 1e-5};
    int bad = 0;
    for (int k = 0; k < 4; ++k) {
        std::printf("%s %-16s normwise relative error %.2e (tolerance %.0e)\n", worst[k] <= tol[k] ? "ok " : "BAD",
                    names[k], worst[k], tol[k]);
        bad += worst[k] > tol[k];
    }
    std::printf("%s\n", bad ? "FAILED" : "ALL OK");
    return bad ? 1 : 0;
}
