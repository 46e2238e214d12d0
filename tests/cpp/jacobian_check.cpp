// Host build of the 18-DOF step adjoint (csrc/d18_adjoint.h) read out as the Jacobian of one step.
//
// Reads from stdin "steering throttle braking dt" and then a full 75-field TMeasy state record
// (tire LF, RF, LR, RR: xe ye xedot yedot omega dOmega xt rStat fx fy fz vsx vsy My engTor; vehicle:
// x y u v psi wz phi wx udot vdot wxdot wzdot crankOmega dOmega_crank gear).  The step starts from
// that record as the reference's PerturbRhsFun sees it (wheeled_vehicle_models/18dof/
// dof18_halfImplicit.cpp:679-710: freshly constructed states, so gear 0, lagged udot = vdot = 0 and
// xedot = yedot = 0; only the packed states are taken from the record).  Seeding the adjoint with each
// packed output state in turn gives one row of d(step)/d(packed input state) and
// d(step)/d(steering, throttle).  Output: the stepped u and v, then n rows of n + 2 numbers in the
// reference's packY order (:320-341; n = 21 with a torque converter, 20 without).
// With a third argument "asis" the step starts from the record exactly as given (gear, lagged
// accelerations and deflection rates included) and the output is the layout of lfdm_step_jacobian:
// 21 rows of 24 numbers (packY order of a vehicle with a torque converter, then steering, throttle,
// braking; the crankOmega column zero without a converter), no stepped u and v.
// Usage: jacobian_check <vehicle.json> <tmeasy.json> [asis] < input
#include <cstdio>
#include <string>
#include <vector>

#include "../../low-fidelity-dynamic-models_amd/csrc/d18_adjoint.h"
#include "../../low-fidelity-dynamic-models_amd/csrc/params.h"

using namespace lfdm;
using namespace lfdm::dev;

namespace {

// Addresses of the packed states, in packY order.
template <class S>
std::vector<double*> packed(S& s, bool tc) {
    std::vector<double*> p;
    for (int i = 0; i < 4; ++i) {
        p.push_back(&s.xe[i]);
        p.push_back(&s.ye[i]);
    }
    for (int i = 0; i < 4; ++i) p.push_back(&s.omega[i]);
    if (tc) p.push_back(&s.crank);
    for (double* v : {&s.x, &s.y, &s.u, &s.v, &s.psi, &s.wz, &s.phi, &s.wx}) p.push_back(v);
    return p;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s vehicle.json tmeasy.json < input\n", argv[0]);
        return 2;
    }
    std::vector<double> flat;
    std::string err;
    PSet P;
    if (!params_from_json(18, argv[1], argv[2], "", 0, flat, err) ||
        !pset_from_flat(18, flat.data(), (int)flat.size(), 0, P, err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 2;
    }
    if (P.tire_nr) {
        std::fprintf(stderr, "the step adjoint covers the TMeasy tire only\n");
        return 2;
    }
    double steer, thr, brk, dt, rec[75];
    if (std::scanf("%lf %lf %lf %lf", &steer, &thr, &brk, &dt) != 4) return 2;
    for (double& v : rec)
        if (std::scanf("%lf", &v) != 1) return 2;

    const bool asis = argc > 3 && std::string(argv[3]) == "asis";
    S18 s{};
    for (int i = 0; i < 4; ++i) {
        s.xe[i] = rec[15 * i];
        s.ye[i] = rec[15 * i + 1];
        s.omega[i] = rec[15 * i + 4];
        if (asis) {
            s.xedot[i] = rec[15 * i + 2];
            s.yedot[i] = rec[15 * i + 3];
        }
    }
    const double* v = rec + 60;
    s.x = v[0]; s.y = v[1]; s.u = v[2]; s.v = v[3]; s.psi = v[4]; s.wz = v[5]; s.phi = v[6]; s.wx = v[7];
    s.crank = P.tcbool || asis ? v[12] : 0.0;
    s.gear = asis ? (int)v[14] : 0;
    if (asis) {
        s.udot = v[8];
        s.vdot = v[9];
    }

    K18 k;
    k18_init(P, P.m, P.jz, P.jx, P.tire.dfy0Pn, P.tire.dfy0P2n, 1.0, P.maxBrakeTorque, k);
    if (asis) {
        for (int r = 0; r < 21; ++r) {
            S18b sb;
            s18b_zero(sb);
            *packed(sb, true)[r] = 1.0;
            K18b kb;
            k18b_zero(kb);
            double gs = 0, gt = 0, gb = 0;
            step18_adj(P, P.d, k, s, steer, thr, brk, dt, sb, gs, gt, gb, kb);
            const std::vector<double*> col = packed(sb, true);
            for (int j = 0; j < 21; ++j) std::printf("%.17g ", (j == 12 && !P.tcbool) ? 0.0 : *col[j]);
            std::printf("%.17g %.17g %.17g\n", gs, gt, gb);
        }
        return 0;
    }
    S18 s1 = s;
    step18<false, false>(P, P.d, k, s1, steer, thr, brk, dt, nullptr);
    std::printf("%.17g %.17g\n", s1.u, s1.v);

    const bool tc = P.tcbool != 0;
    const int n = (int)packed(s, tc).size();
    for (int r = 0; r < n; ++r) {
        S18b sb;
        s18b_zero(sb);
        *packed(sb, tc)[r] = 1.0;
        K18b kb;
        k18b_zero(kb);
        double gs = 0, gt = 0, gb = 0;
        step18_adj(P, P.d, k, s, steer, thr, brk, dt, sb, gs, gt, gb, kb);
        const std::vector<double*> col = packed(sb, tc);
        for (int j = 0; j < n; ++j) std::printf("%.17g ", *col[j]);
        std::printf("%.17g %.17g\n", gs, gt);
    }
    return 0;
}
