// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load liblfdm_oracle.so; the product library never links or calls it.
//
// Flat C entry points over the CPU restatement (orc18.h and siblings).  Every batch entry point
// runs vehicles independently (OpenMP over vehicles, one vehicle's whole horizon per task).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "orc_count.h"
#include "orc_dual.h"
#include "orc11.h"
#include "orc18.h"
#include "orc24.h"

namespace orc {
thread_local OpCount g_ops;
thread_local SeenCall g_seen[512];
thread_local int g_nseen = 0;
}

using namespace orc;

namespace {

// Per-vehicle override fields, in order: m, jz, jx, dfy0Pn, dfy0P2n, torque_scale, maxBrakeTorque.
constexpr int kOverrides = 7;

template <class T>
void apply_overrides_18(d18::Vehicle<T>& s, const double* ovr, int n, int i) {
    if (!ovr) return;
    s.vp.m = T(ovr[0 * n + i]);
    s.vp.jz = T(ovr[1 * n + i]);
    s.vp.jx = T(ovr[2 * n + i]);
    s.tp.dfy0Pn = T(ovr[3 * n + i]);
    s.tp.dfy0P2n = T(ovr[4 * n + i]);
    const double sc = ovr[5 * n + i];
    for (auto& e : s.vp.torque) e.y = e.y * sc;
    s.vp.maxBrakeTorque = T(ovr[6 * n + i]);
}

// Control sources. mode 0: per-step array ctrl[(k*3 + c)*n + i] (c = steering, throttle, braking);
// mode 1: one driver table (rows of t, steering, throttle, braking; imeta[0] = rows);
// mode 2: per-vehicle knot program (imeta = {ns, np}, dmeta = {ds, dp}; knots per vehicle
//         laid out [ns + 2 np][n]).
struct Controls {
    int mode;
    const double* ctrl;
    const int* imeta;
    const double* dmeta;
    int n;
    void at(int i, int k, double t, double& s, double& th, double& b) const {
        if (mode == 0) {
            s = ctrl[((size_t)k * 3 + 0) * n + i];
            th = ctrl[((size_t)k * 3 + 1) * n + i];
            b = ctrl[((size_t)k * 3 + 2) * n + i];
        } else if (mode == 1) {
            driver_input(t, reinterpret_cast<const DriverRow*>(ctrl), imeta[0], s, th, b);
        } else {
            const int ns = imeta[0], np = imeta[1];
            const int w = ns + 2 * np;
            double kn[256];
            for (int j = 0; j < w && j < 256; ++j) kn[j] = ctrl[(size_t)j * n + i];
            knot_input(t, kn, ns, dmeta[0], np, dmeta[1], s, th, b);
        }
    }
};


// Per-model plumbing for the generic runner below.
struct M18 {
    template <class T> using V = d18::Vehicle<T>;
    static int rec(bool nr) { return d18::record_len(nr); }
    template <class T> static void params(V<T>& s, const double* flat) { d18::read_params(flat, s.nr, s.vp, s.tp); }
    template <class T> static void ovr(V<T>& s, const double* o, int n, int i) { apply_overrides_18(s, o, n, i); }
    template <class T> static void step(V<T>& s, const T& st, const T& th, const T& br, double dt) {
        d18::step(s, st, th, br, dt);
    }
    template <class T> static void get(const V<T>& s, double* o) { d18::to_record(s, o); }
    template <class T> static void set(V<T>& s, const double* o) { d18::from_record(s, o); }
};
struct M11 {
    template <class T> using V = d11::Vehicle<T>;
    static int rec(bool nr) { return d11::record_len(nr); }
    template <class T> static void params(V<T>& s, const double* flat) { d11::read_params(flat, s.nr, s.vp, s.tp); }
    template <class T> static void ovr(V<T>& s, const double* o, int n, int i) {
        if (!o) return;
        s.vp.m = T(o[0 * n + i]);
        s.vp.jz = T(o[1 * n + i]);
        s.tp.dfy0Pn = T(o[3 * n + i]);
        s.tp.dfy0P2n = T(o[4 * n + i]);
        const double sc = o[5 * n + i];
        for (auto& e : s.vp.torque) e.y = e.y * sc;
        s.vp.maxBrakeTorque = T(o[6 * n + i]);
    }
    template <class T> static void step(V<T>& s, const T& st, const T& th, const T& br, double dt) {
        d11::step(s, st, th, br, dt);
    }
    template <class T> static void get(const V<T>& s, double* o) { d11::to_record(s, o); }
    template <class T> static void set(V<T>& s, const double* o) { d11::from_record(s, o); }
};
struct M24 {
    template <class T> using V = d24::Vehicle<T>;
    static int rec(bool nr) { return d24::record_len(nr); }
    template <class T> static void params(V<T>& s, const double* flat) {
        d24::read_params(flat, s.nr, s.vp, s.tp, s.sp);
    }
    template <class T> static void ovr(V<T>& s, const double* o, int n, int i) {
        if (!o) return;
        s.vp.m = T(o[0 * n + i]);
        s.vp.jz = T(o[1 * n + i]);
        s.vp.jx = T(o[2 * n + i]);
        s.tp.dfy0Pn = T(o[3 * n + i]);
        s.tp.dfy0P2n = T(o[4 * n + i]);
        const double sc = o[5 * n + i];
        for (auto& e : s.vp.torque) e.y = e.y * sc;
        s.vp.maxBrakeTorque = T(o[6 * n + i]);
    }
    template <class T> static void step(V<T>& s, const T& st, const T& th, const T& br, double dt) {
        d24::step(s, st, th, br, dt);
    }
    template <class T> static void get(const V<T>& s, double* o) { d24::to_record(s, o); }
    template <class T> static void set(V<T>& s, const double* o) { d24::from_record(s, o); }
};

// horizon (optional): vehicle i takes min(nsteps, horizon[i]) steps; samples after its last step repeat
// its final record.
template <class M>
int run_model(int nr, const double* flat, const double* ovr, const double* init, int n, double t0, double dt,
              int nsteps, int mode, const double* ctrl, const int* imeta, const double* dmeta, int every,
              double* out, double* fin, int threads, const int* horizon = nullptr) {
    const int rl = M::rec(nr != 0);
    Controls cs{mode, ctrl, imeta, dmeta, n};
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        typename M::template V<double> s;
        s.nr = nr != 0;
        M::params(s, flat);
        M::ovr(s, ovr, n, i);
        if (init) {
            M::set(s, init + (size_t)i * rl);
        } else {
            std::vector<double> z(rl, 0.0);
            M::set(s, z.data());
        }
        double t = t0;
        int rec = 0;
        const int my = horizon ? std::min(nsteps, std::max(0, horizon[i])) : nsteps;
        for (int k = 0; k < nsteps; ++k) {
            if (k < my) {
                double st, th, br;
                cs.at(i, k, t, st, th, br);
                M::step(s, st, th, br, dt);
                t = t + dt;
            }
            if (out && every > 0 && (k + 1) % every == 0) {
                M::get(s, out + ((size_t)rec * n + i) * rl);
                ++rec;
            }
        }
        if (fin) M::get(s, fin + (size_t)i * rl);
    }
    return 0;
}

}  // namespace

extern "C" {

int orc18_record_len(int nr) { return d18::record_len(nr != 0); }

// Runs n vehicles of one 18-DOF parameter set for nsteps steps of dt from t0 (t += dt per step,
// as dof18_halfImplicit.cpp:229).  init: [n][record] or NULL (default-constructed states).
// out (optional): every `every` steps, the records of all vehicles, [sample][n][record].
// fin (optional): final records [n][record].  Returns 0.
int orc18_run(int nr, const double* flat, const double* ovr, const double* init, int n, double t0, double dt,
              int nsteps, int mode, const double* ctrl, const int* imeta, const double* dmeta, int every, double* out,
              double* fin, int threads) {
    return run_model<M18>(nr, flat, ovr, init, n, t0, dt, nsteps, mode, ctrl, imeta, dmeta, every, out, fin,
                          threads);
}

int orc11_record_len(int nr) { return d11::record_len(nr != 0); }
int orc11_run(int nr, const double* flat, const double* ovr, const double* init, int n, double t0, double dt,
              int nsteps, int mode, const double* ctrl, const int* imeta, const double* dmeta, int every, double* out,
              double* fin, int threads) {
    return run_model<M11>(nr, flat, ovr, init, n, t0, dt, nsteps, mode, ctrl, imeta, dmeta, every, out, fin,
                          threads);
}

// Same as orcNN_run for model `dof` (11, 18 or 24) with per-vehicle horizons horizon[n] (steps from t0).
int orc_run_h(int dof, int nr, const double* flat, const double* ovr, const double* init, int n, double t0, double dt,
              int nsteps, const int* horizon, int mode, const double* ctrl, const int* imeta, const double* dmeta,
              int every, double* out, double* fin, int threads) {
    if (dof == 11)
        return run_model<M11>(nr, flat, ovr, init, n, t0, dt, nsteps, mode, ctrl, imeta, dmeta, every, out, fin,
                              threads, horizon);
    if (dof == 18)
        return run_model<M18>(nr, flat, ovr, init, n, t0, dt, nsteps, mode, ctrl, imeta, dmeta, every, out, fin,
                              threads, horizon);
    if (dof == 24)
        return run_model<M24>(nr, flat, ovr, init, n, t0, dt, nsteps, mode, ctrl, imeta, dmeta, every, out, fin,
                              threads, horizon);
    return -1;
}

int orc24_record_len(int nr) { return d24::record_len(nr != 0); }
int orc24_run(int nr, const double* flat, const double* ovr, const double* init, int n, double t0, double dt,
              int nsteps, int mode, const double* ctrl, const int* imeta, const double* dmeta, int every, double* out,
              double* fin, int threads) {
    return run_model<M24>(nr, flat, ovr, init, n, t0, dt, nsteps, mode, ctrl, imeta, dmeta, every, out, fin,
                          threads);
}
// initializeTireSus on n records in place (vehicle i uses the per-vehicle parameters of ovr, if any).
int orc24_init(int nr, const double* flat, const double* ovr, int n, double* recs) {
    const int rl = d24::record_len(nr != 0);
    for (int i = 0; i < n; ++i) {
        d24::Vehicle<double> s;
        s.nr = nr != 0;
        M24::params(s, flat);
        M24::ovr(s, ovr, n, i);
        d24::from_record(s, recs + (size_t)i * rl);
        d24::init_tire_sus(s);
        d24::to_record(s, recs + (size_t)i * rl);
    }
    return 0;
}

// Forward-mode tangents of the tracking loss L = sum_k dt ((x_k - xr_k)^2 + (y_k - yr_k)^2) of n 18-DOF
// vehicles, (x_k, y_k) the position after step k (the integrand of the reference's FSA tracking cost,
// dof18_sundials.cpp:868-884), over nsteps steps of per-step controls ctrl[nsteps][3][n] from the records
// init[n][rec] with per-vehicle parameters ovr[7][n] (m, jz, jx, dfy0Pn, dfy0P2n, torque scale,
// maxBrakeTorque), xyref[nsteps][2][n].  Three tangent directions per vehicle: 0 along dctrl (layout of
// ctrl), 1 along dovr (layout of ovr), 2 along dinit (layout of init).  Writes loss[n] and
// dloss[3][n] (the directional derivatives).
int orc18_loss_tangents(int nr, const double* flat, const double* ovr, const double* init, int n, int nsteps,
                        double dt, const double* ctrl, const double* xyref, const double* dctrl, const double* dovr,
                        const double* dinit, double* loss, double* dloss, int threads) {
    using D = Dual<3>;
    const int rl = d18::record_len(nr != 0);
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        d18::Vehicle<D> s;
        s.nr = nr != 0;
        d18::read_params(flat, s.nr, s.vp, s.tp);
        auto par = [&](int f) {
            D x(ovr[(size_t)f * n + i]);
            x.d[1] = dovr[(size_t)f * n + i];
            return x;
        };
        s.vp.m = par(0);
        s.vp.jz = par(1);
        s.vp.jx = par(2);
        s.tp.dfy0Pn = par(3);
        s.tp.dfy0P2n = par(4);
        const D sc = par(5);
        for (auto& e : s.vp.torque) e.y = e.y * sc;
        s.vp.maxBrakeTorque = par(6);
        d18::from_record_src(s, [&](int k) {
            D x(init[(size_t)i * rl + k]);
            x.d[2] = dinit[(size_t)i * rl + k];
            return x;
        });
        D L(0.0);
        for (int k = 0; k < nsteps; ++k) {
            D c[3];
            for (int q = 0; q < 3; ++q) {
                c[q] = D(ctrl[((size_t)k * 3 + q) * n + i]);
                c[q].d[0] = dctrl[((size_t)k * 3 + q) * n + i];
            }
            d18::step(s, c[0], c[1], c[2], dt);
            const D ex = s.vs.x - xyref[((size_t)k * 2 + 0) * n + i];
            const D ey = s.vs.y - xyref[((size_t)k * 2 + 1) * n + i];
            L = L + dt * (ex * ex + ey * ey);
        }
        loss[i] = L.v;
        for (int j = 0; j < 3; ++j) dloss[(size_t)j * n + i] = L.d[j];
    }
    return 0;
}

// Tangents of one operation of the dual scalar at x (and y for the two-argument ones), for the unit
// tests of its derivative rules: op 0 sin, 1 cos, 2 tan, 3 tanh, 4 sqrt, 5 fabs, 6 atan2, 7 hypot,
// 8 x / y, 9 x * y.  Returns the value; *dx, *dy receive the partial derivatives.
double orc_dual_op(int op, double x, double y, double* dx, double* dy) {
    using D = Dual<2>;
    D a(x), b(y), r;
    a.d[0] = 1.0;
    b.d[1] = 1.0;
    switch (op) {
        case 0: r = sin(a); break;
        case 1: r = cos(a); break;
        case 2: r = tan(a); break;
        case 3: r = tanh(a); break;
        case 4: r = sqrt(a); break;
        case 5: r = fabs(a); break;
        case 6: r = atan2(a, b); break;
        case 7: r = hypot(a, b); break;
        case 8: r = a / b; break;
        default: r = a * b; break;
    }
    *dx = r.d[0];
    *dy = r.d[1];
    return r.v;
}

// Counts operations over nsteps of one vehicle (same arguments as orc18_run with n = 1).
// counts: add, mul, div, sqrt, sin, cos, tan, atan2, hypot, tanh (totals over the run).
int orc18_count_ops(int nr, const double* flat, const double* init, double t0, double dt, int nsteps, int mode,
                    const double* ctrl, const int* imeta, const double* dmeta, long long* counts) {
    const int rl = d18::record_len(nr != 0);
    Controls cs{mode, ctrl, imeta, dmeta, 1};
    d18::Vehicle<CD> s;
    s.nr = nr != 0;
    d18::read_params(flat, s.nr, s.vp, s.tp);
    std::vector<double> z(rl, 0.0);
    d18::from_record(s, init ? init : z.data());
    g_ops = OpCount{};
    double t = t0;
    for (int k = 0; k < nsteps; ++k) {
        double st, th, br;
        cs.at(0, k, t, st, th, br);
        new_step();
        d18::step(s, CD(st), CD(th), CD(br), dt);
        t = t + dt;
    }
    const long long c[] = {g_ops.add, g_ops.mul,  g_ops.div,   g_ops.sqrt,  g_ops.sin,
                           g_ops.cos, g_ops.tan,  g_ops.atan2, g_ops.hypot, g_ops.tanh};
    std::memcpy(counts, c, sizeof(c));
    return 0;
}

// Same tallies for any model (dof 11, 18 or 24) from an initial record (NULL: zero state).
int orc_count_ops(int dof, int nr, const double* flat, const double* init, double t0, double dt, int nsteps, int mode,
                  const double* ctrl, const int* imeta, const double* dmeta, long long* counts) {
    auto run = [&](auto model) {
        using M = decltype(model);
        const int rl = M::rec(nr != 0);
        Controls cs{mode, ctrl, imeta, dmeta, 1};
        typename M::template V<CD> s;
        s.nr = nr != 0;
        M::params(s, flat);
        std::vector<double> z(rl, 0.0);
        M::set(s, init ? init : z.data());
        g_ops = OpCount{};
        double t = t0;
        for (int k = 0; k < nsteps; ++k) {
            double st, th, br;
            cs.at(0, k, t, st, th, br);
            new_step();
            M::step(s, CD(st), CD(th), CD(br), dt);
            t = t + dt;
        }
    };
    if (dof == 11) run(M11{});
    else if (dof == 18) run(M18{});
    else if (dof == 24) run(M24{});
    else return -1;
    const long long c[] = {g_ops.add,  g_ops.mul,  g_ops.div,   g_ops.sqrt,   g_ops.sin,  g_ops.cos,
                           g_ops.tan,  g_ops.atan2, g_ops.hypot, g_ops.tanh,  g_ops.dsin, g_ops.dcos,
                           g_ops.dtan, g_ops.datan2, g_ops.dhypot, g_ops.dtanh};
    std::memcpy(counts, c, sizeof(c));
    return 0;
}

// Tallies of single reference functions, for checks against counts made by hand from the reference
// source: computeVehRHS (dof18.cpp:1019-1078) at a record with given tire forces, and tmxy_combined
// (dof18.cpp:247-288) at one operating point.  counts as orc_count_ops.
void orc_count_veh_rhs18(const double* flat, const double* rec, const double* fx, const double* fy,
                         long long* counts) {
    d18::Vehicle<CD> s;
    d18::read_params(flat, false, s.vp, s.tp);
    d18::from_record(s, rec);
    CD fxc[4], fyc[4];
    for (int i = 0; i < 4; ++i) {
        fxc[i] = CD(fx[i]);
        fyc[i] = CD(fy[i]);
    }
    g_ops = OpCount{};
    d18::veh_rhs(s.vs, s.vp, fxc, fyc);
    counts[0] = g_ops.add; counts[1] = g_ops.mul; counts[2] = g_ops.div; counts[3] = g_ops.sqrt;
}

void orc_count_tmxy(double s, double df0, double sm, double fm, double ss, double fs, long long* counts) {
    CD f, fos;
    g_ops = OpCount{};
    d18::tmxy_combined(f, fos, CD(s), CD(df0), CD(sm), CD(fm), CD(ss), CD(fs));
    counts[0] = g_ops.add; counts[1] = g_ops.mul; counts[2] = g_ops.div; counts[3] = g_ops.sqrt;
}

// ---- building blocks, exposed for closed-form and invariant checks ----

void orc_tmxy(double s, double df0, double sm, double fm, double ss, double fs, double* f, double* fos) {
    d18::tmxy_combined(*f, *fos, s, df0, sm, fm, ss, fs);
}
double orc_interp_l(double fz, double w1, double w2, double pn) { return interp_l(fz, w1, w2, pn); }
double orc_interp_q(double fz, double w1, double w2, double pn) { return interp_q(fz, w1, w2, pn); }
double orc_sine_step(double x, double x1, double y1, double x2, double y2) { return sine_step(x, x1, y1, x2, y2); }
double orc_map_y(const double* xy, int n, double x) {
    std::vector<MapE<double>> m(n);
    for (int i = 0; i < n; ++i) m[i] = MapE<double>{xy[2 * i], xy[2 * i + 1]};
    return map_y(m, x);
}
void orc_diff_split(double torque, double sl, double sr, double* tl, double* tr) {
    d18::differential_split(torque, 2.0, sl, sr, *tl, *tr);
}
// Vertical loads of one vehicle state (record layout) for an 18-DOF parameter vector.
void orc18_loads(int nr, const double* flat, const double* rec, double* loads) {
    d18::Vehicle<double> s;
    s.nr = nr != 0;
    d18::read_params(flat, s.nr, s.vp, s.tp);
    d18::from_record(s, rec);
    d18::tire_loads(s, loads);
}
// Chassis accelerations (udot, vdot, wxdot, wzdot) for given tire forces in the vehicle frame.
void orc18_veh_rhs(int nr, const double* flat, const double* rec, const double* fx, const double* fy, double* acc) {
    d18::Vehicle<double> s;
    s.nr = nr != 0;
    d18::read_params(flat, s.nr, s.vp, s.tp);
    d18::from_record(s, rec);
    d18::veh_rhs(s.vs, s.vp, fx, fy);
    acc[0] = s.vs.udot; acc[1] = s.vs.vdot; acc[2] = s.vs.wxdot; acc[3] = s.vs.wzdot;
}

}  // extern "C"
