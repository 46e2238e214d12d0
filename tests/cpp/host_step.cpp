// Host build of the solver's step code (csrc/d{11,18,24}_device.h through the model traits of
// csrc/integrate.h), for CPU tests of its two arithmetic forms.
//
// On the host, qdiv and qsqrt are `/` and sqrt and every transcendental is glibc's, so the STRICT
// instantiation performs exactly the reference's operations in the reference's order: its results
// must equal the oracle's (and so the reference CPU solver's) bit for bit, which pins the strict
// form's formulas without a GPU.  The fast instantiation shows what the reformulations alone change.
// Built as tests/cpp -> lfdm_amd/libhost_step.so (make hoststep); loaded by tests/test_strict_host.py.
// Also exports the fast form's polynomial sine and cosine for CPU checks of their formulas.
#include <cstring>
#include <string>
#include <vector>

#include "../../low-fidelity-dynamic-models_amd/csrc/integrate.h"
#include "../../low-fidelity-dynamic-models_amd/csrc/params.h"

using namespace lfdm;

namespace {

// n vehicles of one parameter set, nsteps steps of per-step controls ctrl[nsteps][3][n] from the
// records init[n][rec]; final records into fin[n][rec].  Mirrors one launch of the integration
// kernel: all steps but the last keep only the carried state, the last one the full record.
template <class M, bool NR, bool STRICT>
void run(const PSet& P, const double* ovr, const double* init, int n, int rec, int nsteps, const double* ctrl,
         double dt, double* fin) {
    const size_t S = (size_t)n;
    std::vector<double> st((size_t)rec * S), ov((size_t)kOverrides * S);
    for (int i = 0; i < n; ++i)
        for (int f = 0; f < rec; ++f) st[(size_t)f * S + i] = init[(size_t)i * rec + f];
    const double dflt[kOverrides] = {P.m, P.jz, P.jx, P.tire.dfy0Pn, P.tire.dfy0P2n, 1.0, P.maxBrakeTorque};
    for (int f = 0; f < kOverrides; ++f)
        for (int i = 0; i < n; ++i) ov[(size_t)f * S + i] = ovr ? ovr[(size_t)f * n + i] : dflt[f];
    for (int i = 0; i < n; ++i) {
        typename M::S s;
        typename M::K k;
        M::template load<NR>(st.data(), S, i, s, k);
        M::kinit(P, ov.data() + i, S, k);
        typename M::X x;
        for (int j = 0; j < nsteps; ++j) {
            const double* c = ctrl + (size_t)j * 3 * n + i;
            const kern::Ctl cc{c[0], c[n], c[2 * n]};
            if (j + 1 < nsteps)
                M::template step<NR, false, STRICT>(P, P.d, k, s, cc, dt, nullptr);
            else
                M::template step<NR, true, STRICT>(P, P.d, k, s, cc, dt, &x);
        }
        if (nsteps > 0) M::template store<NR>(st.data(), S, i, P, s, k, x);
    }
    for (int i = 0; i < n; ++i)
        for (int f = 0; f < rec; ++f) fin[(size_t)i * rec + f] = st[(size_t)f * S + i];
}

template <class M>
void run_model(const PSet& P, int nr, int strict, const double* ovr, const double* init, int n, int rec, int nsteps,
               const double* ctrl, double dt, double* fin) {
    if (nr) {
        if (strict) run<M, true, true>(P, ovr, init, n, rec, nsteps, ctrl, dt, fin);
        else run<M, true, false>(P, ovr, init, n, rec, nsteps, ctrl, dt, fin);
    } else {
        if (strict) run<M, false, true>(P, ovr, init, n, rec, nsteps, ctrl, dt, fin);
        else run<M, false, false>(P, ovr, init, n, rec, nsteps, ctrl, dt, fin);
    }
}

}  // namespace

extern "C" {

// Returns 0, or -1 with a message in err (cap bytes) when the parameter vector does not parse.
int host_step_run(int dof, int nr, int strict, const double* flat, int len, const double* ovr, const double* init,
                  int n, int rec, int nsteps, const double* ctrl, double dt, double* fin, char* err, int cap) {
    PSet P;
    std::string e;
    if (!pset_from_flat(dof, flat, len, nr, P, e)) {
        std::snprintf(err, cap, "%s", e.c_str());
        return -1;
    }
    if (dof == 18) run_model<kern::M18>(P, nr, strict, ovr, init, n, rec, nsteps, ctrl, dt, fin);
    else if (dof == 11) run_model<kern::M11>(P, nr, strict, ovr, init, n, rec, nsteps, ctrl, dt, fin);
    else if (dof == 24) run_model<kern::M24>(P, nr, strict, ovr, init, n, rec, nsteps, ctrl, dt, fin);
    else {
        std::snprintf(err, cap, "unsupported model");
        return -1;
    }
    return 0;
}

// The fast form's sine and cosine as the host build of the step evaluates them (the polynomial code the
// device runs, with the C library's fma): out[3 n] = sincos_fast(a) sine, cosine, then sin_2pi_unit(v).
void host_fast_trig(const double* a, const double* v, int n, double* out) {
    for (int i = 0; i < n; ++i) {
        double s, c;
        dev::sincos_fast(a[i], s, c);
        out[i] = s;
        out[n + i] = c;
        out[2 * n + i] = dev::sin_2pi_unit(v[i]);
    }
}

// Word k of the 2/pi table of the large-argument reduction.
unsigned int host_two_over_pi_word(int k) { return dev::two_over_pi_word(k); }

}  // extern "C"
