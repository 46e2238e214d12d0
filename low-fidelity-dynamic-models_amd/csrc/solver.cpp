// Host side of the C ABI: solver handles, groups, device layout, launches and CSV output.
//
// Mirrors d11/d18/d24SolverHalfImplicitGPU (wheeled_vehicle_models/{11,18,24}dof-gpu/
// dof*_halfImplicit_gpu.cu) call for call.  The differences are in where things live: every Construct group becomes one
// parameter record shared by its vehicles (not one copy per vehicle), groups are padded to whole
// 64-wide wavefronts on the device, the state is SoA, and the time base is carried across
// launches exactly as the reference CPU solver accumulates it (t += dt every step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/lfdm_amd.h"
#include "kernels.h"
#include "params.h"

using namespace lfdm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

// Sets the calling thread's last-error text from other translation units of the library.
int lfdm::set_last_error(int code, const char* msg) { return fail(code, msg); }

namespace {

#define HIPCHK(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(LFDM_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int record_len(int dof, int tire_nr) {
    if (dof == 11) return tire_nr ? 2 * 11 + 12 : 2 * 15 + 12;
    if (dof == 18) return tire_nr ? 4 * 11 + 15 : 4 * 15 + 15;
    if (dof == 24) return tire_nr ? 4 * 19 + 4 * 24 + 24 : 4 * 23 + 4 * 24 + 24;
    return -1;
}

// Output channels of the device response (dof11_halfImplicit_gpu.cu:378, dof18 :372,
// dof24 :430) and their CSV column names (Write, dof11 :561-579, dof18 :583-596, dof24 :608-640).
int out_channels(int dof) { return dof == 11 ? 9 : (dof == 18 ? 13 : 15); }

const char* const kCols11[] = {"time", "x", "y", "vx", "vy", "yaw", "yaw_rate", "wf", "wr"};
const char* const kCols18[] = {"time", "x", "y", "vx", "vy", "roll", "yaw", "roll_rate", "yaw_rate",
                               "wlf", "wrf", "wlr", "wrr"};
const char* const kCols24[] = {"time", "x",        "y",        "vx",         "vy",  "roll", "yaw", "pitch",
                               "roll_rate", "yaw_rate", "pitch_rate", "wlf", "wrf",  "wlr", "wrr"};

struct Group {
    std::vector<double> flat;
    int tire_nr = 0;
    PSet pset;
    std::vector<DrvRow> rows;
    std::vector<int> users;  // vehicles of this set, ascending (a Construct call's range, less vehicles given
                             // parameters of their own, plus the vehicles given exactly this set)
};

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct lfdm_solver {
    int dof = 18;
    unsigned total = 0;
    int tire_type = -1;
    std::vector<Group> groups;
    std::vector<int> group_of;  // parameter set of every constructed vehicle
    unsigned n_params = 0;
    std::vector<double> init;  // [n_states][rec]
    unsigned n_states = 0;
    double step = 1e-3, tend = 0.0, ksim = 2.0, host_dump = 10.0;
    unsigned tpb = 32;
    // output
    bool output = false;
    std::string out_file;
    double dtout = 0.0;
    bool store_all = false;
    unsigned num_outs = 0;
    std::vector<unsigned> which_outs;
    std::vector<std::ostringstream> csv;
    // controls
    int ctrl_mode = kCtrlTable;
    int ns = 0, np = 0;
    double ds = 0, dp = 0;
    std::vector<double> knots_host;  // [ns + 2 np][total]
    bool knots_dirty = false;
    int steps_n = 0;
    int steps_pos = 0;  // next per-step control row consumed by lfdm_advance
    std::vector<double> steps_host;  // [steps_n][3][total]
    const double* steps_ext = nullptr;
    bool steps_shared = false;  // steps_ext is [steps_n][3] for the whole fleet
    bool steps_dirty = false;
    std::vector<double> ovr_host;  // [7][total]
    bool ovr_dirty = false;
    bool ovr_device_newer = false;  // lfdm_set_overrides_device wrote the device copy after ovr_host
    double t_now = 0.0;
    // per-vehicle horizons (steps from t = 0; empty: none) and the slot layout
    std::vector<int> horizon_host;
    bool horizon_dirty = false;
    int layout = LFDM_LAYOUT_GROUPED;
    int arith = LFDM_ARITH_FAST;  // arithmetic of the integration step
    // device
    bool built = false;
    int S = 0, rec = 0;
    std::vector<int> user_slot;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool stream_set = false;  // the caller chose `stream` (NULL = the default stream)
    PSet* d_sets = nullptr;
    int* d_wave_set = nullptr;
    int* d_slot_user = nullptr;
    int* d_user_slot = nullptr;
    double* d_state = nullptr;
    double* d_ovr = nullptr;
    int* d_horizon = nullptr;    // [S]
    int* d_wave_hmax = nullptr;  // [S / 64]
    DrvRow* d_drv = nullptr;
    double* d_knots = nullptr;
    double* d_steps = nullptr;
    size_t steps_cap = 0;
    int* d_out_steps = nullptr;
    size_t out_steps_cap = 0;
    double* d_out = nullptr;
    size_t out_cap = 0;
    // sampled-vehicle output: the columns of the written vehicles, gathered on the device
    double* d_outc = nullptr;
    size_t outc_cap = 0;
    int* d_which = nullptr;
    double* d_rec = nullptr;
    size_t rec_cap = 0;
    std::vector<int> outs_host;  // output steps of the launch in flight (source of an async copy)
    // differentiation form
    double* d_ckpt = nullptr;
    size_t ckpt_cap = 0;
    double* d_tape = nullptr;
    size_t tape_cap = 0;
    double* d_carry = nullptr;
    size_t carry_cap = 0;
    int adj_nsteps = 0, adj_K = 0;
    unsigned long long adj_gen = 0;  // adjoint forwards run on this handle (and pending ones invalidated)
    int adj_G = 1;                 // segments per backward group, fixed by the forward
    bool tape_holds_last = false;  // the tape holds the last group, written by the forward
    size_t adj_tape_budget = 0;  // bytes the backward's tape may take (0: automatic, adj_budget)
    int simds = 0;  // SIMDs of the device (4 per CU), queried once
    const double* adj_ctrl = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    double last_ms = 0.0;
};

// A mixed fleet: solver handles (buckets, typically one per model) integrated together by one HIP
// graph, one branch per bucket.
struct lfdm_fleet {
    struct Cfg {
        int block = 256;
        int spl = 0;    // steps per launch, 0 = the whole run in one launch
        int waves = 0;  // waves per SIMD of the kernel, 0 = the solver's rule
    };
    std::vector<lfdm_solver*> buckets;
    std::vector<Cfg> cfg;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool stream_set = false;  // the caller chose `stream` (NULL = the default stream)
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> pre, join;
    hipEvent_t fork = nullptr, post = nullptr, ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<char> sig;  // launch records and block sizes the graph was captured with
    bool timed = false;
};

namespace {

int check(lfdm_solver* h) {
    if (!h) return fail(LFDM_E_ARG, "null solver handle");
    return LFDM_OK;
}

// A pending adjoint (forward run, backward not yet) reads the per-vehicle parameters again: changing
// them makes it unusable.
void invalidate_adjoint(lfdm_solver* h) {
    if (h->adj_nsteps > 0) {
        h->adj_nsteps = 0;
        ++h->adj_gen;
    }
}

int ensure_buffer(double*& p, size_t& cap, size_t n) {
    if (n <= cap && p) return LFDM_OK;
    dfree(p);
    HIPCHK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(double)));
    cap = n;
    return LFDM_OK;
}

int ensure_stream(lfdm_solver* h) {
    if (!h->ev0) {
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || ndev == 0)
            return fail(LFDM_E_HIP, "no HIP device available: the solver runs only on an AMD GPU");
    }
    if (!h->stream_set && !h->own_stream) {
        HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->own_stream = true;
    }
    if (!h->ev0) {
        HIPCHK(hipEventCreate(&h->ev0));
        HIPCHK(hipEventCreate(&h->ev1));
    }
    return LFDM_OK;
}

// Builds the device image of the fleet on first use.
int ensure_device(lfdm_solver* h) {
    if (h->built) return LFDM_OK;
    if (h->n_params != h->total)
        return fail(LFDM_E_STATE, "Construct covered " + std::to_string(h->n_params) + " of " +
                                      std::to_string(h->total) + " vehicles");
    if (h->n_states != h->total)
        return fail(LFDM_E_STATE, "Initialize covered " + std::to_string(h->n_states) + " of " +
                                      std::to_string(h->total) + " vehicles");
    int rc = ensure_stream(h);
    if (rc) return rc;
    const int rec = h->rec;
    // slots: every group padded to whole wavefronts
    int S = 0;
    std::vector<int> slot_base(h->groups.size());
    for (size_t g = 0; g < h->groups.size(); ++g) {
        slot_base[g] = S;
        S += (int)((h->groups[g].users.size() + kWave - 1) / kWave) * kWave;
    }
    h->S = S;
    std::vector<int> slot_user(S, -1), wave_set(S / kWave, 0);
    h->user_slot.assign(h->total, 0);
    std::vector<PSet> sets(h->groups.size());
    std::vector<DrvRow> pool;
    const bool by_horizon = h->layout == LFDM_LAYOUT_BY_HORIZON && !h->horizon_host.empty();
    for (size_t g = 0; g < h->groups.size(); ++g) {
        const Group& G = h->groups[g];
        std::vector<int> order(G.users);
        if (by_horizon)  // longest horizon first: lanes of a wavefront stop together, long wavefronts start first
            std::stable_sort(order.begin(), order.end(),
                             [&](int a, int b) { return h->horizon_host[a] > h->horizon_host[b]; });
        for (size_t i = 0; i < order.size(); ++i) {
            slot_user[slot_base[g] + i] = order[i];
            h->user_slot[order[i]] = slot_base[g] + (int)i;
        }
        const int waves = (int)((order.size() + kWave - 1) / kWave);
        for (int w = 0; w < waves; ++w) wave_set[slot_base[g] / kWave + w] = (int)g;
        sets[g] = G.pset;
        sets[g].drv_off = (int)pool.size();
        sets[g].drv_rows = (int)G.rows.size();
        pool.insert(pool.end(), G.rows.begin(), G.rows.end());
    }
    if (pool.empty()) pool.push_back(DrvRow{0, 0, 0, 0});
    // state SoA; padding slots take their group's first vehicle
    std::vector<double> st((size_t)rec * S, 0.0);
    for (int s = 0; s < S; ++s) {
        int u = slot_user[s];
        if (u < 0) {
            int g = wave_set[s / kWave];
            u = h->groups[g].users[0];
        }
        for (int f = 0; f < rec; ++f) st[(size_t)f * S + s] = h->init[(size_t)u * rec + f];
    }
    // per-vehicle parameters: group values unless overridden
    std::vector<double> ovr((size_t)kOverrides * S, 0.0);
    for (int s = 0; s < S; ++s) {
        const PSet& p = sets[wave_set[s / kWave]];
        const double dflt[kOverrides] = {p.m, p.jz, p.jx, p.tire.dfy0Pn, p.tire.dfy0P2n, 1.0, p.maxBrakeTorque};
        for (int f = 0; f < kOverrides; ++f) ovr[(size_t)f * S + s] = dflt[f];
        const int u = slot_user[s];
        if (u >= 0 && !h->ovr_host.empty())
            for (int f = 0; f < kOverrides; ++f) ovr[(size_t)f * S + s] = h->ovr_host[(size_t)f * h->total + u];
    }
    HIPCHK(hipMalloc((void**)&h->d_sets, sets.size() * sizeof(PSet)));
    HIPCHK(hipMalloc((void**)&h->d_wave_set, wave_set.size() * sizeof(int)));
    HIPCHK(hipMalloc((void**)&h->d_slot_user, (size_t)S * sizeof(int)));
    HIPCHK(hipMalloc((void**)&h->d_user_slot, (size_t)h->total * sizeof(int)));
    HIPCHK(hipMalloc((void**)&h->d_state, st.size() * sizeof(double)));
    HIPCHK(hipMalloc((void**)&h->d_ovr, ovr.size() * sizeof(double)));
    HIPCHK(hipMalloc((void**)&h->d_drv, pool.size() * sizeof(DrvRow)));
    HIPCHK(hipMemcpyAsync(h->d_sets, sets.data(), sets.size() * sizeof(PSet), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_wave_set, wave_set.data(), wave_set.size() * sizeof(int), hipMemcpyHostToDevice,
                          h->stream));
    HIPCHK(hipMemcpyAsync(h->d_slot_user, slot_user.data(), (size_t)S * sizeof(int), hipMemcpyHostToDevice,
                          h->stream));
    HIPCHK(hipMemcpyAsync(h->d_user_slot, h->user_slot.data(), (size_t)h->total * sizeof(int),
                          hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_state, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_ovr, ovr.data(), ovr.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_drv, pool.data(), pool.size() * sizeof(DrvRow), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->built = true;
    h->ovr_dirty = false;
    h->horizon_dirty = true;
    return LFDM_OK;
}

// Drops the device image (keeping the current states as the records to rebuild it from), so the
// next call lays the slots out again.
int drop_device_image(lfdm_solver* h) {
    if (!h->built) return LFDM_OK;
    std::vector<double> cur((size_t)h->total * h->rec);
    if (int rc = lfdm_get_states(h, cur.data())) return rc;
    h->init.swap(cur);
    h->n_states = h->total;
    if (h->ovr_device_newer) {  // the device copy is the current one: keep it for the rebuild
        std::vector<double> ovr((size_t)kOverrides * h->S);
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(ovr.data(), h->d_ovr, ovr.size() * sizeof(double), hipMemcpyDeviceToHost));
        h->ovr_host.assign((size_t)kOverrides * h->total, 0.0);
        for (unsigned u = 0; u < h->total; ++u)
            for (int f = 0; f < kOverrides; ++f)
                h->ovr_host[(size_t)f * h->total + u] = ovr[(size_t)f * h->S + h->user_slot[u]];
        h->ovr_device_newer = false;
    }
    dfree(h->d_sets);
    dfree(h->d_wave_set);
    dfree(h->d_slot_user);
    dfree(h->d_user_slot);
    dfree(h->d_state);
    dfree(h->d_ovr);
    dfree(h->d_drv);
    dfree(h->d_horizon);
    dfree(h->d_wave_hmax);
    h->built = false;
    invalidate_adjoint(h);  // a pending adjoint refers to the old slots
    h->tape_holds_last = false;
    return LFDM_OK;
}

int sync_inputs(lfdm_solver* h) {
    if (h->ovr_dirty) {
        // rebuild the slot-ordered override array from the host copy
        std::vector<double> ovr((size_t)kOverrides * h->S);
        std::vector<int> slot_user(h->S, -1);
        for (unsigned u = 0; u < h->total; ++u) slot_user[h->user_slot[u]] = (int)u;
        HIPCHK(hipMemcpy(ovr.data(), h->d_ovr, ovr.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int s = 0; s < h->S; ++s)
            if (slot_user[s] >= 0)
                for (int f = 0; f < kOverrides; ++f)
                    ovr[(size_t)f * h->S + s] = h->ovr_host[(size_t)f * h->total + slot_user[s]];
        HIPCHK(hipMemcpyAsync(h->d_ovr, ovr.data(), ovr.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        h->ovr_dirty = false;
    }
    if (h->horizon_dirty) {
        dfree(h->d_horizon);
        dfree(h->d_wave_hmax);
        if (!h->horizon_host.empty()) {
            std::vector<int> hs(h->S, 0), wm(h->S / kWave, 0);
            for (unsigned u = 0; u < h->total; ++u) {
                const int slot = h->user_slot[u];
                hs[slot] = h->horizon_host[u];
                wm[slot / kWave] = std::max(wm[slot / kWave], hs[slot]);
            }
            HIPCHK(hipMalloc((void**)&h->d_horizon, hs.size() * sizeof(int)));
            HIPCHK(hipMalloc((void**)&h->d_wave_hmax, wm.size() * sizeof(int)));
            HIPCHK(hipMemcpyAsync(h->d_horizon, hs.data(), hs.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipMemcpyAsync(h->d_wave_hmax, wm.data(), wm.size() * sizeof(int), hipMemcpyHostToDevice,
                                  h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        }
        h->horizon_dirty = false;
    }
    if (h->knots_dirty) {
        dfree(h->d_knots);
        HIPCHK(hipMalloc((void**)&h->d_knots, h->knots_host.size() * sizeof(double)));
        HIPCHK(hipMemcpyAsync(h->d_knots, h->knots_host.data(), h->knots_host.size() * sizeof(double),
                              hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        h->knots_dirty = false;
    }
    if (h->steps_dirty) {
        int rc = ensure_buffer(h->d_steps, h->steps_cap, h->steps_host.size());
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(h->d_steps, h->steps_host.data(), h->steps_host.size() * sizeof(double),
                              hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        h->steps_dirty = false;
    }
    return LFDM_OK;
}

// Step count of one launch of the reference kernel: t from t0 while t < t0 + ksim - step/10
// (dof18_halfImplicit_gpu.cu:892-948).
int chunk_steps(double t0, double ksim, double step) {
    double t = t0;
    const double end = (t + ksim) - step / 10.;
    int n = 0;
    while (t < end) {
        t += step;
        ++n;
    }
    return n;
}

// Steps after which the reference kernel writes an output row: kernel_time accumulates from 0 and
// a row is written when |kernel_time - (stored + 1) dtout| < 1e-7 (.cu:950-957).
std::vector<int> output_steps(int nsteps, double step, double dtout) {
    std::vector<int> r;
    double kt = 0;
    unsigned stored = 0;
    for (int k = 0; k < nsteps; ++k) {
        kt += step;
        if (std::abs(kt - (stored + 1) * dtout) < 1e-7) {
            r.push_back(k + 1);
            ++stored;
        }
    }
    return r;
}

// SIMDs of the current device (4 per CU), queried once per handle.
int max_waves(int dof);

int simd_count(lfdm_solver* h) {
    if (h->simds > 0) return h->simds;
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    h->simds = 4 * std::max(cus, 1);
    return h->simds;
}

// Launch record of one integration launch of h (no outputs).  Zero-filled first, so records can be
// compared bytewise.
Launch make_launch(lfdm_solver* h, int mode, double t0, int nsteps, double cs = 0, double ct = 0, double cb = 0,
                   const double* steps_ptr = nullptr) {
    Launch L;
    std::memset(&L, 0, sizeof(L));
    L.sets = h->d_sets;
    L.wave_set = h->d_wave_set;
    L.slot_user = h->d_slot_user;
    L.state = h->d_state;
    L.ovr = h->d_ovr;
    L.drv = h->d_drv;
    L.knots = h->d_knots;
    L.steps = steps_ptr;
    L.t0 = t0;
    L.dt = h->step;
    L.c_s = cs;
    L.c_t = ct;
    L.c_b = cb;
    L.ds = h->ds;
    L.dp = h->dp;
    L.ns = h->ns;
    L.np = h->np;
    L.nsteps = nsteps;
    L.S = h->S;
    L.n_user = (int)h->total;
    L.mode = mode;
    L.nr = h->tire_type == LFDM_TIRE_TMEASYNR;
    L.steps_shared = (mode == kCtrlSteps && h->steps_shared) ? 1 : 0;
    if (h->d_horizon) {
        L.horizon = h->d_horizon;
        L.wave_hmax = h->d_wave_hmax;
        L.step_base = (int)std::llround(t0 / h->step);
    }
    // a fleet of at most w wavefronts per SIMD, w below the model's default, runs the kernels built for
    // w waves per SIMD: their larger register budget spills less, and no SIMD can take more of the
    // fleet's wavefronts than w (at two per SIMD, the 18-DOF kernel built for three ran 1.3x slower)
    const int simds = simd_count(h);
    const long long per = simds > 0 ? ((long long)h->S / kWave + simds - 1) / simds : 0;
    L.waves = (per > 0 && per < max_waves(h->dof)) ? (int)per : 0;
    L.strict = h->arith == LFDM_ARITH_STRICT ? 1 : 0;
    return L;
}

int block_size(unsigned tpb) {
    const int block = (int)((tpb + kWave - 1) / kWave) * kWave;
    return std::min(256, std::max(64, block));
}

hipError_t launch_dof(int dof, const Launch& L, int block, hipStream_t stream) {
    return dof == 11 ? launch_d11(L, block, stream) : dof == 18 ? launch_d18(L, block, stream) : launch_d24(L, block, stream);
}

int launch(lfdm_solver* h, int mode, double t0, int nsteps, const std::vector<int>* outs, double cs = 0,
           double ct = 0, double cb = 0, const double* steps_ptr = nullptr) {
    Launch L = make_launch(h, mode, t0, nsteps, cs, ct, cb, steps_ptr);
    if (outs && !outs->empty()) {
        h->outs_host = *outs;
        outs = &h->outs_host;
        if (h->out_steps_cap < outs->size()) {
            if (h->d_out_steps) (void)hipFree(h->d_out_steps);
            HIPCHK(hipMalloc((void**)&h->d_out_steps, outs->size() * sizeof(int)));
            h->out_steps_cap = outs->size();
        }
        HIPCHK(hipMemcpyAsync(h->d_out_steps, outs->data(), outs->size() * sizeof(int), hipMemcpyHostToDevice,
                              h->stream));
        L.out_steps = h->d_out_steps;
        L.n_out = (int)outs->size();
        L.out = h->d_out;
    }
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    hipError_t e = launch_dof(h->dof, L, block_size(h->tpb), h->stream);
    if (e != hipSuccess) return fail(LFDM_E_HIP, std::string("integration launch failed: ") + hipGetErrorString(e));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    h->timed = true;
    return LFDM_OK;
}

// ---- CSV output, as d18SolverHalfImplicitGPU::Write / WriteToFile (.cu:552-663) ----

void csv_header(lfdm_solver* h) {
    const unsigned nw = h->store_all ? h->total : h->num_outs;
    const int nc = out_channels(h->dof);
    const char* const* cols = h->dof == 11 ? kCols11 : (h->dof == 18 ? kCols18 : kCols24);
    for (unsigned i = 0; i < nw; ++i) {
        std::ostringstream& o = h->csv[i];
        for (int c = 0; c < nc; ++c) o << cols[c] << ",";
        o << std::endl;
        for (int c = 0; c < nc; ++c) o << 0 << ",";
        o << std::endl;
    }
}

// Appends `rows` output samples to the CSV streams.  With every vehicle written, `host` holds
// [sample][channel][vehicle] rows; with sampled vehicles it holds, per launch of `dct` samples,
// [written vehicle][sample][channel] (the device gathers only their columns).
void csv_rows(lfdm_solver* h, const std::vector<double>& host, unsigned rows, unsigned dct) {
    const unsigned nw = h->store_all ? h->total : h->num_outs;
    const size_t N = h->total;
    const int nc = out_channels(h->dof);
    for (unsigned i = 0; i < nw; ++i) {
        std::ostringstream& o = h->csv[i];
        for (unsigned r = 0; r < rows; ++r) {
            const size_t off = h->store_all ? (size_t)r * N * nc + i
                                            : (((size_t)(r / dct) * nw + i) * dct + r % dct) * nc;
            const size_t stride = h->store_all ? N : 1;
            for (int c = 0; c < nc; ++c) o << host[off + stride * c] << ",";
            o << std::endl;
        }
    }
}

int csv_flush(lfdm_solver* h) {
    const unsigned nw = h->store_all ? h->total : h->num_outs;
    for (unsigned i = 0; i < nw; ++i) {
        const unsigned v = h->store_all ? i : h->which_outs[i];
        const std::string name = h->out_file + "_" + std::to_string(v) + ".csv";
        std::ofstream f(name.c_str());
        if (!f) return fail(LFDM_E_IO, "cannot write " + name);
        f << h->csv[i].str();
        h->csv[i].str("");
        h->csv[i].clear();
    }
    return LFDM_OK;
}

// Control source of h ready for nsteps steps from step 0 (what Solve checks).
int check_controls(lfdm_solver* h, int nsteps) {
    if (h->ctrl_mode == kCtrlTable)
        for (const Group& g : h->groups)
            if (g.rows.empty())
                return fail(LFDM_E_STATE, "No controls provided, please use construct to pass path to driver inputs");
    if (h->ctrl_mode == kCtrlSteps && nsteps > h->steps_n)
        return fail(LFDM_E_STATE, "per-step controls cover " + std::to_string(h->steps_n) + " steps, " +
                                      std::to_string(nsteps) + " requested");
    return LFDM_OK;
}

// Launch records that integrate h nsteps steps from t = 0 (step 0) in launches of at most spl
// steps, the time accumulated step by step as Solve does.  Ends with the final time in *t_end.
std::vector<Launch> run_launches(lfdm_solver* h, int nsteps, int spl, int waves, double* t_end) {
    const double* steps = h->ctrl_mode == kCtrlSteps ? (h->steps_ext ? h->steps_ext : h->d_steps) : nullptr;
    const size_t row = h->steps_shared ? 3 : 3 * (size_t)h->total;
    std::vector<Launch> r;
    double t = 0.0;
    for (int done = 0; done < nsteps;) {
        const int n = std::min(spl > 0 ? spl : nsteps, nsteps - done);
        Launch L = make_launch(h, h->ctrl_mode, t, n, 0, 0, 0, steps ? steps + (size_t)done * row : nullptr);
        if (L.horizon) L.step_base = done;
        if (waves > 0) L.waves = waves;
        r.push_back(L);
        for (int k = 0; k < n; ++k) t += h->step;
        done += n;
    }
    if (t_end) *t_end = t;
    return r;
}

// Lane-steps a layout would run: sum over wavefronts of 64 x (the wavefront's longest horizon,
// capped at nsteps); vehicles without horizons run nsteps.
double predicted_lane_steps(const lfdm_solver* h, int layout, int nsteps) {
    double sum = 0;
    for (const Group& G : h->groups) {
        std::vector<int> hz(G.users.size(), nsteps);
        if (!h->horizon_host.empty())
            for (size_t i = 0; i < G.users.size(); ++i) hz[i] = std::min(h->horizon_host[G.users[i]], nsteps);
        if (layout == LFDM_LAYOUT_BY_HORIZON) std::stable_sort(hz.begin(), hz.end(), std::greater<int>());
        for (size_t w = 0; w < hz.size(); w += kWave) {
            int m = 0;
            for (size_t i = w; i < std::min(hz.size(), w + kWave); ++i) m = std::max(m, hz[i]);
            sum += (double)kWave * m;
        }
    }
    return sum;
}

}  // namespace

extern "C" {

const char* lfdm_last_error(void) { return g_err.c_str(); }

int lfdm_record_len(int dof, int tire_type) {
    const int r = record_len(dof, tire_type == LFDM_TIRE_TMEASYNR);
    if (r < 0) return fail(LFDM_E_UNSUPPORTED, "unsupported model: " + std::to_string(dof) + "-DOF");
    return r;
}

int lfdm_params_from_json(int dof, const char* veh, const char* tire, const char* sus, int tire_type, double* out,
                          int cap, int* len) {
    if (!veh || !tire || !len) return fail(LFDM_E_ARG, "null argument");
    if (tire_type != LFDM_TIRE_TMEASY && tire_type != LFDM_TIRE_TMEASYNR) return fail(LFDM_E_ARG, "bad tire type");
    if (dof != 11 && dof != 18 && dof != 24)
        return fail(LFDM_E_UNSUPPORTED, "unsupported model: " + std::to_string(dof) + "-DOF");
    if (dof == 24 && !sus) return fail(LFDM_E_ARG, "the 24-DOF model needs a suspension JSON file");
    std::vector<double> flat;
    std::string err;
    const bool ok = params_from_json(dof, veh, tire, sus ? sus : "", tire_type, flat, err);
    if (!ok) return fail(err.rfind("cannot open", 0) == 0 ? LFDM_E_IO : LFDM_E_PARSE, err);
    *len = (int)flat.size();
    if (!out || cap < (int)flat.size()) return fail(LFDM_E_CAPACITY, "output buffer too small");
    std::copy(flat.begin(), flat.end(), out);
    return LFDM_OK;
}

int lfdm_driver_file_load(const char* path, double* rows, int cap_rows, int* nrows) {
    if (!path || !nrows) return fail(LFDM_E_ARG, "null argument");
    std::vector<DrvRow> r;
    std::string err;
    if (!load_driver_file(path, r, err)) return fail(LFDM_E_IO, err);
    *nrows = (int)r.size();
    if (!rows || cap_rows < (int)r.size()) return fail(LFDM_E_CAPACITY, "row buffer too small");
    std::memcpy(rows, r.data(), r.size() * sizeof(DrvRow));
    return LFDM_OK;
}

int lfdm_create(int dof, unsigned total, lfdm_solver** out) {
    if (!out) return fail(LFDM_E_ARG, "null output handle");
    *out = nullptr;
    if (dof != 11 && dof != 18 && dof != 24)
        return fail(LFDM_E_UNSUPPORTED, "unsupported model: " + std::to_string(dof) + "-DOF");
    if (total == 0) return fail(LFDM_E_ARG, "total vehicle count must be positive");
    lfdm_solver* h = new lfdm_solver();
    h->dof = dof;
    h->total = total;
    *out = h;
    return LFDM_OK;
}

int lfdm_destroy(lfdm_solver* h) {
    if (!h) return LFDM_OK;
    if (h->stream_set || h->own_stream) (void)hipStreamSynchronize(h->stream);
    dfree(h->d_sets);
    dfree(h->d_wave_set);
    dfree(h->d_slot_user);
    dfree(h->d_user_slot);
    dfree(h->d_state);
    dfree(h->d_ovr);
    dfree(h->d_drv);
    dfree(h->d_knots);
    dfree(h->d_steps);
    dfree(h->d_out_steps);
    dfree(h->d_out);
    dfree(h->d_outc);
    dfree(h->d_which);
    dfree(h->d_rec);
    dfree(h->d_horizon);
    dfree(h->d_wave_hmax);
    dfree(h->d_ckpt);
    dfree(h->d_tape);
    dfree(h->d_carry);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return LFDM_OK;
}

int lfdm_construct_flat(lfdm_solver* h, const double* flat, int len, int tire_type, unsigned n, const double* rows,
                        int nrows) {
    if (int rc = check(h)) return rc;
    if (!flat || len <= 0) return fail(LFDM_E_ARG, "empty parameter vector");
    if (tire_type != LFDM_TIRE_TMEASY && tire_type != LFDM_TIRE_TMEASYNR) return fail(LFDM_E_ARG, "bad tire type");
    if (h->built) return fail(LFDM_E_STATE, "Construct after the fleet is on the device");
    if (h->tire_type >= 0 && h->tire_type != tire_type)
        return fail(LFDM_E_ARG, "every group of a solver must use the same tire type");
    if ((unsigned long long)n + h->n_params > h->total)
        return fail(LFDM_E_CAPACITY, "Construct exceeds the total vehicle count of the solver");
    if (n == 0) return fail(LFDM_E_ARG, "vehicle count must be positive");
    Group g;
    g.flat.assign(flat, flat + len);
    g.tire_nr = tire_type;
    std::string err;
    if (!pset_from_flat(h->dof, flat, len, tire_type, g.pset, err)) return fail(LFDM_E_ARG, err);
    if (rows && nrows > 0) g.rows.assign(reinterpret_cast<const DrvRow*>(rows), reinterpret_cast<const DrvRow*>(rows) + nrows);
    g.users.resize(n);
    std::iota(g.users.begin(), g.users.end(), (int)h->n_params);
    h->group_of.resize(h->n_params + n, (int)h->groups.size());
    h->groups.push_back(std::move(g));
    h->n_params += n;
    h->tire_type = tire_type;
    h->rec = record_len(h->dof, tire_type);
    return LFDM_OK;
}

int lfdm_construct(lfdm_solver* h, const char* veh, const char* tire, const char* sus, unsigned n,
                   const char* driver_file, int tire_type) {
    if (int rc = check(h)) return rc;
    double buf[4096];
    int len = 0;
    int rc = lfdm_params_from_json(h->dof, veh, tire, sus, tire_type, buf, 4096, &len);
    if (rc) return rc;
    std::vector<DrvRow> rows;
    if (driver_file && driver_file[0]) {
        std::string err;
        if (!load_driver_file(driver_file, rows, err)) return fail(LFDM_E_IO, err);
    }
    return lfdm_construct_flat(h, buf, len, tire_type, n, rows.empty() ? nullptr : &rows[0].t, (int)rows.size());
}

namespace {

// Appends n vehicles whose records are rec(i) (stride 0: one record for all, as Initialize).
int append_states(lfdm_solver* h, const double* records, unsigned n, size_t stride) {
    if (!records) return fail(LFDM_E_ARG, "null record");
    if (h->tire_type < 0) return fail(LFDM_E_STATE, "Initialize before Construct");
    if ((unsigned long long)n + h->n_states > h->total)
        return fail(LFDM_E_CAPACITY, "Initialize exceeds the total vehicle count of the solver");
    for (unsigned i = 0; i < (stride ? n : 1u); ++i)
        for (int f = 0; f < h->rec; ++f)
            if (!std::isfinite(records[i * stride + f])) return fail(LFDM_E_ARG, "initial state holds a non-finite value");
    const unsigned first = h->n_states;
    if (h->dof == 24 && (unsigned long long)first + n > h->n_params)
        return fail(LFDM_E_STATE, "24-DOF Initialize needs its vehicles constructed first");
    for (unsigned i = 0; i < n; ++i) h->init.insert(h->init.end(), records + i * stride, records + i * stride + h->rec);
    h->n_states += n;
    if (h->dof == 24) {
        // initializeTireSus per vehicle (dof24_halfImplicit_gpu.cu:320-333), with the vehicle's own
        // mass when per-vehicle parameters were set before Initialize
        for (unsigned i = first; i < first + n; ++i) {
            const Group& G = h->groups[h->group_of[i]];
            const double m = h->ovr_host.empty() ? G.pset.m : h->ovr_host[i];
            init_tire_sus24(G.pset, m, h->tire_type == LFDM_TIRE_TMEASYNR, h->init.data() + (size_t)i * h->rec);
        }
    }
    if (h->built) {
        // the fleet is already on the device: write these records into their slots
        std::vector<int> slots(n);
        for (unsigned i = 0; i < n; ++i) slots[i] = h->user_slot[first + i];
        size_t need = (size_t)n * h->rec;
        if (int rc = ensure_buffer(h->d_rec, h->rec_cap, need + n)) return rc;
        HIPCHK(hipMemcpyAsync(h->d_rec, h->init.data() + (size_t)first * h->rec, need * sizeof(double),
                              hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->d_rec + need, slots.data(), n * sizeof(int), hipMemcpyHostToDevice, h->stream));
        hipError_t e = launch_scatter(h->d_state, h->S, h->rec, reinterpret_cast<int*>(h->d_rec + need), (int)n,
                                      h->d_rec, h->stream);
        if (e != hipSuccess) return fail(LFDM_E_HIP, hipGetErrorString(e));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return LFDM_OK;
}

}  // namespace

int lfdm_initialize(lfdm_solver* h, const double* record, unsigned n) {
    if (int rc = check(h)) return rc;
    return append_states(h, record, n, 0);
}

int lfdm_initialize_batch(lfdm_solver* h, const double* records, unsigned n) {
    if (int rc = check(h)) return rc;
    return append_states(h, records, n, (size_t)h->rec);
}

int lfdm_set_end_time(lfdm_solver* h, double v) {
    if (int rc = check(h)) return rc;
    if (!(v > 0) || !std::isfinite(v)) return fail(LFDM_E_ARG, "end time must be positive");
    h->tend = v;
    return LFDM_OK;
}
int lfdm_set_time_step(lfdm_solver* h, double v) {
    if (int rc = check(h)) return rc;
    if (!(v > 0) || !std::isfinite(v)) return fail(LFDM_E_ARG, "time step must be positive");
    h->step = v;
    return LFDM_OK;
}
int lfdm_set_kernel_sim_time(lfdm_solver* h, double v) {
    if (int rc = check(h)) return rc;
    if (!(v > 0) || !std::isfinite(v)) return fail(LFDM_E_ARG, "kernel sim time must be positive");
    h->ksim = v;
    return LFDM_OK;
}
int lfdm_set_host_dump_time(lfdm_solver* h, double v) {
    if (int rc = check(h)) return rc;
    if (!(v > 0) || !std::isfinite(v)) return fail(LFDM_E_ARG, "host dump time must be positive");
    h->host_dump = v;
    return LFDM_OK;
}
int lfdm_set_threads_per_block(lfdm_solver* h, unsigned v) {
    if (int rc = check(h)) return rc;
    if (v == 0 || v > 1024) return fail(LFDM_E_ARG, "threads per block must be 1..1024");
    h->tpb = v;
    return LFDM_OK;
}
int lfdm_get_step(lfdm_solver* h, double* v) {
    if (int rc = check(h)) return rc;
    if (!v) return fail(LFDM_E_ARG, "null output");
    *v = h->step;
    return LFDM_OK;
}
int lfdm_get_end_time(lfdm_solver* h, double* v) {
    if (int rc = check(h)) return rc;
    if (!v) return fail(LFDM_E_ARG, "null output");
    *v = h->tend;
    return LFDM_OK;
}
int lfdm_get_kernel_sim_time(lfdm_solver* h, double* v) {
    if (int rc = check(h)) return rc;
    if (!v) return fail(LFDM_E_ARG, "null output");
    *v = h->ksim;
    return LFDM_OK;
}
int lfdm_get_tire_type(lfdm_solver* h, int* v) {
    if (int rc = check(h)) return rc;
    if (!v) return fail(LFDM_E_ARG, "null output");
    *v = h->tire_type < 0 ? LFDM_TIRE_TMEASY : h->tire_type;
    return LFDM_OK;
}

int lfdm_set_output(lfdm_solver* h, const char* file, double freq, int store_all, unsigned no_outs) {
    if (int rc = check(h)) return rc;
    if (!file) return fail(LFDM_E_ARG, "null output file");
    if (!(freq > 0) || !std::isfinite(freq)) return fail(LFDM_E_ARG, "output frequency must be positive");
    h->output = true;
    h->store_all = store_all != 0;
    if (!h->store_all) {
        if (no_outs > h->total) no_outs = h->total;  // the reference warns and clamps (.cu:350-356)
        h->num_outs = no_outs;
        // vehicles chosen as the reference does: shuffle of 0..N-1 with mt19937 seeded 68 (.cu:359-369)
        float some_seed = 68;
        std::mt19937 generator(some_seed);
        std::vector<int> numbers(h->total);
        std::iota(numbers.begin(), numbers.end(), 0);
        std::shuffle(numbers.begin(), numbers.end(), generator);
        h->which_outs.assign(numbers.begin(), numbers.begin() + no_outs);
    } else {
        h->num_outs = h->total;
    }
    h->csv = std::vector<std::ostringstream>(h->num_outs);
    h->out_file = file;
    h->dtout = 1.0 / freq;
    return LFDM_OK;
}

int lfdm_solve(lfdm_solver* h) {
    if (int rc = check(h)) return rc;
    if (h->tend == 0.) return fail(LFDM_E_STATE, "Final time not set, please use SetEndTime function");
    if (int rc = ensure_device(h)) return rc;
    if (int rc = sync_inputs(h)) return rc;
    const int mode = h->ctrl_mode;
    if (mode == kCtrlTable)
        for (const Group& g : h->groups)
            if (g.rows.empty())
                return fail(LFDM_E_STATE, "No controls provided, please use construct to pass path to driver inputs");
    const double* steps = h->steps_ext ? h->steps_ext : h->d_steps;
    const size_t N = h->total;
    const size_t nc = (size_t)out_channels(h->dof);
    unsigned dct = 0, hct = 0;
    std::vector<double> host;
    // columns copied to the host per output sample: every vehicle, or only the written ones
    const size_t W = h->store_all ? N : (size_t)h->num_outs;
    if (h->output) {
        dct = (unsigned)std::ceil(h->ksim / h->dtout);
        hct = (unsigned)std::ceil(h->host_dump / h->dtout);
        if (int rc = ensure_buffer(h->d_out, h->out_cap, (size_t)dct * nc * N)) return rc;
        HIPCHK(hipMemsetAsync(h->d_out, 0, (size_t)dct * nc * N * sizeof(double), h->stream));
        if (!h->store_all) {
            if (int rc = ensure_buffer(h->d_outc, h->outc_cap, (size_t)dct * nc * W)) return rc;
            std::vector<int> which(h->which_outs.begin(), h->which_outs.end());
            dfree(h->d_which);
            HIPCHK(hipMalloc((void**)&h->d_which, std::max<size_t>(W, 1) * sizeof(int)));
            HIPCHK(hipMemcpyAsync(h->d_which, which.data(), W * sizeof(int), hipMemcpyHostToDevice, h->stream));
        }
        host.assign((size_t)std::max(hct, dct) * nc * W, 0.0);
        for (auto& s : h->csv) { s.str(""); s.clear(); }
        csv_header(h);
    }
    double current = 0.0;
    double t_cont = 0.0;
    int step_index = 0;
    unsigned launches = 0;
    double since = 0.0;
    while (current < h->tend) {
        const double kend = current + h->ksim;
        const int n = chunk_steps(current, h->ksim, h->step);
        if (mode == kCtrlSteps && step_index + n > h->steps_n)
            return fail(LFDM_E_STATE, "per-step controls cover " + std::to_string(h->steps_n) + " steps, Solve needs " +
                                          std::to_string(step_index + n));
        std::vector<int> outs;
        if (h->output) outs = output_steps(n, h->step, h->dtout);
        if (outs.size() > dct) outs.resize(dct);
        const size_t row = h->steps_shared ? 3 : 3 * N;
        const double* sp = mode == kCtrlSteps ? steps + (size_t)step_index * row : nullptr;
        if (int rc = launch(h, mode, t_cont, n, &outs, 0, 0, 0, sp)) return rc;
        for (int k = 0; k < n; ++k) t_cont += h->step;
        step_index += n;
        current = kend;
        since += h->ksim;
        if (h->output) {
            const size_t chunk = (size_t)dct * nc * W;
            const size_t filled = chunk * launches;
            if (host.size() < filled + chunk) host.resize(filled + chunk, 0.0);
            const double* src = h->d_out;
            if (!h->store_all) {  // [written vehicle][sample][channel] of this launch
                HIPCHK(launch_gather(h->d_out, (int)N, (int)(dct * nc), h->d_which, (int)W, h->d_outc, h->stream));
                src = h->d_outc;
            }
            HIPCHK(hipMemcpyAsync(host.data() + filled, src, chunk * sizeof(double), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            ++launches;
            if (std::abs(since - h->host_dump) < 1e-6) {
                csv_rows(h, host, hct, dct);
                since = 0.;
                launches = 0;
            }
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->t_now = t_cont;
    if (h->output) {
        if (launches != 0) csv_rows(h, host, launches * dct, dct);
        if (int rc = csv_flush(h)) return rc;
    }
    return LFDM_OK;
}

int lfdm_solve_step(lfdm_solver* h, double t, double steering, double throttle, double braking, double* t_out) {
    if (int rc = check(h)) return rc;
    if (h->output)
        return fail(LFDM_E_STATE,
                    "Cannot get csv file output if SolveStep is called, please access sim_states through GetSimSate");
    if (int rc = ensure_device(h)) return rc;
    if (int rc = sync_inputs(h)) return rc;
    const int n = chunk_steps(t, h->ksim, h->step);
    if (int rc = launch(h, kCtrlConst, t, n, nullptr, steering, throttle, braking)) return rc;
    double tt = t;
    for (int k = 0; k < n; ++k) tt += h->step;
    h->t_now = tt;
    if (t_out) *t_out = t + h->ksim;
    return LFDM_OK;
}

int lfdm_get_sim_state(lfdm_solver* h, unsigned vehicle, double* record) {
    if (int rc = check(h)) return rc;
    if (!record) return fail(LFDM_E_ARG, "null record");
    if (vehicle >= h->total) return fail(LFDM_E_ARG, "Vehicle index out of bounds");
    if (!h->built) {
        if (vehicle >= h->n_states) return fail(LFDM_E_STATE, "vehicle not initialized");
        std::memcpy(record, h->init.data() + (size_t)vehicle * h->rec, h->rec * sizeof(double));
        return LFDM_OK;
    }
    if (int rc = ensure_buffer(h->d_rec, h->rec_cap, (size_t)h->rec + 1)) return rc;
    int slot = h->user_slot[vehicle];
    HIPCHK(hipMemcpyAsync(h->d_rec + h->rec, &slot, sizeof(int), hipMemcpyHostToDevice, h->stream));
    hipError_t e = launch_gather(h->d_state, h->S, h->rec, reinterpret_cast<int*>(h->d_rec + h->rec), 1, h->d_rec,
                                 h->stream);
    if (e != hipSuccess) return fail(LFDM_E_HIP, hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(record, h->d_rec, h->rec * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return LFDM_OK;
}

int lfdm_get_states(lfdm_solver* h, double* records) {
    if (int rc = check(h)) return rc;
    if (!records) return fail(LFDM_E_ARG, "null output");
    if (!h->built) {
        if (h->n_states != h->total) return fail(LFDM_E_STATE, "not every vehicle is initialized");
        std::memcpy(records, h->init.data(), h->init.size() * sizeof(double));
        return LFDM_OK;
    }
    const size_t need = (size_t)h->total * h->rec;
    if (int rc = ensure_buffer(h->d_rec, h->rec_cap, need)) return rc;
    hipError_t e = launch_gather(h->d_state, h->S, h->rec, h->d_user_slot, (int)h->total, h->d_rec, h->stream);
    if (e != hipSuccess) return fail(LFDM_E_HIP, hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(records, h->d_rec, need * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return LFDM_OK;
}

int lfdm_set_states(lfdm_solver* h, const double* records) {
    if (int rc = check(h)) return rc;
    if (!records) return fail(LFDM_E_ARG, "null input");
    if (h->tire_type < 0) return fail(LFDM_E_STATE, "set_states before Construct");
    const size_t need = (size_t)h->total * h->rec;
    h->init.assign(records, records + need);
    h->n_states = h->total;
    if (!h->built) return LFDM_OK;
    if (int rc = ensure_buffer(h->d_rec, h->rec_cap, need)) return rc;
    HIPCHK(hipMemcpyAsync(h->d_rec, records, need * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipError_t e = launch_scatter(h->d_state, h->S, h->rec, h->d_user_slot, (int)h->total, h->d_rec, h->stream);
    if (e != hipSuccess) return fail(LFDM_E_HIP, hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(h->stream));
    return LFDM_OK;
}

int lfdm_set_overrides(lfdm_solver* h, const double* ovr) {
    if (int rc = check(h)) return rc;
    if (!ovr) return fail(LFDM_E_ARG, "null overrides");
    for (size_t i = 0; i < (size_t)kOverrides * h->total; ++i)
        if (!std::isfinite(ovr[i])) return fail(LFDM_E_ARG, "per-vehicle parameters must be finite");
    h->ovr_host.assign(ovr, ovr + (size_t)kOverrides * h->total);
    if (h->built) h->ovr_dirty = true;
    h->ovr_device_newer = false;
    invalidate_adjoint(h);
    return LFDM_OK;
}

int lfdm_get_vehicle_params(lfdm_solver* h, unsigned vehicle, double* out, int cap, int* len) {
    if (int rc = check(h)) return rc;
    if (!out || !len) return fail(LFDM_E_ARG, "null output");
    if (vehicle >= h->n_params) return fail(LFDM_E_ARG, "vehicle " + std::to_string(vehicle) + " not constructed");
    const Group& G = h->groups[h->group_of[vehicle]];
    const int n = (int)G.flat.size();
    *len = n;
    if (cap < n) return fail(LFDM_E_CAPACITY, "parameter vector needs " + std::to_string(n) + " doubles");
    std::copy(G.flat.begin(), G.flat.end(), out);
    double ovr[kOverrides];
    bool have = false;
    if (h->built && h->ovr_device_newer) {
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int f = 0; f < kOverrides; ++f)
            HIPCHK(hipMemcpy(&ovr[f], h->d_ovr + (size_t)f * h->S + h->user_slot[vehicle], sizeof(double),
                             hipMemcpyDeviceToHost));
        have = true;
    } else if (!h->ovr_host.empty()) {
        for (int f = 0; f < kOverrides; ++f) ovr[f] = h->ovr_host[(size_t)f * h->total + vehicle];
        have = true;
    }
    if (have) {  // the vehicle's own values of the per-vehicle parameters
        FlatSlots fs;
        std::string err;
        if (!flat_slots(h->dof, out, n, G.tire_nr == LFDM_TIRE_TMEASYNR, fs, err)) return fail(LFDM_E_PARSE, err);
        out[fs.m] = ovr[0];
        out[fs.jz] = ovr[1];
        if (fs.jx >= 0) out[fs.jx] = ovr[2];
        out[fs.dfy0Pn] = ovr[3];
        out[fs.dfy0P2n] = ovr[4];
        for (int k = 0; k < fs.torque_n; ++k) out[fs.torque + 2 * k + 1] = out[fs.torque + 2 * k + 1] * ovr[5];
        out[fs.max_brake] = ovr[6];
    }
    return LFDM_OK;
}

int lfdm_set_vehicle_params(lfdm_solver* h, unsigned vehicle, const double* flat, int len) {
    if (int rc = check(h)) return rc;
    if (!flat || len <= 0) return fail(LFDM_E_ARG, "empty parameter vector");
    if (vehicle >= h->n_params) return fail(LFDM_E_ARG, "vehicle " + std::to_string(vehicle) + " not constructed");
    PSet ps;
    std::string err;
    if (!pset_from_flat(h->dof, flat, len, h->tire_type, ps, err)) return fail(LFDM_E_ARG, err);
    if (int rc = drop_device_image(h)) return rc;  // the slot layout changes
    const int g0 = h->group_of[vehicle];
    const std::vector<double> fv(flat, flat + len);
    int g1 = -1;
    for (size_t g = 0; g < h->groups.size(); ++g) {
        const Group& G = h->groups[g];
        if (G.flat.size() == fv.size() && std::memcmp(G.flat.data(), fv.data(), fv.size() * sizeof(double)) == 0 &&
            G.rows.size() == h->groups[g0].rows.size() &&
            std::memcmp(G.rows.data(), h->groups[g0].rows.data(), G.rows.size() * sizeof(DrvRow)) == 0) {
            g1 = (int)g;
            break;
        }
    }
    if (g1 < 0) {  // a set of its own, keeping the vehicle's driver table
        Group ng;
        ng.flat = fv;
        ng.tire_nr = h->tire_type;
        ng.pset = ps;
        ng.rows = h->groups[g0].rows;
        h->groups.push_back(std::move(ng));
        g1 = (int)h->groups.size() - 1;
    }
    if (g1 != g0) {
        std::vector<int>& from = h->groups[g0].users;
        from.erase(std::lower_bound(from.begin(), from.end(), (int)vehicle));
        std::vector<int>& to = h->groups[g1].users;
        to.insert(std::lower_bound(to.begin(), to.end(), (int)vehicle), (int)vehicle);
        h->group_of[vehicle] = g1;
    }
    if (!h->ovr_host.empty()) {  // the per-vehicle parameters take the new set's values
        const PSet& p = h->groups[g1].pset;
        const double v[kOverrides] = {p.m, p.jz, p.jx, p.tire.dfy0Pn, p.tire.dfy0P2n, 1.0, p.maxBrakeTorque};
        for (int f = 0; f < kOverrides; ++f) h->ovr_host[(size_t)f * h->total + vehicle] = v[f];
    }
    return LFDM_OK;
}

int lfdm_set_overrides_device(lfdm_solver* h, const double* device_ovr) {
    if (int rc = check(h)) return rc;
    if (!device_ovr) return fail(LFDM_E_ARG, "null overrides");
    if (int rc = ensure_device(h)) return rc;
    if (int rc = sync_inputs(h)) return rc;
    HIPCHK(launch_scatter_fields(h->d_ovr, h->S, kOverrides, h->d_user_slot, (int)h->total, device_ovr, h->stream));
    h->ovr_device_newer = true;
    invalidate_adjoint(h);
    return LFDM_OK;
}

int lfdm_set_states_device(lfdm_solver* h, const double* device_records) {
    if (int rc = check(h)) return rc;
    if (!device_records) return fail(LFDM_E_ARG, "null input");
    if (int rc = ensure_device(h)) return rc;
    HIPCHK(launch_scatter(h->d_state, h->S, h->rec, h->d_user_slot, (int)h->total, device_records, h->stream));
    return LFDM_OK;
}

int lfdm_adjoint_generation(lfdm_solver* h, unsigned long long* gen) {
    if (int rc = check(h)) return rc;
    if (!gen) return fail(LFDM_E_ARG, "null output");
    *gen = h->adj_gen;
    return LFDM_OK;
}

int lfdm_set_controls_knots(lfdm_solver* h, int ns, double ds, int np, double dp, const double* knots) {
    if (int rc = check(h)) return rc;
    if (!knots || ns < 2 || np < 1 || !(ds > 0) || !(dp > 0) || ns + 2 * np > 4096)
        return fail(LFDM_E_ARG, "bad knot program");
    h->ns = ns; h->np = np; h->ds = ds; h->dp = dp;
    h->knots_host.assign(knots, knots + (size_t)(ns + 2 * np) * h->total);
    h->knots_dirty = true;
    h->ctrl_mode = kCtrlKnots;
    return LFDM_OK;
}

int lfdm_set_controls_steps(lfdm_solver* h, int nsteps, const double* ctrl) {
    if (int rc = check(h)) return rc;
    if (!ctrl || nsteps <= 0) return fail(LFDM_E_ARG, "bad per-step controls");
    h->steps_n = nsteps;
    h->steps_pos = 0;
    h->steps_host.assign(ctrl, ctrl + (size_t)nsteps * 3 * h->total);
    h->steps_ext = nullptr;
    h->steps_shared = false;
    h->steps_dirty = true;
    h->ctrl_mode = kCtrlSteps;
    return LFDM_OK;
}

int lfdm_set_controls_steps_device(lfdm_solver* h, int nsteps, const double* dctrl) {
    if (int rc = check(h)) return rc;
    if (!dctrl || nsteps <= 0) return fail(LFDM_E_ARG, "bad per-step controls");
    h->steps_n = nsteps;
    h->steps_pos = 0;
    h->steps_ext = dctrl;
    h->steps_shared = false;
    h->steps_dirty = false;
    h->ctrl_mode = kCtrlSteps;
    return LFDM_OK;
}

int lfdm_set_controls_shared_device(lfdm_solver* h, int nsteps, const double* dctrl) {
    if (int rc = check(h)) return rc;
    if (!dctrl || nsteps <= 0) return fail(LFDM_E_ARG, "bad per-step controls");
    h->steps_n = nsteps;
    h->steps_pos = 0;
    h->steps_ext = dctrl;
    h->steps_shared = true;
    h->steps_dirty = false;
    h->ctrl_mode = kCtrlSteps;
    return LFDM_OK;
}

int lfdm_functor_chunk_begin(lfdm_solver* h, double t, int nsteps, void* record, int record_size, int* block,
                             int* grid, void** stream) {
    if (int rc = check(h)) return rc;
    if (!record || !block || !grid || !stream || nsteps < 0) return fail(LFDM_E_ARG, "bad functor chunk arguments");
    if (record_size != (int)sizeof(Launch))
        return fail(LFDM_E_ARG, "launch record size " + std::to_string(record_size) + " differs from the library's " +
                                    std::to_string(sizeof(Launch)) + " (header and library of different versions)");
    if (h->output)
        return fail(LFDM_E_STATE, "SolveStep with a controls functor has no CSV output; read states with GetSimState");
    if (int rc = ensure_device(h)) return rc;
    if (int rc = sync_inputs(h)) return rc;
    h->t_now = t;
    Launch L = make_launch(h, kCtrlFunctor, t, nsteps);
    std::memcpy(record, &L, sizeof(L));
    *block = block_size(h->tpb);
    *grid = (h->S + *block - 1) / *block;
    *stream = (void*)h->stream;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    return LFDM_OK;
}

int lfdm_functor_chunk_done(lfdm_solver* h, int nsteps) {
    if (int rc = check(h)) return rc;
    for (int k = 0; k < nsteps; ++k) h->t_now += h->step;
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    h->timed = true;
    return LFDM_OK;
}

int lfdm_chunk_steps(lfdm_solver* h, double t, int* nsteps) {
    if (int rc = check(h)) return rc;
    if (!nsteps) return fail(LFDM_E_ARG, "null output");
    *nsteps = chunk_steps(t, h->ksim, h->step);
    return LFDM_OK;
}

int lfdm_initialized_count(lfdm_solver* h, unsigned* count) {
    if (int rc = check(h)) return rc;
    if (!count) return fail(LFDM_E_ARG, "null output");
    *count = h->n_states;
    return LFDM_OK;
}

int lfdm_set_controls_table(lfdm_solver* h) {
    if (int rc = check(h)) return rc;
    h->ctrl_mode = kCtrlTable;
    return LFDM_OK;
}

int lfdm_set_time(lfdm_solver* h, double t) {
    if (int rc = check(h)) return rc;
    h->t_now = t;
    h->steps_pos = 0;
    return LFDM_OK;
}

int lfdm_set_horizons(lfdm_solver* h, const int* steps) {
    if (int rc = check(h)) return rc;
    std::vector<int> hz;
    if (steps) {
        hz.assign(steps, steps + h->total);
        for (int v : hz)
            if (v < 0) return fail(LFDM_E_ARG, "horizons must be non-negative step counts");
    }
    const bool relayout = h->layout == LFDM_LAYOUT_BY_HORIZON && (!hz.empty() || !h->horizon_host.empty());
    if (relayout)
        if (int rc = drop_device_image(h)) return rc;
    h->horizon_host.swap(hz);
    h->horizon_dirty = true;
    return LFDM_OK;
}

int lfdm_set_layout(lfdm_solver* h, int layout) {
    if (int rc = check(h)) return rc;
    if (layout != LFDM_LAYOUT_GROUPED && layout != LFDM_LAYOUT_BY_HORIZON) return fail(LFDM_E_ARG, "bad layout");
    if (layout == h->layout) return LFDM_OK;
    if (!h->horizon_host.empty())
        if (int rc = drop_device_image(h)) return rc;
    h->layout = layout;
    return LFDM_OK;
}

int lfdm_set_arithmetic(lfdm_solver* h, int arith) {
    if (int rc = check(h)) return rc;
    if (arith != LFDM_ARITH_FAST && arith != LFDM_ARITH_STRICT) return fail(LFDM_E_ARG, "unknown arithmetic");
    h->arith = arith;
    return LFDM_OK;
}

int lfdm_get_arithmetic(lfdm_solver* h, int* arith) {
    if (int rc = check(h)) return rc;
    if (!arith) return fail(LFDM_E_ARG, "null output");
    *arith = h->arith;
    return LFDM_OK;
}

int lfdm_get_layout(lfdm_solver* h, int* layout) {
    if (int rc = check(h)) return rc;
    if (!layout) return fail(LFDM_E_ARG, "null output");
    *layout = h->layout;
    return LFDM_OK;
}

int lfdm_advance(lfdm_solver* h, int nsteps) {
    if (int rc = check(h)) return rc;
    if (nsteps <= 0) return fail(LFDM_E_ARG, "step count must be positive");
    if (int rc = ensure_device(h)) return rc;
    if (int rc = sync_inputs(h)) return rc;
    const int mode = h->ctrl_mode;
    if (mode == kCtrlTable)
        for (const Group& g : h->groups)
            if (g.rows.empty()) return fail(LFDM_E_STATE, "no driver table for a group");
    if (mode == kCtrlSteps && h->steps_pos + nsteps > h->steps_n)
        return fail(LFDM_E_STATE, "per-step controls cover " + std::to_string(h->steps_n - h->steps_pos) +
                                      " more steps, " + std::to_string(nsteps) + " requested");
    const double* steps = nullptr;
    if (mode == kCtrlSteps)
        steps = (h->steps_ext ? h->steps_ext : h->d_steps) +
                (size_t)h->steps_pos * (h->steps_shared ? 3 : 3 * (size_t)h->total);
    if (int rc = launch(h, mode, h->t_now, nsteps, nullptr, 0, 0, 0, steps)) return rc;
    if (mode == kCtrlSteps) h->steps_pos += nsteps;
    for (int k = 0; k < nsteps; ++k) h->t_now += h->step;
    return LFDM_OK;
}

namespace {

int adjoint_supported(lfdm_solver* h) {
    if (h->dof != 18) return fail(LFDM_E_UNSUPPORTED, "the adjoint is available for the 18-DOF model");
    if (h->tire_type != LFDM_TIRE_TMEASY) return fail(LFDM_E_UNSUPPORTED, "the adjoint needs TMeasy tires");
    for (const Group& g : h->groups)
        if (g.pset.nonLinearSteer || g.pset.throttleMod)
            return fail(LFDM_E_UNSUPPORTED, "the adjoint needs linear steering and no throttleMod");
    return LFDM_OK;
}

AdjLaunch adj_launch(lfdm_solver* h) {
    AdjLaunch L{};
    L.sets = h->d_sets;
    L.wave_set = h->d_wave_set;
    L.slot_user = h->d_slot_user;
    L.state = h->d_state;
    L.ovr = h->d_ovr;
    L.ctrl = h->adj_ctrl;
    L.ckpt = h->d_ckpt;
    L.tape = h->d_tape;
    L.carry = h->d_carry;
    L.dt = h->step;
    L.nsteps = h->adj_nsteps;
    L.K = h->adj_K;
    L.S = h->S;
    L.n_user = (int)h->total;
    return L;
}


// The tape budget of a forward: the handle's, or, left automatic, half of the device memory free at
// the forward (counting the tape the handle already holds), so that on a 288 GB device the tape of
// tens of thousands of vehicles over thousands of steps fits whole.
size_t adj_budget(lfdm_solver* h) {
    if (h->adj_tape_budget) return h->adj_tape_budget;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return size_t(16) << 30;
    return (free_b + h->tape_cap * sizeof(double)) / 2;
}

// Segments per backward group: as many as the forward's tape budget holds, at least one and at most
// all.  The forward tapes the group the backward walks first, so every segment beyond the first group
// costs one recompute; with the whole horizon in the budget the backward recomputes nothing.
int adj_group(lfdm_solver* h, int nseg) {
    const size_t per_seg = (size_t)h->adj_K * kAdjTapeRow * (size_t)h->S * sizeof(double);
    const size_t G = adj_budget(h) / std::max<size_t>(per_seg, 1);
    return (int)std::max<size_t>(1, std::min<size_t>(G, (size_t)nseg));
}

}  // namespace

int lfdm_adjoint_forward(lfdm_solver* h, int nsteps, int checkpoint_every, const double* dctrl, double* dxy) {
    if (int rc = check(h)) return rc;
    if (nsteps <= 0 || checkpoint_every <= 0 || !dctrl || !dxy) return fail(LFDM_E_ARG, "bad adjoint arguments");
    if (int rc = adjoint_supported(h)) return rc;
    if (int rc = ensure_device(h)) return rc;
    if (int rc = sync_inputs(h)) return rc;
    const int K = std::min(checkpoint_every, nsteps);
    const size_t nseg = (size_t)((nsteps + K - 1) / K);
    if (int rc = ensure_buffer(h->d_ckpt, h->ckpt_cap, nseg * kAdjCkptRow * (size_t)h->S)) return rc;
    const int simds = simd_count(h);
    if (simds < 0) return simds;
    h->adj_nsteps = nsteps;
    h->adj_K = K;
    h->adj_ctrl = dctrl;
    h->tape_holds_last = false;
    // the forward also tapes the group the backward walks first
    const int G = adj_group(h, (int)nseg);
    h->adj_G = G;
    if (int rc = ensure_buffer(h->d_tape, h->tape_cap, (size_t)G * K * kAdjTapeRow * (size_t)h->S)) return rc;
    AdjLaunch L = adj_launch(h);
    L.xy = dxy;
    L.fwd_tape_lo = (int)nseg - G;
    // a fleet of at most one wavefront per SIMD gets the whole register file per lane
    const int waves = (size_t)h->S / kWave <= (size_t)simds ? 1 : 2;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    hipError_t e = launch_adj_forward(L, waves, h->stream);
    if (e != hipSuccess) return fail(LFDM_E_HIP, std::string("adjoint forward launch failed: ") + hipGetErrorString(e));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    h->timed = true;
    h->tape_holds_last = true;
    ++h->adj_gen;
    return LFDM_OK;
}

int lfdm_adjoint_set_tape_budget(lfdm_solver* h, unsigned long long bytes) {
    if (int rc = check(h)) return rc;
    h->adj_tape_budget = (size_t)bytes;
    return LFDM_OK;
}

int lfdm_adjoint_tape_bytes(lfdm_solver* h, unsigned long long* bytes, int* segments_per_group) {
    if (int rc = check(h)) return rc;
    if (!bytes) return fail(LFDM_E_ARG, "null output");
    if (h->adj_nsteps <= 0) return fail(LFDM_E_STATE, "no adjoint forward has run on this handle");
    *bytes = (unsigned long long)h->adj_G * (unsigned long long)h->adj_K * kAdjTapeRow * (unsigned long long)h->S *
             sizeof(double);
    if (segments_per_group) *segments_per_group = h->adj_G;
    return LFDM_OK;
}

int lfdm_step_jacobian(lfdm_solver* h, const double* dctrl, double* djac, int* rows, int* cols) {
    if (int rc = check(h)) return rc;
    if (!dctrl || !djac) return fail(LFDM_E_ARG, "null controls or Jacobian buffer");
    if (int rc = adjoint_supported(h)) return rc;
    if (int rc = ensure_device(h)) return rc;
    if (int rc = sync_inputs(h)) return rc;
    AdjLaunch L = adj_launch(h);
    L.ctrl = dctrl;
    L.jac = djac;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    HIPCHK(launch_step_jacobian(L, h->stream));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    if (rows) *rows = kJacRows;
    if (cols) *cols = kJacCols;
    return LFDM_OK;
}

int lfdm_adjoint_backward(lfdm_solver* h, const double* dxy_bar, double* dctrl_bar, double* dstate_bar,
                          double* dparam_bar) {
    if (int rc = check(h)) return rc;
    if (!dxy_bar || !dctrl_bar || !dstate_bar || !dparam_bar) return fail(LFDM_E_ARG, "null adjoint buffer");
    if (h->adj_nsteps <= 0)
        return fail(LFDM_E_STATE, "no adjoint forward pending: none ran, or the per-vehicle parameters or the "
                                  "slot layout changed since");
    const int nseg = (h->adj_nsteps + h->adj_K - 1) / h->adj_K;
    const int G = h->adj_G;
    if (int rc = ensure_buffer(h->d_carry, h->carry_cap, (size_t)kAdjCarry * (size_t)h->S)) return rc;
    AdjLaunch L = adj_launch(h);
    L.xy_bar = dxy_bar;
    L.ctrl_bar = dctrl_bar;
    L.state_bar = dstate_bar;
    L.param_bar = dparam_bar;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int hi = nseg; hi > 0; hi -= G) {
        L.seg_hi = hi;
        L.seg_lo = std::max(hi - G, 0);
        hipError_t e = hipSuccess;
        if (!(hi == nseg && h->tape_holds_last)) e = launch_adj_recompute(L, h->stream);
        if (e == hipSuccess) e = launch_adj_reverse(L, h->stream);
        if (e != hipSuccess)
            return fail(LFDM_E_HIP, std::string("adjoint backward launch failed: ") + hipGetErrorString(e));
    }
    if (nseg > G) h->tape_holds_last = false;  // later groups overwrote it
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    h->timed = true;
    return LFDM_OK;
}

int lfdm_set_stream(lfdm_solver* h, void* stream) {
    if (int rc = check(h)) return rc;
    if (h->own_stream && h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
        h->own_stream = false;
    }
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
    h->stream_set = true;
    return LFDM_OK;
}

int lfdm_synchronize(lfdm_solver* h) {
    if (int rc = check(h)) return rc;
    if (!h->stream_set && !h->own_stream) return LFDM_OK;  // nothing launched yet
    HIPCHK(hipStreamSynchronize(h->stream));
    return LFDM_OK;
}

int lfdm_last_launch_ms(lfdm_solver* h, double* ms) {
    if (int rc = check(h)) return rc;
    if (!ms) return fail(LFDM_E_ARG, "null output");
    if (!h->timed) return fail(LFDM_E_STATE, "no launch recorded");
    HIPCHK(hipEventSynchronize(h->ev1));
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, h->ev0, h->ev1));
    *ms = f;
    return LFDM_OK;
}

// ---- mixed fleets ------------------------------------------------------------------------------

namespace {

int fleet_check(lfdm_fleet* f) {
    if (!f) return fail(LFDM_E_ARG, "null fleet handle");
    return LFDM_OK;
}

// Occupancies the model's integration kernels are built for (24-DOF: one wave per SIMD only).
int max_waves(int dof) { return dof == 24 ? kMaxWaves24 : (dof == 18 ? kMaxWaves18 : kMaxWaves11); }

void fleet_drop_graph(lfdm_fleet* f) {
    if (f->exec) (void)hipGraphExecDestroy(f->exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    f->exec = nullptr;
    f->graph = nullptr;
    f->sig.clear();
}

int fleet_prepare(lfdm_fleet* f, int nsteps) {
    if (f->buckets.empty()) return fail(LFDM_E_STATE, "the fleet has no buckets");
    if (nsteps <= 0) return fail(LFDM_E_ARG, "step count must be positive");
    for (lfdm_solver* h : f->buckets) {
        if (int rc = ensure_device(h)) return rc;
        if (int rc = sync_inputs(h)) return rc;
        if (int rc = check_controls(h, nsteps)) return rc;
    }
    if (!f->stream_set && !f->own_stream) {
        HIPCHK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        f->own_stream = true;
    }
    if (!f->fork) {
        HIPCHK(hipEventCreateWithFlags(&f->fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&f->post, hipEventDisableTiming));
        HIPCHK(hipEventCreate(&f->ev0));
        HIPCHK(hipEventCreate(&f->ev1));
    }
    while (f->branch.size() < f->buckets.size()) {
        hipStream_t st;
        hipEvent_t a, b;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&a, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&b, hipEventDisableTiming));
        f->branch.push_back(st);
        f->pre.push_back(a);
        f->join.push_back(b);
    }
    return LFDM_OK;
}

// The fleet's stream waits for work pending on the buckets' own streams (and, after fleet_end,
// the buckets' streams wait for the fleet's).
int fleet_begin(lfdm_fleet* f) {
    for (size_t b = 0; b < f->buckets.size(); ++b) {
        HIPCHK(hipEventRecord(f->pre[b], f->buckets[b]->stream));
        HIPCHK(hipStreamWaitEvent(f->stream, f->pre[b], 0));
    }
    return LFDM_OK;
}

int fleet_end(lfdm_fleet* f) {
    HIPCHK(hipEventRecord(f->post, f->stream));
    for (lfdm_solver* h : f->buckets) HIPCHK(hipStreamWaitEvent(h->stream, f->post, 0));
    return LFDM_OK;
}

// Whether the fleet's graph chains its buckets one after another rather than running them side by
// side: when every bucket has at least as many wavefronts as its kernels' occupancy lets the chip hold
// at once, each fills the chip by itself, and running the buckets side by side only mixes kernels of
// different register and LDS budgets on the CUs (the 2,000,000-vehicle mixed 11/18/24-DOF fleet: 1.47
// s side by side, 1.42 s one after another).  Smaller buckets run side by side, filling the chip
// together.
bool fleet_serial(lfdm_fleet* f, const std::vector<std::vector<Launch>>& ls) {
    if (f->buckets.size() < 2) return false;
    for (size_t b = 0; b < f->buckets.size(); ++b) {
        lfdm_solver* h = f->buckets[b];
        const long long simds = simd_count(h);
        const int w = (!ls[b].empty() && ls[b][0].waves > 0) ? ls[b][0].waves : max_waves(h->dof);
        if (simds <= 0 || (long long)(h->S / kWave) < simds * w) return false;
    }
    return true;
}

int launch_list(lfdm_solver* h, const std::vector<Launch>& ls, int block, hipStream_t st) {
    for (const Launch& L : ls) {
        hipError_t e = launch_dof(h->dof, L, block, st);
        if (e != hipSuccess) return fail(LFDM_E_HIP, std::string("integration launch failed: ") + hipGetErrorString(e));
    }
    return LFDM_OK;
}

}  // namespace

int lfdm_fleet_create(lfdm_fleet** out) {
    if (!out) return fail(LFDM_E_ARG, "null output handle");
    *out = new lfdm_fleet();
    return LFDM_OK;
}

int lfdm_fleet_destroy(lfdm_fleet* f) {
    if (!f) return LFDM_OK;
    if (f->stream_set || f->own_stream) (void)hipStreamSynchronize(f->stream);
    fleet_drop_graph(f);
    for (hipStream_t st : f->branch) (void)hipStreamDestroy(st);
    for (hipEvent_t e : f->pre) (void)hipEventDestroy(e);
    for (hipEvent_t e : f->join) (void)hipEventDestroy(e);
    for (hipEvent_t e : {f->fork, f->post, f->ev0, f->ev1})
        if (e) (void)hipEventDestroy(e);
    if (f->own_stream && f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
    return LFDM_OK;
}

int lfdm_fleet_add(lfdm_fleet* f, lfdm_solver* h) {
    if (int rc = fleet_check(f)) return rc;
    if (int rc = check(h)) return rc;
    for (lfdm_solver* b : f->buckets)
        if (b == h) return fail(LFDM_E_ARG, "solver already in the fleet");
    f->buckets.push_back(h);
    f->cfg.emplace_back();
    fleet_drop_graph(f);
    return (int)f->buckets.size() - 1;
}

int lfdm_fleet_set_stream(lfdm_fleet* f, void* stream) {
    if (int rc = fleet_check(f)) return rc;
    if (f->own_stream && f->stream) {
        (void)hipStreamSynchronize(f->stream);
        (void)hipStreamDestroy(f->stream);
    }
    f->stream = (hipStream_t)stream;
    f->own_stream = false;
    f->stream_set = true;
    return LFDM_OK;
}

int lfdm_fleet_set_config(lfdm_fleet* f, int bucket, int block, int steps_per_launch, int waves_per_simd) {
    if (int rc = fleet_check(f)) return rc;
    if (bucket < 0 || bucket >= (int)f->buckets.size()) return fail(LFDM_E_ARG, "bad bucket index");
    if (block < 64 || block > 256 || block % kWave) return fail(LFDM_E_ARG, "block size must be 64, 128, 192 or 256");
    if (steps_per_launch < 0) return fail(LFDM_E_ARG, "steps per launch must be >= 0");
    if (waves_per_simd < 0 || waves_per_simd > max_waves(f->buckets[bucket]->dof))
        return fail(LFDM_E_ARG, "waves per SIMD must be 0 (automatic) or 1 to " +
                                    std::to_string(max_waves(f->buckets[bucket]->dof)) + " for the " +
                                    std::to_string(f->buckets[bucket]->dof) + "-DOF kernel");
    f->cfg[bucket].block = block;
    f->cfg[bucket].spl = steps_per_launch;
    f->cfg[bucket].waves = waves_per_simd;
    return LFDM_OK;
}

int lfdm_fleet_get_config(lfdm_fleet* f, int bucket, int* block, int* steps_per_launch, int* waves_per_simd,
                          int* layout) {
    if (int rc = fleet_check(f)) return rc;
    if (bucket < 0 || bucket >= (int)f->buckets.size()) return fail(LFDM_E_ARG, "bad bucket index");
    if (block) *block = f->cfg[bucket].block;
    if (steps_per_launch) *steps_per_launch = f->cfg[bucket].spl;
    if (waves_per_simd) *waves_per_simd = f->cfg[bucket].waves;
    if (layout) *layout = f->buckets[bucket]->layout;
    return LFDM_OK;
}

int lfdm_fleet_run(lfdm_fleet* f, int nsteps) {
    if (int rc = fleet_check(f)) return rc;
    if (int rc = fleet_prepare(f, nsteps)) return rc;
    std::vector<std::vector<Launch>> ls(f->buckets.size());
    std::vector<double> t_end(f->buckets.size());
    std::vector<char> sig;
    for (size_t b = 0; b < f->buckets.size(); ++b) {
        ls[b] = run_launches(f->buckets[b], nsteps, f->cfg[b].spl, f->cfg[b].waves, &t_end[b]);
        const char* blk = reinterpret_cast<const char*>(&f->cfg[b].block);
        sig.insert(sig.end(), blk, blk + sizeof(int));
        for (const Launch& L : ls[b]) {
            const char* p = reinterpret_cast<const char*>(&L);
            sig.insert(sig.end(), p, p + sizeof(Launch));
        }
    }
    const bool serial = fleet_serial(f, ls);
    sig.push_back(serial ? 1 : 0);
    if (!f->exec || sig != f->sig) {
        fleet_drop_graph(f);
        // one chain of launches per bucket, either one after another on the fleet's stream or side by
        // side on branches forked from it and joined back
        HIPCHK(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
        int rc = LFDM_OK;
        hipError_t e = hipEventRecord(f->fork, f->stream);
        for (size_t b = 0; e == hipSuccess && rc == LFDM_OK && b < f->buckets.size(); ++b) {
            if (serial) {
                rc = launch_list(f->buckets[b], ls[b], f->cfg[b].block, f->stream);
                continue;
            }
            e = hipStreamWaitEvent(f->branch[b], f->fork, 0);
            if (e == hipSuccess) rc = launch_list(f->buckets[b], ls[b], f->cfg[b].block, f->branch[b]);
            if (e == hipSuccess && rc == LFDM_OK) e = hipEventRecord(f->join[b], f->branch[b]);
            if (e == hipSuccess && rc == LFDM_OK) e = hipStreamWaitEvent(f->stream, f->join[b], 0);
        }
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(f->stream, &g);
        if (rc != LFDM_OK || e != hipSuccess || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc != LFDM_OK) return rc;
            return fail(LFDM_E_HIP, std::string("graph capture failed: ") +
                                        hipGetErrorString(e != hipSuccess ? e : ee));
        }
        f->graph = g;
        HIPCHK(hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0));
        f->sig.swap(sig);
    }
    if (int rc = fleet_begin(f)) return rc;
    HIPCHK(hipEventRecord(f->ev0, f->stream));
    HIPCHK(hipGraphLaunch(f->exec, f->stream));
    HIPCHK(hipEventRecord(f->ev1, f->stream));
    f->timed = true;
    if (int rc = fleet_end(f)) return rc;
    for (size_t b = 0; b < f->buckets.size(); ++b) f->buckets[b]->t_now = t_end[b];
    return LFDM_OK;
}

int lfdm_fleet_autotune(lfdm_fleet* f, int nsteps, int trial_steps) {
    if (int rc = fleet_check(f)) return rc;
    if (trial_steps <= 0) return fail(LFDM_E_ARG, "trial step count must be positive");
    if (f->buckets.empty()) return fail(LFDM_E_STATE, "the fleet has no buckets");
    for (size_t b = 0; b < f->buckets.size(); ++b) {
        lfdm_solver* h = f->buckets[b];
        // layout from the lane-steps each would run (results do not depend on it)
        if (!h->horizon_host.empty()) {
            const double g = predicted_lane_steps(h, LFDM_LAYOUT_GROUPED, nsteps);
            const double s = predicted_lane_steps(h, LFDM_LAYOUT_BY_HORIZON, nsteps);
            if (int rc = lfdm_set_layout(h, s < 0.99 * g ? LFDM_LAYOUT_BY_HORIZON : LFDM_LAYOUT_GROUPED)) return rc;
        }
    }
    const int T = std::min(trial_steps, nsteps);
    if (int rc = fleet_prepare(f, nsteps)) return rc;
    if (int rc = fleet_begin(f)) return rc;
    // launch shape from timed trials of T steps from the current states (restored after each)
    for (size_t b = 0; b < f->buckets.size(); ++b) {
        lfdm_solver* h = f->buckets[b];
        const size_t n = (size_t)h->rec * h->S;
        double* saved = nullptr;
        HIPCHK(hipMalloc((void**)&saved, n * sizeof(double)));
        hipError_t e = hipMemcpyAsync(saved, h->d_state, n * sizeof(double), hipMemcpyDeviceToDevice, f->stream);
        float best = 0;
        lfdm_fleet::Cfg pick = f->cfg[b];
        int rc = LFDM_OK;
        const int blocks[3] = {64, 128, 256};
        const int parts[2] = {1, 4};
        for (int trial = 0; trial < max_waves(h->dof) * 6 && rc == LFDM_OK && e == hipSuccess; ++trial) {
            const int wv = 1 + trial / 6, bi = (trial / 2) % 3, pi = trial % 2;
            const std::vector<Launch> ls = run_launches(h, T, (T + parts[pi] - 1) / parts[pi], wv, nullptr);
            float ms = 0;
            for (int rep = 0; rep < 2 && rc == LFDM_OK && e == hipSuccess; ++rep) {
                e = hipMemcpyAsync(h->d_state, saved, n * sizeof(double), hipMemcpyDeviceToDevice, f->stream);
                if (e == hipSuccess) e = hipEventRecord(f->ev0, f->stream);
                if (e == hipSuccess) rc = launch_list(h, ls, blocks[bi], f->stream);
                if (e == hipSuccess && rc == LFDM_OK) e = hipEventRecord(f->ev1, f->stream);
                if (e == hipSuccess && rc == LFDM_OK) e = hipEventSynchronize(f->ev1);
                float x = 0;
                if (e == hipSuccess && rc == LFDM_OK) e = hipEventElapsedTime(&x, f->ev0, f->ev1);
                if (rep == 0 || x < ms) ms = x;
            }
            if (e == hipSuccess && rc == LFDM_OK && (best == 0 || ms < best)) {
                best = ms;
                pick.block = blocks[bi];
                pick.spl = parts[pi] == 1 ? 0 : (nsteps + parts[pi] - 1) / parts[pi];
                pick.waves = wv;
            }
        }
        if (e == hipSuccess) e = hipMemcpyAsync(h->d_state, saved, n * sizeof(double), hipMemcpyDeviceToDevice, f->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
        (void)hipFree(saved);
        if (rc != LFDM_OK) return rc;
        if (e != hipSuccess) return fail(LFDM_E_HIP, std::string("autotune trial failed: ") + hipGetErrorString(e));
        f->cfg[b] = pick;
    }
    f->timed = false;
    return fleet_end(f);
}

int lfdm_fleet_synchronize(lfdm_fleet* f) {
    if (int rc = fleet_check(f)) return rc;
    if (f->stream_set || f->own_stream) HIPCHK(hipStreamSynchronize(f->stream));
    return LFDM_OK;
}

int lfdm_fleet_last_ms(lfdm_fleet* f, double* ms) {
    if (int rc = fleet_check(f)) return rc;
    if (!ms) return fail(LFDM_E_ARG, "null output");
    if (!f->timed) return fail(LFDM_E_STATE, "no fleet run recorded");
    HIPCHK(hipEventSynchronize(f->ev1));
    float x = 0;
    HIPCHK(hipEventElapsedTime(&x, f->ev0, f->ev1));
    *ms = x;
    return LFDM_OK;
}

}  // extern "C"
