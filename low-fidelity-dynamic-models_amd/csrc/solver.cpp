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
    unsigned n = 0;
    unsigned user_base = 0;
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
    double t_now = 0.0;
    // device
    bool built = false;
    int S = 0, rec = 0;
    std::vector<int> user_slot;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    PSet* d_sets = nullptr;
    int* d_wave_set = nullptr;
    int* d_slot_user = nullptr;
    int* d_user_slot = nullptr;
    double* d_state = nullptr;
    double* d_ovr = nullptr;
    DrvRow* d_drv = nullptr;
    double* d_knots = nullptr;
    double* d_steps = nullptr;
    size_t steps_cap = 0;
    int* d_out_steps = nullptr;
    size_t out_steps_cap = 0;
    double* d_out = nullptr;
    size_t out_cap = 0;
    double* d_rec = nullptr;
    size_t rec_cap = 0;
    std::vector<int> outs_host;  // output steps of the launch in flight (source of an async copy)
    // differentiation form
    double* d_ckpt = nullptr;
    size_t ckpt_cap = 0;
    double* d_seg = nullptr;
    size_t seg_cap = 0;
    int adj_nsteps = 0, adj_K = 0;
    const double* adj_ctrl = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    double last_ms = 0.0;
};

namespace {

int check(lfdm_solver* h) {
    if (!h) return fail(LFDM_E_ARG, "null solver handle");
    return LFDM_OK;
}

int ensure_buffer(double*& p, size_t& cap, size_t n) {
    if (n <= cap && p) return LFDM_OK;
    dfree(p);
    HIPCHK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(double)));
    cap = n;
    return LFDM_OK;
}

int ensure_stream(lfdm_solver* h) {
    if (h->stream == nullptr && !h->own_stream) {
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || ndev == 0)
            return fail(LFDM_E_HIP, "no HIP device available: the solver runs only on an AMD GPU");
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
        S += (int)((h->groups[g].n + kWave - 1) / kWave) * kWave;
    }
    h->S = S;
    std::vector<int> slot_user(S, -1), wave_set(S / kWave, 0);
    h->user_slot.assign(h->total, 0);
    std::vector<PSet> sets(h->groups.size());
    std::vector<DrvRow> pool;
    for (size_t g = 0; g < h->groups.size(); ++g) {
        const Group& G = h->groups[g];
        for (unsigned i = 0; i < G.n; ++i) {
            slot_user[slot_base[g] + i] = (int)(G.user_base + i);
            h->user_slot[G.user_base + i] = slot_base[g] + (int)i;
        }
        const int waves = (int)((G.n + kWave - 1) / kWave);
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
            u = (int)h->groups[g].user_base;
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

int launch(lfdm_solver* h, int mode, double t0, int nsteps, const std::vector<int>* outs, double cs = 0,
           double ct = 0, double cb = 0, const double* steps_ptr = nullptr) {
    Launch L{};
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
    int block = (int)((h->tpb + kWave - 1) / kWave) * kWave;
    block = std::min(256, std::max(64, block));
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    hipError_t e = h->dof == 11 ? launch_d11(L, block, h->stream)
                   : h->dof == 18 ? launch_d18(L, block, h->stream)
                                  : launch_d24(L, block, h->stream);
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

void csv_rows(lfdm_solver* h, const std::vector<double>& host, unsigned rows) {
    const unsigned nw = h->store_all ? h->total : h->num_outs;
    const size_t N = h->total;
    const int nc = out_channels(h->dof);
    for (unsigned i = 0; i < nw; ++i) {
        const unsigned v = h->store_all ? i : h->which_outs[i];
        std::ostringstream& o = h->csv[i];
        for (unsigned r = 0; r < rows; ++r) {
            const size_t off = (size_t)r * N * nc;
            for (int c = 0; c < nc; ++c) o << host[off + N * c + v] << ",";
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
    if (h->stream) (void)hipStreamSynchronize(h->stream);
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
    dfree(h->d_rec);
    dfree(h->d_ckpt);
    dfree(h->d_seg);
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
    g.n = n;
    g.user_base = h->n_params;
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
        size_t gi = 0;
        for (unsigned i = first; i < first + n; ++i) {
            while (i >= h->groups[gi].user_base + h->groups[gi].n) ++gi;
            const Group& G = h->groups[gi];
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
    if (h->output) {
        dct = (unsigned)std::ceil(h->ksim / h->dtout);
        hct = (unsigned)std::ceil(h->host_dump / h->dtout);
        if (int rc = ensure_buffer(h->d_out, h->out_cap, (size_t)dct * nc * N)) return rc;
        HIPCHK(hipMemsetAsync(h->d_out, 0, (size_t)dct * nc * N * sizeof(double), h->stream));
        host.assign((size_t)std::max(hct, dct) * nc * N, 0.0);
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
            const size_t filled = N * nc * dct * launches;
            if (host.size() < filled + (size_t)dct * nc * N) host.resize(filled + (size_t)dct * nc * N, 0.0);
            HIPCHK(hipMemcpyAsync(host.data() + filled, h->d_out, (size_t)dct * nc * N * sizeof(double),
                                  hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            ++launches;
            if (std::abs(since - h->host_dump) < 1e-6) {
                csv_rows(h, host, hct);
                since = 0.;
                launches = 0;
            }
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->t_now = t_cont;
    if (h->output) {
        if (launches != 0) csv_rows(h, host, launches * dct);
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
    L.seg = h->d_seg;
    L.dt = h->step;
    L.nsteps = h->adj_nsteps;
    L.K = h->adj_K;
    L.S = h->S;
    L.n_user = (int)h->total;
    return L;
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
    if (int rc = ensure_buffer(h->d_ckpt, h->ckpt_cap, nseg * 32 * (size_t)h->S)) return rc;
    if (int rc = ensure_buffer(h->d_seg, h->seg_cap, (size_t)K * 56 * (size_t)h->S)) return rc;
    h->adj_nsteps = nsteps;
    h->adj_K = K;
    h->adj_ctrl = dctrl;
    AdjLaunch L = adj_launch(h);
    L.xy = dxy;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    hipError_t e = launch_adj_forward(L, h->stream);
    if (e != hipSuccess) return fail(LFDM_E_HIP, std::string("adjoint forward launch failed: ") + hipGetErrorString(e));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    h->timed = true;
    return LFDM_OK;
}

int lfdm_adjoint_backward(lfdm_solver* h, const double* dxy_bar, double* dctrl_bar, double* dstate_bar,
                          double* dparam_bar) {
    if (int rc = check(h)) return rc;
    if (!dxy_bar || !dctrl_bar || !dstate_bar || !dparam_bar) return fail(LFDM_E_ARG, "null adjoint buffer");
    if (h->adj_nsteps <= 0) return fail(LFDM_E_STATE, "adjoint backward before adjoint forward");
    AdjLaunch L = adj_launch(h);
    L.xy_bar = dxy_bar;
    L.ctrl_bar = dctrl_bar;
    L.state_bar = dstate_bar;
    L.param_bar = dparam_bar;
    // fields of the initial-record adjoint that the dynamics never read stay zero
    HIPCHK(hipMemsetAsync(dstate_bar, 0, (size_t)h->rec * h->total * sizeof(double), h->stream));
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    hipError_t e = launch_adj_backward(L, h->stream);
    if (e != hipSuccess) return fail(LFDM_E_HIP, std::string("adjoint backward launch failed: ") + hipGetErrorString(e));
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
    if (!stream) return LFDM_OK;
    h->own_stream = false;
    return LFDM_OK;
}

int lfdm_synchronize(lfdm_solver* h) {
    if (int rc = check(h)) return rc;
    if (!h->stream && !h->own_stream) return LFDM_OK;
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

}  // extern "C"
