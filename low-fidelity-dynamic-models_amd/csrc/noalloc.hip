// lfdm_integrate_noalloc: the integration of one parameter set's fleet without any device allocation,
// for callers that cannot allocate (code under HIP graph capture, serving loops that forbid it).
//
// It is the handle path's Initialize (records taken as given) + set_overrides + per-step device controls
// + advance + get_states in one call, and runs the same kernels (the parameter set, the wavefront and slot
// maps, the SoA state image and the per-vehicle parameters, which the handle path keeps in allocations of
// its own, live in one file-scope device array instead).  Every call on a GPU uses that GPU's copy of the
// array: calls that overlap on one GPU race, and each call rewrites every byte it reads before reading it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/lfdm_amd.h"
#include "kernels.h"
#include "params.h"

namespace lfdm {
namespace {

// 64 MiB per GPU, held until the process ends: ~100,000 18-DOF, ~190,000 11-DOF or ~37,000 24-DOF vehicles.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices];  // device address of `scratch` on each device, looked up once
int g_simds[kMaxDevices];

// `scratch` names a host-side stand-in; kernels get the device address from hipGetSymbolAddress.
int scratch_of_current_device(unsigned char** ptr, int* simds) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_last_error(LFDM_E_HIP, "no HIP device available: the solver runs only on an AMD GPU");
    if (dev < 0 || dev >= kMaxDevices) return set_last_error(LFDM_E_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_last_error(LFDM_E_HIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return set_last_error(LFDM_E_HIP, hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char*>(p);
        g_simds[dev] = 4 * std::max(cus, 1);
    }
    *ptr = g_scratch[dev];
    *simds = g_simds[dev];
    return LFDM_OK;
}

// The parameter set travels in kernel arguments (no host buffer stays referenced after the launch).
struct Chunk {
    unsigned char b[1024];
};

__global__ void put_bytes(unsigned char* __restrict__ dst, Chunk c, int len) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) dst[i] = c.b[i];
}

// Slot maps, SoA state and per-vehicle parameters of one wavefront-padded parameter set: slot i holds
// vehicle i; padding slots hold vehicle 0's record and the set's own parameters (as in the handle path).
__global__ void prep_fleet(int n, int S, int rec, const double* __restrict__ in, const double* __restrict__ ovr_in,
                           const PSet* __restrict__ P, int* __restrict__ wave_set, int* __restrict__ slot_user,
                           double* __restrict__ state, double* __restrict__ ovr) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= S) return;
    if ((slot & (kWave - 1)) == 0) wave_set[slot / kWave] = 0;
    const bool real = slot < n;
    slot_user[slot] = real ? slot : -1;
    const int src = real ? slot : 0;
    for (int f = 0; f < rec; ++f) state[(size_t)f * S + slot] = in[(size_t)src * rec + f];
    const double dflt[kOverrides] = {P->m, P->jz, P->jx, P->tire.dfy0Pn, P->tire.dfy0P2n, 1.0, P->maxBrakeTorque};
    for (int f = 0; f < kOverrides; ++f)
        ovr[(size_t)f * S + slot] = (real && ovr_in) ? ovr_in[(size_t)f * n + slot] : dflt[f];
}

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

struct Layout {
    size_t pset, wave, slot, state, ovr, total;
};

Layout layout(long long S, int rec) {
    Layout L;
    L.pset = 0;
    L.wave = up256(sizeof(PSet));
    L.slot = L.wave + up256((size_t)(S / kWave) * sizeof(int));
    L.state = L.slot + up256((size_t)S * sizeof(int));
    L.ovr = L.state + up256((size_t)rec * S * sizeof(double));
    L.total = L.ovr + (size_t)kOverrides * S * sizeof(double);
    return L;
}

#define NACHK(call)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return set_last_error(LFDM_E_HIP, (std::string(#call) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

}  // namespace
}  // namespace lfdm

using namespace lfdm;

extern "C" int lfdm_integrate_noalloc(int dof, const double* flat, int len, int tire_type, unsigned n,
                                      const double* device_records_in, const double* device_ovr,
                                      const double* device_ctrl, int nsteps, double dt, int arith,
                                      double* device_records_out, void* stream) {
    if (dof != 11 && dof != 18 && dof != 24) return set_last_error(LFDM_E_ARG, "model must be 11, 18 or 24");
    if (tire_type != LFDM_TIRE_TMEASY && tire_type != LFDM_TIRE_TMEASYNR)
        return set_last_error(LFDM_E_ARG, "unknown tire type");
    if (arith != LFDM_ARITH_FAST && arith != LFDM_ARITH_STRICT) return set_last_error(LFDM_E_ARG, "unknown arithmetic");
    if (n == 0 || nsteps < 0 || !(dt > 0.0)) return set_last_error(LFDM_E_ARG, "need n > 0, nsteps >= 0 and dt > 0");
    if (!flat || !device_records_in || !device_records_out || (nsteps > 0 && !device_ctrl))
        return set_last_error(LFDM_E_ARG, "null parameter vector, record or control pointer");
    const int rec = lfdm_record_len(dof, tire_type);
    const long long S = ((long long)n + kWave - 1) / kWave * kWave;
    const Layout lay = layout(S, rec);
    if (lay.total > kScratchBytes) {
        const long long per = (long long)(rec + kOverrides) * 8 + 4 + 1;
        const std::string msg = "lfdm_integrate_noalloc holds at most " + std::to_string(kScratchBytes >> 20) +
                                " MiB: about " + std::to_string((long long)(kScratchBytes - 8192) / per) + " " +
                                std::to_string(dof) + "-DOF vehicles, " + std::to_string(n) + " requested";
        return set_last_error(LFDM_E_CAPACITY, msg.c_str());
    }
    PSet P;
    std::string err;
    if (!pset_from_flat(dof, flat, len, tire_type == LFDM_TIRE_TMEASYNR, P, err))
        return set_last_error(LFDM_E_PARSE, err.c_str());
    P.drv_off = 0;
    P.drv_rows = 0;
    unsigned char* base = nullptr;
    int simds = 0;
    if (int rc = scratch_of_current_device(&base, &simds)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the parameter set, in kernel-argument chunks
    const unsigned char* src = reinterpret_cast<const unsigned char*>(&P);
    for (size_t o = 0; o < sizeof(PSet); o += sizeof(Chunk)) {
        Chunk c;
        const int m = (int)std::min(sizeof(Chunk), sizeof(PSet) - o);
        std::memset(c.b, 0, sizeof(c.b));
        std::memcpy(c.b, src + o, (size_t)m);
        hipLaunchKernelGGL(put_bytes, dim3((m + 255) / 256), dim3(256), 0, st, base + lay.pset + o, c, m);
        NACHK(hipGetLastError());
    }
    const PSet* dP = reinterpret_cast<const PSet*>(base + lay.pset);
    int* wave_set = reinterpret_cast<int*>(base + lay.wave);
    int* slot_user = reinterpret_cast<int*>(base + lay.slot);
    double* state = reinterpret_cast<double*>(base + lay.state);
    double* ovr = reinterpret_cast<double*>(base + lay.ovr);
    hipLaunchKernelGGL(prep_fleet, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st, (int)n, (int)S, rec,
                       device_records_in, device_ovr, dP, wave_set, slot_user, state, ovr);
    NACHK(hipGetLastError());
    if (nsteps > 0) {
        Launch L;
        std::memset(&L, 0, sizeof(L));
        L.sets = dP;
        L.wave_set = wave_set;
        L.slot_user = slot_user;
        L.state = state;
        L.ovr = ovr;
        L.steps = device_ctrl;
        L.t0 = 0.0;
        L.dt = dt;
        L.nsteps = nsteps;
        L.S = (int)S;
        L.n_user = (int)n;
        L.mode = kCtrlSteps;
        L.nr = tire_type == LFDM_TIRE_TMEASYNR;
        {  // the handle path's rule: the wavefronts per SIMD the fleet fills, below the model's default
            const long long per = simds > 0 ? ((long long)(S / kWave) + simds - 1) / simds : 0;
            const int wmax = dof == 11 ? kMaxWaves11 : dof == 18 ? kMaxWaves18 : kMaxWaves24;
            L.waves = (per > 0 && per < wmax) ? (int)per : 0;
        }
        L.strict = arith == LFDM_ARITH_STRICT;
        const hipError_t e = dof == 11 ? launch_d11(L, kWave, st) : dof == 18 ? launch_d18(L, kWave, st)
                                                                            : launch_d24(L, kWave, st);
        NACHK(e);
    }
    // slot_user[0 .. n) is the identity: vehicle i's record comes from slot i
    NACHK(launch_gather(state, (int)S, rec, slot_user, (int)n, device_records_out, st));
    return LFDM_OK;
}
