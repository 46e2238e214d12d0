// Record gather/scatter between the SoA state [field][slot] and packed per-vehicle records.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace lfdm {

namespace {

__global__ void gather_records(const double* __restrict__ state, int S, int rec, const int* __restrict__ slots, int n,
                               double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * rec) return;
    const int v = (int)(i / rec), f = (int)(i % rec);
    out[i] = state[(size_t)f * S + slots[v]];
}

__global__ void scatter_records(double* __restrict__ state, int S, int rec, const int* __restrict__ slots, int n,
                                const double* __restrict__ in) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * rec) return;
    const int v = (int)(i / rec), f = (int)(i % rec);
    state[(size_t)f * S + slots[v]] = in[i];
}

__global__ void scatter_field_rows(double* __restrict__ dst, int S, int F, const int* __restrict__ user_slot, int n,
                                   const double* __restrict__ src) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * F) return;
    const int f = (int)(i / n), v = (int)(i % n);
    dst[(size_t)f * S + user_slot[v]] = src[i];
}

}  // namespace

hipError_t launch_scatter_fields(double* dst, int S, int F, const int* user_slot, int n, const double* src,
                                 hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const long long total = (long long)n * F;
    const int block = 256;
    const int grid = (int)((total + block - 1) / block);
    hipLaunchKernelGGL(scatter_field_rows, grid, block, 0, stream, dst, S, F, user_slot, n, src);
    return hipGetLastError();
}

hipError_t launch_gather(const double* state, int S, int rec, const int* slots, int n, double* out,
                         hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const long long total = (long long)n * rec;
    const int block = 256;
    const int grid = (int)((total + block - 1) / block);
    hipLaunchKernelGGL(gather_records, grid, block, 0, stream, state, S, rec, slots, n, out);
    return hipGetLastError();
}

hipError_t launch_scatter(double* state, int S, int rec, const int* slots, int n, const double* in,
                          hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const long long total = (long long)n * rec;
    const int block = 256;
    const int grid = (int)((total + block - 1) / block);
    hipLaunchKernelGGL(scatter_records, grid, block, 0, stream, state, S, rec, slots, n, in);
    return hipGetLastError();
}

}  // namespace lfdm
