// 18-DOF integration kernels: every tire model x control source instantiation.
#include "integrate.h"

namespace lfdm {

hipError_t launch_d18(const Launch& L, int block, hipStream_t stream) {
    if (L.S <= 0 || L.nsteps <= 0) return hipSuccess;
    return kern::launch_model<kern::M18>(L, block, stream);
}

}  // namespace lfdm
