// dop_tiled_bf16_v1.hip -- bf16 instantiations of the tiled interior kernels, one point per lane
// (one translation unit per storage type and lane width, so they compile in parallel).
#include "dop_tiled_impl.hpp"

namespace dop {

hipError_t launch_tiled_bf16_v1(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2)
{
    return tiled::launch_interior_k<bf16, 1>(p, A, du, u, accumulate, s, s2);
}

}  // namespace dop
