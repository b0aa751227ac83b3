// dop_tiled_f32_v4.hip -- float instantiations of the tiled interior kernels, four points per lane
// (one translation unit per storage type and lane width, so they compile in parallel).
#include "dop_tiled_impl.hpp"

namespace dop {

hipError_t launch_tiled_f32_v4(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2)
{
    return tiled::launch_interior_k<float, 4>(p, A, du, u, accumulate, s, s2);
}

}  // namespace dop
