// dop_tiled_f64.hip -- double instantiations of the tiled interior kernels.
#include "dop_tiled_impl.hpp"

namespace dop {

hipError_t launch_tiled_f64(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                           int accumulate, hipStream_t s)
{
    switch (A.VX) {
    case 1: return tiled::launch_interior_k<double, 1>(p, A, du, u, accumulate, s);
    case 2: return tiled::launch_interior_k<double, 2>(p, A, du, u, accumulate, s);
    case 4: return tiled::launch_interior_k<double, 4>(p, A, du, u, accumulate, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace dop
