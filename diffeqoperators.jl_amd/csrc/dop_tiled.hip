// dop_tiled.hip -- dispatch of the tiled interior kernels by storage type
// (kernels: dop_tiled_impl.hpp, instantiated in dop_tiled_{f64,f32,bf16}.hip).
#include <hip/hip_runtime.h>

#include "dop_kernels.hpp"

namespace dop {

hipError_t launch_tiled_f64(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                            int accumulate, hipStream_t s);
hipError_t launch_tiled_f32(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                            int accumulate, hipStream_t s);
hipError_t launch_tiled_bf16(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);

hipError_t launch_tiled(const DevicePlan& p, void* du, const void* u, int accumulate, int variant,
                        hipStream_t s)
{
    FastArgs A;
    if (!fast_variant(p.fast, variant, &A)) return hipErrorInvalidValue;
    switch (p.host.storage) {
    case kF64: return launch_tiled_f64(p, A, du, u, accumulate, s);
    case kF32: return launch_tiled_f32(p, A, du, u, accumulate, s);
    case kBF16: return launch_tiled_bf16(p, A, du, u, accumulate, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace dop
