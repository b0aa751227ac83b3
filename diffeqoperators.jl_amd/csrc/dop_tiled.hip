// dop_tiled.hip -- dispatch of the tiled interior kernels by storage type and points per lane
// (kernels: dop_tiled_impl.hpp, instantiated in dop_tiled_{f64,f32,bf16}_v*.hip).
#include <hip/hip_runtime.h>

#include "dop_kernels.hpp"

namespace dop {

hipError_t launch_tiled_f64_v1(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_f64_v2(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_f64_v4(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_f32_v1(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_f32_v2(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_f32_v4(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_f32_v8(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_bf16_v1(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_bf16_v2(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_bf16_v4(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);
hipError_t launch_tiled_bf16_v8(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s);

hipError_t launch_tiled(const DevicePlan& p, void* du, const void* u, int accumulate, int variant,
                        hipStream_t s)
{
    FastArgs A;
    if (!fast_variant(p.fast, variant, &A)) return hipErrorInvalidValue;
    switch (p.host.storage) {
    case kF64:
        switch (A.VX) {
        case 1: return launch_tiled_f64_v1(p, A, du, u, accumulate, s);
        case 2: return launch_tiled_f64_v2(p, A, du, u, accumulate, s);
        case 4: return launch_tiled_f64_v4(p, A, du, u, accumulate, s);
        }
        break;
    case kF32:
        switch (A.VX) {
        case 1: return launch_tiled_f32_v1(p, A, du, u, accumulate, s);
        case 2: return launch_tiled_f32_v2(p, A, du, u, accumulate, s);
        case 4: return launch_tiled_f32_v4(p, A, du, u, accumulate, s);
        case 8: return launch_tiled_f32_v8(p, A, du, u, accumulate, s);
        }
        break;
    case kBF16:
        switch (A.VX) {
        case 1: return launch_tiled_bf16_v1(p, A, du, u, accumulate, s);
        case 2: return launch_tiled_bf16_v2(p, A, du, u, accumulate, s);
        case 4: return launch_tiled_bf16_v4(p, A, du, u, accumulate, s);
        case 8: return launch_tiled_bf16_v8(p, A, du, u, accumulate, s);
        }
        break;
    }
    return hipErrorInvalidValue;
}

}  // namespace dop
