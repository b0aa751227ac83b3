// dop_tiled.hip -- dispatch of the tiled interior kernels by storage type and points per lane
// (kernels: dop_tiled_impl.hpp, instantiated in dop_tiled_{f64,f32,bf16}_v*.hip).
#include <hip/hip_runtime.h>

#include <mutex>

#include "dop_kernels.hpp"

namespace dop {

hipError_t launch_tiled_f64_v1(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_f64_v2(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_f64_v4(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_f32_v1(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_f32_v2(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_f32_v4(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_f32_v8(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_bf16_v1(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_bf16_v2(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_bf16_v4(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);
hipError_t launch_tiled_bf16_v8(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2);

hipError_t side_stream(int i, hipStream_t* out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static hipStream_t pool[kMaxDevices][kSideStreams] = {};
    if (i < 0 || i >= kSideStreams || !out) return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!pool[dev][i] && (e = hipStreamCreateWithFlags(&pool[dev][i], hipStreamNonBlocking)) != hipSuccess) {
        pool[dev][i] = nullptr;
        return e;
    }
    *out = pool[dev][i];
    return hipSuccess;
}

namespace {

// The second stream of a 3-D apply: the shell (latency-bound, few points) runs there beside the
// tiles (bandwidth-bound); they write disjoint points.  The stream is the library's side stream
// 0; the fork and join events are per thread and device (created on first use and kept), so
// threads that apply at once order their own shells correctly on the shared stream.
struct ShellStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

hipError_t shell_stream(ShellStream** out)
{
    constexpr int kMaxDevices = 64;
    thread_local ShellStream per[kMaxDevices];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    ShellStream& S = per[dev];
    if (!S.s) {
        if ((e = hipEventCreateWithFlags(&S.fork, hipEventDisableTiming)) != hipSuccess) return e;
        if ((e = hipEventCreateWithFlags(&S.join, hipEventDisableTiming)) != hipSuccess) return e;
        if ((e = side_stream(0, &S.s)) != hipSuccess) return e;
    }
    *out = &S;
    return hipSuccess;
}

hipError_t launch_tiled_on(const DevicePlan& p, const FastArgs& A, void* du, const void* u, int accumulate,
                           hipStream_t s, hipStream_t s2);

}  // namespace

hipError_t launch_tiled(const DevicePlan& p, void* du, const void* u, int accumulate, int variant,
                        hipStream_t s)
{
    FastArgs A;
    if (!fast_variant(p.fast, variant, &A)) return hipErrorInvalidValue;
    if (p.fast.kind != 3 || p.fast.strips.nblocks == 0) return launch_tiled_on(p, A, du, u, accumulate, s, s);
    // fork: the shell on the second stream after everything before it on s; join before
    // anything after it on s
    ShellStream* S = nullptr;
    hipError_t e = shell_stream(&S);
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(S->fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(S->s, S->fork, 0)) != hipSuccess) return e;
    const hipError_t le = launch_tiled_on(p, A, du, u, accumulate, s, S->s);
    if ((e = hipEventRecord(S->join, S->s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(s, S->join, 0)) != hipSuccess) return e;
    return le;
}

namespace {

hipError_t launch_tiled_on(const DevicePlan& p, const FastArgs& A, void* du, const void* u, int accumulate,
                           hipStream_t s, hipStream_t s2)
{
    switch (p.host.storage) {
    case kF64:
        switch (A.VX) {
        case 1: return launch_tiled_f64_v1(p, A, du, u, accumulate, s, s2);
        case 2: return launch_tiled_f64_v2(p, A, du, u, accumulate, s, s2);
        case 4: return launch_tiled_f64_v4(p, A, du, u, accumulate, s, s2);
        }
        break;
    case kF32:
        switch (A.VX) {
        case 1: return launch_tiled_f32_v1(p, A, du, u, accumulate, s, s2);
        case 2: return launch_tiled_f32_v2(p, A, du, u, accumulate, s, s2);
        case 4: return launch_tiled_f32_v4(p, A, du, u, accumulate, s, s2);
        case 8: return launch_tiled_f32_v8(p, A, du, u, accumulate, s, s2);
        }
        break;
    case kBF16:
        switch (A.VX) {
        case 1: return launch_tiled_bf16_v1(p, A, du, u, accumulate, s, s2);
        case 2: return launch_tiled_bf16_v2(p, A, du, u, accumulate, s, s2);
        case 4: return launch_tiled_bf16_v4(p, A, du, u, accumulate, s, s2);
        case 8: return launch_tiled_bf16_v8(p, A, du, u, accumulate, s, s2);
        }
        break;
    }
    return hipErrorInvalidValue;
}

}  // namespace

}  // namespace dop
