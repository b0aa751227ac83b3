// dop_kernels.hpp -- launchers of the gfx950 kernels (defined in dop_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>

#include "dop_fast.hpp"
#include "dop_plan.hpp"

namespace dop {

// Everything a launch needs: the host copy of the plan (table pointers are device
// pointers), its device copy, and the tiled-path description.
struct DevicePlan {
    PlanDev host{};
    const PlanDev* dev = nullptr;
    FastPlan fast;
    std::atomic<int> variant{0};   // tile-shape candidate chosen by the autotuner
    int generic_only = 0;   // every point on k_apply_generic (no tiles, no single-axis kernel)
};

// Four consecutive values as one vector load / store (element-aligned: addresses need not be
// 16-byte aligned).
template <typename T>
struct Quad;
template <>
struct Quad<float> {
    typedef float V __attribute__((ext_vector_type(4), aligned(4)));
    __device__ static void ld(const float* p, float* o)
    {
        const V v = *reinterpret_cast<const V*>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
    __device__ static void st(float* p, const double* v)
    {
        V o;
        o.x = (float)v[0]; o.y = (float)v[1]; o.z = (float)v[2]; o.w = (float)v[3];
        *reinterpret_cast<V*>(p) = o;
    }
};
template <>
struct Quad<double> {
    typedef double V __attribute__((ext_vector_type(2), aligned(8)));
    __device__ static void ld(const double* p, double* o)
    {
        const V a = *reinterpret_cast<const V*>(p);
        const V b = *reinterpret_cast<const V*>(p + 2);
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    }
    __device__ static void st(double* p, const double* v)
    {
        V a, b;
        a.x = v[0]; a.y = v[1]; b.x = v[2]; b.y = v[3];
        *reinterpret_cast<V*>(p) = a;
        *reinterpret_cast<V*>(p + 2) = b;
    }
};

constexpr int kBlock = 256;        // threads per block
constexpr int kChunkPts = 1024;    // points per chunk of a persistent launch

// Single-plan apply: the tiled interior + shell when the plan has one, else the single-axis
// kernel when one axis carries operators, else generic.  variant < 0 uses the plan's tuned
// variant; force_generic = 1 (or the plan's generic_only) runs the generic kernel.
hipError_t launch_apply(const DevicePlan& p, void* du, const void* u, int accumulate,
                        int variant, int force_generic, hipStream_t s);

// Number of tile-shape candidates of the plan (autotuner).
int apply_variants(const DevicePlan& p);

// *dst = v on stream s (the new values of a plan, carried by the launch itself).
hipError_t launch_set_plan(PlanDev* dst, const PlanDev& v, hipStream_t s);

// The library's own streams for side work on the current device (non-blocking, default
// priority), created on first use and kept for the life of the process: the shell of a 3-D
// apply runs on stream 0, the server's size classes on streams 0, 1, 2.  HIP maps a process's
// streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default) and streams that share a queue
// run in order; with the caller's stream, three streams of the library's own fill four queues,
// where a stream per user (the shell's and each server's) would make two of them share one.
constexpr int kSideStreams = 3;
hipError_t side_stream(int i, hipStream_t* out);

}  // namespace dop
