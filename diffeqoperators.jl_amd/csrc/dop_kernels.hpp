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
