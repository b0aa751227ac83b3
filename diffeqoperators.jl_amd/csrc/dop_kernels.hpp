// dop_kernels.hpp -- launchers of the gfx950 kernels (defined in dop_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "dop_fast.hpp"
#include "dop_plan.hpp"

namespace dop {

// Everything a launch needs: the host copy of the plan (table pointers are device
// pointers), its device copy, and the tiled-path description.
struct DevicePlan {
    PlanDev host{};
    const PlanDev* dev = nullptr;
    FastPlan fast;
    int variant = 0;   // tile-shape candidate chosen by the autotuner
};

// One work item of a persistent batched launch: a plan bound to its buffers.
struct BatchItemDev {
    const PlanDev* plan;
    void* du;
    const void* u;
    int64_t total;    // output points
    int64_t chunk0;   // first chunk index of this item
};

// One problem of a batched RK4 integration.
struct RkItemDev {
    const PlanDev* plan;
    void* u;          // state (updated in place at the last stage)
    void* s1;         // stage inputs (double-buffered)
    void* s2;
    void* acc;        // running RK4 sum
    double scale;     // du/dt = scale * L Q u
    double dt;
    int64_t total;
    int64_t chunk0;
};

constexpr int kBlock = 256;        // threads per block
constexpr int kChunkPts = 1024;    // points per chunk of a persistent launch

// Single-plan apply: the tiled interior + shell when the plan has one, else generic.
// variant < 0 uses the plan's tuned variant; force_generic = 1 runs the generic kernel.
hipError_t launch_apply(const DevicePlan& p, void* du, const void* u, int accumulate,
                        int variant, int force_generic, hipStream_t s);

// Number of tile-shape candidates of the plan (autotuner).
int apply_variants(const DevicePlan& p);

hipError_t launch_batch(int storage, const BatchItemDev* items, int n, int64_t nchunks,
                        int accumulate, int grid, hipStream_t s);

hipError_t launch_rk4_stage(int storage, const RkItemDev* items, int n, int64_t nchunks,
                            int stage, int grid, hipStream_t s);

}  // namespace dop
