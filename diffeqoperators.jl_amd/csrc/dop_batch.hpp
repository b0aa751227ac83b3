// dop_batch.hpp -- work list of a persistent batched launch (see dop_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "dop_fast.hpp"
#include "dop_kernels.hpp"
#include "dop_plan.hpp"

namespace dop {

// One problem of a batch: its plan, buffers and (2-D) tile geometry.
struct BProb {
    const PlanDev* plan;
    int32_t plan_accum;
    int32_t tiled;        // 2-D DPP tiles cover the interior box
    const void* src;      // apply: u
    void* out;            // apply: du
    void* u;              // RK4: state and stage buffers
    void* s1;
    void* s2;
    void* acc;
    double scale, dt;     // RK4: du/dt = scale * L Q u, step dt
    int32_t K, chunk;
    int32_t lo[2], hi[2]; // interior box
    int32_t nseg, gx;     // wave segments (64-2H outputs each) per row; blocks across x
    int32_t uniform[2];
    int64_t n0, n1, ys;   // extents and row stride
    const double* w64[2];
    const float* w32[2];
    ShellArgs shell;      // points for the generic evaluator (the whole domain if !tiled)
};

// kind 0: tile (a = block column, b = row chunk); kind 1: generic chunk a of the shell
struct BItem {
    int32_t prob, kind, a, b;
};

// Fills probs[i] geometry from plan i and appends its work items for row chunk `chunk`.
void batch_geometry(const DevicePlan& p, int chunk, BProb* B);
void batch_items(const std::vector<BProb>& probs, std::vector<BItem>* items);

// stage < 0: plain apply; 0..3: RK4 stage
hipError_t launch_batch_items(int storage, const BProb* probs, const BItem* items, int nitems,
                              int stage, int accumulate, int grid, hipStream_t s);

}  // namespace dop
