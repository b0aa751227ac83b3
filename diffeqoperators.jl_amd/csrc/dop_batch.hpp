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
    int32_t tiled;        // 1: 2-D DPP tiles cover the interior box, 2: 1-D DPP tiles
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

constexpr int kMaxBatchTileK = 11;

// kind 0: 2-D tile (a = block column, b = row chunk); kind 1: generic chunk a of the shell;
// kind 2: 1-D tile (a = block column)
struct BItem {
    int32_t prob, kind, a, b;
};

// A run of items of one kind (0: 2-D tile, 1: generic points, 2: 1-D tile), window width K (tiles) and
// accumulator; one launch each.
struct BGroup {
    int32_t kind, K, acc, first, count;
};

// Fills probs[i] geometry from plan i.
void batch_geometry(const DevicePlan& p, int chunk, BProb* B);
// Work items of all problems for row chunk B.chunk, sorted into groups.
void batch_items(const std::vector<BProb>& probs, std::vector<BItem>* items,
                 std::vector<BGroup>* groups);

// stage < 0: plain apply; 0..3: RK4 stage.  One launch per group, in order, on stream s.
// one group's launch (the groups of a stage are independent: they may run concurrently)
hipError_t launch_batch_group(int storage, const BGroup& g, const BProb* probs,
                              const BItem* items, int stage, int accumulate, int grid,
                              hipStream_t s);
hipError_t launch_batch_items_forked(int storage, const BProb* probs, const BItem* items,
                                     const std::vector<BGroup>& groups, int stage, int accumulate,
                                     int grid, hipStream_t s, const hipStream_t* ws,
                                     const hipEvent_t* evs, int ns);
hipError_t launch_batch_items(int storage, const BProb* probs, const BItem* items,
                              const std::vector<BGroup>& groups, int stage, int accumulate,
                              int grid, hipStream_t s);

}  // namespace dop
