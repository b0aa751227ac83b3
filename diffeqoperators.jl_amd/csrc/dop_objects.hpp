// dop_objects.hpp -- the objects behind the opaque handles of include/dop.h.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dop.h"
#include "dop_batch.hpp"
#include "dop_build.hpp"
#include "dop_kernels.hpp"

namespace dop {
// records the thread-local message returned by dop_last_error(); returns code
int set_error(int code, const std::string& msg);
}  // namespace dop

struct dop_plan {
    dop::DevicePlan d;            // host copy (device table pointers), device copy, tiled path
    dop::PlanDev* dev = nullptr;  // device copy of d.host
    dop::HostPlan H;              // host tables of the current values (dop_plan_update)
    std::atomic<int> tuned{0};    // variant chosen by the autotuner (or the tuning cache)
    std::mutex tune_mu;           // one tuning at a time
    std::vector<void*> bufs;
    int64_t in_elems = 0, out_elems = 0;
    int device = 0;
    ~dop_plan()
    {
        for (void* b : bufs) (void)hipFree(b);
        if (dev) (void)hipFree(dev);
    }
};

// A batch (apply or RK4): problem table + work list on the device, host copies kept for
// re-chunking by the autotuner.
struct dop_batch {
    std::vector<dop::BProb> probs;
    std::vector<dop::BItem> items;
    std::vector<dop::BGroup> groups;
    dop::BProb* d_probs = nullptr;
    dop::BItem* d_items = nullptr;
    int nitems = 0;
    int storage = 0;
    int grid = 0;
    int chunk = 16;
    ~dop_batch()
    {
        if (d_probs) (void)hipFree(d_probs);
        if (d_items) (void)hipFree(d_items);
    }
};

struct dop_rk4 {
    dop_batch b;
    std::vector<void*> bufs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~dop_rk4()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : bufs) (void)hipFree(p);
    }
};

