// dop_batch.hip -- dispatch of a batch's launch groups to the per-storage-type units.
#include "dop_batch.hpp"

namespace dop {

hipError_t launch_batch_group_f64(const BGroup&, const BProb*, const BItem*, int, int, int,
                                  hipStream_t);
hipError_t launch_batch_group_f32(const BGroup&, const BProb*, const BItem*, int, int, int,
                                  hipStream_t);
hipError_t launch_batch_group_bf16(const BGroup&, const BProb*, const BItem*, int, int, int,
                                   hipStream_t);

hipError_t launch_batch_group(int storage, const BGroup& g, const BProb* probs,
                              const BItem* items, int stage, int accumulate, int grid,
                              hipStream_t s)
{
    if (g.count < 1) return hipSuccess;
    switch (storage) {
    case kF64: return launch_batch_group_f64(g, probs, items, stage, accumulate, grid, s);
    case kF32: return launch_batch_group_f32(g, probs, items, stage, accumulate, grid, s);
    case kBF16: return launch_batch_group_bf16(g, probs, items, stage, accumulate, grid, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_items(int storage, const BProb* probs, const BItem* items,
                              const std::vector<BGroup>& groups, int stage, int accumulate,
                              int grid, hipStream_t s)
{
    for (const BGroup& g : groups) {
        const hipError_t e = launch_batch_group(storage, g, probs, items, stage, accumulate, grid, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// The groups of a stage on ns streams forked from s and joined back (events: ns + 1, the
// fork first).  The groups write disjoint problems, so they may run side by side and a small
// group does not wait for a large one; each goes to the least-loaded stream by work items.
// Works eagerly and under stream capture (then the fork and joins become graph edges).
hipError_t launch_batch_items_forked(int storage, const BProb* probs, const BItem* items,
                                     const std::vector<BGroup>& groups, int stage, int accumulate,
                                     int grid, hipStream_t s, const hipStream_t* ws,
                                     const hipEvent_t* evs, int ns)
{
    hipError_t e;
    if ((e = hipEventRecord(evs[0], s)) != hipSuccess) return e;
    for (int i = 0; i < ns; ++i)
        if ((e = hipStreamWaitEvent(ws[i], evs[0], 0)) != hipSuccess) return e;
    std::vector<int64_t> load(ns, 0);
    std::vector<int> order(groups.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return groups[x].count > groups[y].count; });
    for (int gi : order) {
        const int i = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        if ((e = launch_batch_group(storage, groups[gi], probs, items, stage, accumulate, grid,
                                    ws[i])) != hipSuccess)
            return e;
        load[i] += groups[gi].count;
    }
    for (int i = 0; i < ns; ++i) {
        if ((e = hipEventRecord(evs[1 + i], ws[i])) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s, evs[1 + i], 0)) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace dop
