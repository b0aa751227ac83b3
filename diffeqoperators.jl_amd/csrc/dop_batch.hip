// dop_batch.hip -- dispatch of a batch's launch groups to the per-storage-type units.
#include "dop_batch.hpp"

namespace dop {

hipError_t launch_batch_group_f64(const BGroup&, const BProb*, const BItem*, int, int, int,
                                  hipStream_t);
hipError_t launch_batch_group_f32(const BGroup&, const BProb*, const BItem*, int, int, int,
                                  hipStream_t);
hipError_t launch_batch_group_bf16(const BGroup&, const BProb*, const BItem*, int, int, int,
                                   hipStream_t);

hipError_t launch_batch_items(int storage, const BProb* probs, const BItem* items,
                              const std::vector<BGroup>& groups, int stage, int accumulate,
                              int grid, hipStream_t s)
{
    for (const BGroup& g : groups) {
        if (g.count < 1) continue;
        hipError_t e;
        switch (storage) {
        case kF64: e = launch_batch_group_f64(g, probs, items, stage, accumulate, grid, s); break;
        case kF32: e = launch_batch_group_f32(g, probs, items, stage, accumulate, grid, s); break;
        case kBF16: e = launch_batch_group_bf16(g, probs, items, stage, accumulate, grid, s); break;
        default: e = hipErrorInvalidValue;
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace dop
