// dop_batch_f32.hip -- batched launches for f32 storage.
#include "dop_batch_impl.hpp"

namespace dop {

hipError_t launch_batch_group_f32(const BGroup& g, const BProb* probs, const BItem* items,
                                  int stage, int accumulate, int grid, hipStream_t s)
{
    return batch::launch_group_any<float>(g, probs, items, stage, accumulate, grid, s);
}

}  // namespace dop
