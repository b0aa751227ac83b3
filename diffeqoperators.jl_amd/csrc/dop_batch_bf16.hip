// dop_batch_bf16.hip -- batched launches for bf16 storage.
#include "dop_batch_impl.hpp"

namespace dop {

hipError_t launch_batch_group_bf16(const BGroup& g, const BProb* probs, const BItem* items,
                                  int stage, int accumulate, int grid, hipStream_t s)
{
    return batch::launch_group_any<bf16>(g, probs, items, stage, accumulate, grid, s);
}

}  // namespace dop
