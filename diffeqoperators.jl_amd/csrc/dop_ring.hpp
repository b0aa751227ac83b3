// dop_ring.hpp -- compile-time unrolled steps of a register ring.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace dop {

// Calls f(std::integral_constant<int, J>) for J = J0 .. R-1 until f returns false: the
// steps of a register ring with compile-time slot indices, whatever the body's size.
template <int J, int R, typename F>
__device__ __forceinline__ bool ring_steps(F& f)
{
    if constexpr (J == R) {
        return true;
    } else {
        if (!f(std::integral_constant<int, J>{})) return false;
        return ring_steps<J + 1, R>(f);
    }
}

}  // namespace dop
