// dop_kernels.hip -- gfx950 kernels of the fused operator application.
//
// Bandwidth-bound banded stencils: no MFMA.  Every kernel reads u once per point through
// the cache hierarchy and writes du once; all operators and all axes are summed in
// registers (one pass, where the reference makes one pass plus one temporary per
// operator, composite_operators.jl:64-65).
#include <hip/hip_runtime.h>

#include "dop_eval.hpp"
#include "dop_kernels.hpp"

namespace dop {

// ---------------------------------------------------------------------------------------
// generic path: one thread per output point, any ndim / row program / input kind
template <typename T, int ACC>
__global__ __launch_bounds__(kBlock) void k_apply_generic(const PlanDev* __restrict__ Pp,
                                                         T* __restrict__ du,
                                                         const T* __restrict__ src,
                                                         int accumulate)
{
    const PlanDev& P = *Pp;
    const int64_t total = P.total;
    const int64_t n0 = P.n[0], n1 = P.n[1];
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < total;
         p += (int64_t)gridDim.x * kBlock) {
        int64_t idx[3];
        idx[0] = p % n0;
        const int64_t q = p / n0;
        idx[1] = q % n1;
        idx[2] = q / n1;
        double v = eval_point<T, ACC>(P, src, idx);
        if (accumulate) v += (double)Store<T>::ld(du, p);
        Store<T>::st(du, p, v);
    }
}

template <typename T, int ACC>
static hipError_t launch_generic(const PlanDev& host, const PlanDev* dev, void* du,
                                 const void* u, int accumulate, hipStream_t s)
{
    int64_t blocks = (host.total + kBlock - 1) / kBlock;
    if (blocks > (1 << 22)) blocks = 1 << 22;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_apply_generic<T, ACC>), dim3((unsigned)blocks), dim3(kBlock), 0, s,
                       dev, (T*)du, (const T*)u, accumulate);
    return hipGetLastError();
}

int apply_variants(const DevicePlan& p) { return p.fast.kahan ? 1 : fast_variant_count(p.fast); }

hipError_t launch_tiled(const DevicePlan& p, void* du, const void* u, int accumulate, int variant,
                        hipStream_t s);

static hipError_t launch_generic_any(const PlanDev& host, const PlanDev* dev, void* du,
                                     const void* u, int accumulate, hipStream_t s)
{
    switch (host.storage) {
    case kF64:
        return launch_generic<double, kAccF64>(host, dev, du, u, accumulate, s);
    case kF32:
        if (host.accum == kAccF64) return launch_generic<float, kAccF64>(host, dev, du, u, accumulate, s);
        if (host.accum == kAccF32Kahan)
            return launch_generic<float, kAccF32Kahan>(host, dev, du, u, accumulate, s);
        if (host.accum == kAccF32Wide)
            return launch_generic<float, kAccF32Wide>(host, dev, du, u, accumulate, s);
        return launch_generic<float, kAccF32>(host, dev, du, u, accumulate, s);
    case kBF16:
        if (host.accum == kAccF64) return launch_generic<bf16, kAccF64>(host, dev, du, u, accumulate, s);
        if (host.accum == kAccF32Kahan)
            return launch_generic<bf16, kAccF32Kahan>(host, dev, du, u, accumulate, s);
        if (host.accum == kAccF32Wide)
            return launch_generic<bf16, kAccF32Wide>(host, dev, du, u, accumulate, s);
        return launch_generic<bf16, kAccF32>(host, dev, du, u, accumulate, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_apply(const DevicePlan& p, void* du, const void* u, int accumulate,
                        int variant, int force_generic, hipStream_t s)
{
    // the tiled kernels use 16-byte lane loads/stores: both arrays must be 16-byte aligned
    const bool aligned = ((uintptr_t)du % 16 == 0) && ((uintptr_t)u % 16 == 0);
    if (p.fast.kind >= 2 && !p.fast.kahan && !force_generic && aligned)
        return launch_tiled(p, du, u, accumulate, variant < 0 ? p.variant : variant, s);
    return launch_generic_any(p.host, p.dev, du, u, accumulate, s);
}

}  // namespace dop
