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

int apply_variants(const DevicePlan& p) { return fast_variant_count(p.fast); }

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
        return launch_generic<float, kAccF32>(host, dev, du, u, accumulate, s);
    case kBF16:
        if (host.accum == kAccF64) return launch_generic<bf16, kAccF64>(host, dev, du, u, accumulate, s);
        if (host.accum == kAccF32Kahan)
            return launch_generic<bf16, kAccF32Kahan>(host, dev, du, u, accumulate, s);
        return launch_generic<bf16, kAccF32>(host, dev, du, u, accumulate, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_apply(const DevicePlan& p, void* du, const void* u, int accumulate,
                        int variant, int force_generic, hipStream_t s)
{
    if (p.fast.kind != 0 && !force_generic)
        return launch_tiled(p, du, u, accumulate, variant < 0 ? p.variant : variant, s);
    return launch_generic_any(p.host, p.dev, du, u, accumulate, s);
}

// ---------------------------------------------------------------------------------------
// persistent batched path: blocks walk a chunk list spanning every item of the batch

// item owning chunk c (items sorted by chunk0); the search is block-uniform
template <typename Item>
__device__ inline int find_item(const Item* __restrict__ items, int n, int64_t c)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].chunk0 <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T>
__device__ inline double eval_any(const PlanDev& P, const T* src, const int64_t* idx)
{
    if (P.accum == kAccF64) return eval_point<T, kAccF64>(P, src, idx);
    if (P.accum == kAccF32Kahan) return eval_point<T, kAccF32Kahan>(P, src, idx);
    return eval_point<T, kAccF32>(P, src, idx);
}

template <>
__device__ inline double eval_any<double>(const PlanDev& P, const double* src, const int64_t* idx)
{
    return eval_point<double, kAccF64>(P, src, idx);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_batch(const BatchItemDev* __restrict__ items, int n,
                                                int64_t nchunks, int accumulate)
{
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int it = find_item(items, n, c);
        const BatchItemDev& B = items[it];
        const PlanDev& P = *B.plan;
        const int64_t n0 = P.n[0], n1 = P.n[1];
        const int64_t first = (c - B.chunk0) * kChunkPts;
        for (int j = threadIdx.x; j < kChunkPts; j += kBlock) {
            const int64_t p = first + j;
            if (p >= B.total) break;
            int64_t idx[3];
            idx[0] = p % n0;
            const int64_t q = p / n0;
            idx[1] = q % n1;
            idx[2] = q / n1;
            double v = eval_any<T>(P, (const T*)B.u, idx);
            if (accumulate) v += (double)Store<T>::ld((const T*)B.du, p);
            Store<T>::st((T*)B.du, p, v);
        }
    }
}

hipError_t launch_batch(int storage, const BatchItemDev* items, int n, int64_t nchunks,
                        int accumulate, int grid, hipStream_t s)
{
    if (nchunks < 1) return hipSuccess;
    const unsigned g = (unsigned)(nchunks < grid ? nchunks : grid);
    switch (storage) {
    case kF64: hipLaunchKernelGGL(k_batch<double>, dim3(g), dim3(kBlock), 0, s, items, n, nchunks, accumulate); break;
    case kF32: hipLaunchKernelGGL(k_batch<float>, dim3(g), dim3(kBlock), 0, s, items, n, nchunks, accumulate); break;
    case kBF16: hipLaunchKernelGGL(k_batch<bf16>, dim3(g), dim3(kBlock), 0, s, items, n, nchunks, accumulate); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// classical RK4, one persistent launch per stage for the whole batch:
//   stage 0: k = F(u);  acc = u + dt/6 k;  s1 = u + dt/2 k
//   stage 1: k = F(s1); acc += dt/3 k;     s2 = u + dt/2 k
//   stage 2: k = F(s2); acc += dt/3 k;     s1 = u + dt k
//   stage 3: k = F(s1); u = acc + dt/6 k
// (no buffer is both read through the stencil and written within one stage)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rk4_stage(const RkItemDev* __restrict__ items,
                                                     int n, int64_t nchunks, int stage)
{
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int it = find_item(items, n, c);
        const RkItemDev& R = items[it];
        const PlanDev& P = *R.plan;
        const int64_t n0 = P.n[0], n1 = P.n[1];
        const int64_t first = (c - R.chunk0) * kChunkPts;
        const T* X = stage == 0 ? (const T*)R.u : stage == 1 ? (const T*)R.s1
                   : stage == 2 ? (const T*)R.s2 : (const T*)R.s1;
        T* u = (T*)R.u;
        T* acc = (T*)R.acc;
        T* nxt = stage == 1 ? (T*)R.s2 : (T*)R.s1;
        const double dt = R.dt;
        for (int j = threadIdx.x; j < kChunkPts; j += kBlock) {
            const int64_t p = first + j;
            if (p >= R.total) break;
            int64_t idx[3];
            idx[0] = p % n0;
            const int64_t q = p / n0;
            idx[1] = q % n1;
            idx[2] = q / n1;
            const double k = R.scale * eval_any<T>(P, X, idx);
            if (stage == 3) {
                Store<T>::st(u, p, (double)Store<T>::ld(acc, p) + (dt / 6.0) * k);
            } else {
                const double up = (double)Store<T>::ld(u, p);
                if (stage == 0) Store<T>::st(acc, p, up + (dt / 6.0) * k);
                else Store<T>::st(acc, p, (double)Store<T>::ld(acc, p) + (dt / 3.0) * k);
                const double h = (stage == 2) ? dt : 0.5 * dt;
                Store<T>::st(nxt, p, up + h * k);
            }
        }
    }
}

hipError_t launch_rk4_stage(int storage, const RkItemDev* items, int n, int64_t nchunks,
                            int stage, int grid, hipStream_t s)
{
    if (nchunks < 1) return hipSuccess;
    const unsigned g = (unsigned)(nchunks < grid ? nchunks : grid);
    switch (storage) {
    case kF64: hipLaunchKernelGGL(k_rk4_stage<double>, dim3(g), dim3(kBlock), 0, s, items, n, nchunks, stage); break;
    case kF32: hipLaunchKernelGGL(k_rk4_stage<float>, dim3(g), dim3(kBlock), 0, s, items, n, nchunks, stage); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dop
