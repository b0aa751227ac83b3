// dop_kernels.hip -- gfx950 kernels of the fused operator application.
//
// Bandwidth-bound banded stencils: no MFMA.  Every kernel reads u once per point through
// the cache hierarchy and writes du once; all operators and all axes are summed in
// registers (one pass, where the reference makes one pass plus one temporary per
// operator, composite_operators.jl:64-65).
#include <hip/hip_runtime.h>

#include <type_traits>

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

// one active axis (1-D plans, the padded plans of the vector calculus and nonlinear_diffusion!):
// eval_point's arithmetic on that axis -- a regular row's K taps in order, a boundary row's
// FP64 value -- with the window width a template constant, 32-bit point indices and no loop
// over axes (bitwise equal to k_apply_generic)
constexpr int64_t kLineBlocks = 8192;

template <typename T, int ACC, int K>
__global__ __launch_bounds__(kBlock) void k_apply_line(const PlanDev* __restrict__ Pp,
                                                      T* __restrict__ du,
                                                      const T* __restrict__ src, int accumulate,
                                                      int a)
{
    using W = typename Acc<ACC>::W;
    const PlanDev& P = *Pp;
    const AxisDev& X = P.ax[a];
    const uint32_t total = (uint32_t)P.total;
    const uint32_t n0 = (uint32_t)P.n[0], n1 = (uint32_t)P.n[1];
    const int64_t S = P.ssrc[a];
    const W* wt = Acc<ACC>::table(X);
    const int64_t ws = X.uniform ? 1 : X.n;
    const int jd = X.dmin + (P.ghost ? 0 : 1);
    // a persistent grid: block bx walks its own contiguous span of points (long-lived waves; a
    // wave per point batch would spend its life being launched), and blocks are dispatched to
    // the 8 XCDs in turn, so XCD x takes the x-th contiguous eighth of the spans: neighbouring
    // rows and planes (the taps of strided axes) meet in the same L2
    const uint32_t nb = gridDim.x, per = nb >> 3;
    const uint32_t bx = blockIdx.x < per * 8 ? (blockIdx.x & 7) * per + (blockIdx.x >> 3) : blockIdx.x;
    const uint32_t span = ((total + nb - 1) / nb + kBlock - 1) / kBlock * kBlock;
    const uint32_t pend = min(total, (uint32_t)min((uint64_t)(bx + 1) * span, (uint64_t)total));
    // the plan's fields the loop needs, read once (the point loop then issues no scalar loads)
    const int64_t tsh = P.tshift[a];
    int64_t sst[kMaxDim];
#pragma unroll
    for (int b = 0; b < kMaxDim; ++b) sst[b] = (b != a && b < P.ndim) ? P.ssrc[b] : 0;
    const int rlo = X.rlo, rhi = X.rhi;
    for (uint32_t p = bx * span + threadIdx.x; p < pend; p += kBlock) {
        const uint32_t q = p / n0;
        const uint32_t q2 = q / n1;
        int64_t idx[3];
        idx[0] = p - q * n0;
        idx[1] = q - q2 * n1;
        idx[2] = q2;
        const int64_t base = tsh + idx[0] * sst[0] + idx[1] * sst[1] + idx[2] * sst[2];
        const int r = (int)idx[a];
        Acc<ACC> acc;
        if (r >= rlo && r < rhi) {
            const W* w = wt + (X.uniform ? 0 : r);
            const int64_t p0 = base + (int64_t)(r + jd) * S;
            W wv[K], xv[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                wv[k] = w[k * ws];
                xv[k] = (W)Store<T>::ld(src, p0 + k * S);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) acc.add(wv[k], xv[k]);
        } else {
            acc.add_sum(boundary_row<T>(P, X, src, r, base, S, line_of(X, idx)));
        }
        double v = acc.value();
        if (accumulate) v += (double)Store<T>::ld(du, p);
        Store<T>::st(du, p, v);
    }
}

// Four consecutive points of one row per thread (rows of n0 % 4 == 0 points): the same
// arithmetic as k_apply_line, with the values of the four points' taps read as 4-value vector
// loads (element-aligned; the padded inputs of the vector calculus have odd strides), uniform
// weights read once per thread, and one vector store -- a quarter of the memory instructions,
// which bound the one-point form.
template <typename T, int ACC, int K>
__global__ __launch_bounds__(kBlock) void k_apply_line4(const PlanDev* __restrict__ Pp,
                                                       T* __restrict__ du,
                                                       const T* __restrict__ src, int accumulate,
                                                       int a)
{
    using W = typename Acc<ACC>::W;
    const PlanDev& P = *Pp;
    const AxisDev& X = P.ax[a];
    const uint32_t g0 = (uint32_t)P.n[0] >> 2, n0 = (uint32_t)P.n[0], n1 = (uint32_t)P.n[1];
    const uint32_t rows = (uint32_t)(P.total / P.n[0]);
    const int64_t S = P.ssrc[a];
    const W* wt = Acc<ACC>::table(X);
    const int64_t ws = X.uniform ? 1 : X.n;
    const int jd = X.dmin + (P.ghost ? 0 : 1);
    const int64_t tsh = P.tshift[a];
    int64_t sst[kMaxDim];
#pragma unroll
    for (int b = 0; b < kMaxDim; ++b) sst[b] = (b != a && b < P.ndim) ? P.ssrc[b] : 0;
    const int rlo = X.rlo, rhi = X.rhi;
    const bool uni = X.uniform != 0;
    // uniform weights: read once
    W wu[K];
#pragma unroll
    for (int k = 0; k < K; ++k) wu[k] = uni ? tab(wt, k) : W(0);
    // groups of four points along axis 0: on the active axis 0 the groups [gA, gB) of a row
    // hold interior rows only; the points of the row's other groups (boundary rows and their
    // neighbours) are evaluated one per thread afterwards, so that no wave serialises four
    // boundary rows per thread
    const uint32_t gA = a == 0 ? (uint32_t)(rlo + 3) / 4 : 0;
    const uint32_t gB = a == 0 ? max(gA, (uint32_t)rhi / 4) : g0;
    const uint32_t gin = gB - gA;
    const uint32_t nq = rows * gin;
    const uint32_t nb = gridDim.x, per = nb >> 3;
    const uint32_t bx = blockIdx.x < per * 8 ? (blockIdx.x & 7) * per + (blockIdx.x >> 3) : blockIdx.x;
    const uint32_t span = ((nq + nb - 1) / nb + kBlock - 1) / kBlock * kBlock;
    const uint32_t gend = min(nq, (uint32_t)min((uint64_t)(bx + 1) * span, (uint64_t)nq));
    for (uint32_t g = bx * span + threadIdx.x; g < gend; g += kBlock) {
        const uint32_t row = g / gin;
        const uint32_t gi = gA + (g - row * gin);
        const uint32_t q2 = row / n1;
        int64_t idx[3];
        idx[0] = (int64_t)gi * 4;
        idx[1] = row - q2 * n1;
        idx[2] = q2;
        const int64_t base = tsh + idx[0] * sst[0] + idx[1] * sst[1] + idx[2] * sst[2];
        Acc<ACC> acc[4];
        if (a != 0) {
            // the four points share row r of the axis
            const int r = (int)idx[a];
            if (r >= rlo && r < rhi) {
                const int64_t p0 = base + (int64_t)(r + jd) * S;
                W wv[K];
                T xv[K][4];
                if (uni) {
#pragma unroll
                    for (int k = 0; k < K; ++k) wv[k] = wu[k];
                } else {
#pragma unroll
                    for (int k = 0; k < K; ++k) wv[k] = tab(wt, k * ws + r);
                }
#pragma unroll
                for (int k = 0; k < K; ++k) Quad<T>::ld(src + p0 + k * S, xv[k]);
#pragma unroll
                for (int k = 0; k < K; ++k) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[i].add(wv[k], (W)xv[k][i]);
                }
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    int64_t id[3] = {idx[0] + i, idx[1], idx[2]};
                    acc[i].add_sum(boundary_row<T>(P, X, src, r, base + i, S, line_of(X, id)));
                }
            }
        } else {
            // rows r0 .. r0 + 3, all interior: the window of K + 3 values read as groups of
            // four, the last group ending at the window's end (no value past it is read)
            const int r0 = (int)idx[0];
            constexpr int NW = K + 3;
            constexpr int NG = (NW + 3) / 4;
            T xv[NW];
#pragma unroll
            for (int m = 0; m < NG; ++m) {
                const int o = 4 * m < NW - 4 ? 4 * m : NW - 4;
                T t[4];
                Quad<T>::ld(src + base + (int64_t)(r0 + jd) + o, t);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (o + e >= 4 * m) xv[o + e] = t[e];
            }
            W wv[K][4];
            if (uni) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) wv[k][i] = wu[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) wv[k][i] = tab(wt, k * ws + r0 + i);
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[i].add(wv[k][i], (W)xv[i + k]);
            }
        }
        const int64_t p = (int64_t)row * n0 + idx[0];
        double v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = acc[i].value();
        if (accumulate) {
            T old[4];
            Quad<T>::ld(du + p, old);
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] += (double)old[i];
        }
        Quad<T>::st(du + p, v);
    }
    if (a != 0) return;
    // the points of each row's edge groups, one per thread (k_apply_line's arithmetic)
    const uint32_t ne = 4 * (g0 - gin);
    const uint32_t nE = rows * ne;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < nE; e += nb * kBlock) {
        const uint32_t row = e / ne;
        const uint32_t j = e - row * ne;
        const uint32_t x = j < 4 * gA ? j : 4 * gB + (j - 4 * gA);
        const uint32_t q2 = row / n1;
        int64_t idx[3];
        idx[0] = x;
        idx[1] = row - q2 * n1;
        idx[2] = q2;
        const int64_t base = tsh + idx[1] * sst[1] + idx[2] * sst[2];
        const int r = (int)x;
        Acc<ACC> acc;
        if (r >= rlo && r < rhi) {
            const int64_t p0 = base + (int64_t)(r + jd);
            W wv[K], xv[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                wv[k] = uni ? wu[k] : tab(wt, k * ws + r);
                xv[k] = (W)Store<T>::ld(src, p0 + k);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) acc.add(wv[k], xv[k]);
        } else {
            acc.add_sum(boundary_row<T>(P, X, src, r, base, 1, line_of(X, idx)));
        }
        const int64_t p = (int64_t)row * n0 + x;
        double v = acc.value();
        if (accumulate) v += (double)Store<T>::ld(du, p);
        Store<T>::st(du, p, v);
    }
}

template <typename T, int ACC>
static hipError_t launch_line(const PlanDev& host, const PlanDev* dev, void* du, const void* u,
                              int accumulate, int a, hipStream_t s)
{
    // enough resident blocks for every CU several times over, each with a long span; four
    // points per thread when rows hold whole groups of four (f32 and f64 storage)
    const bool quad = !std::is_same<T, bf16>::value && host.n[0] % 4 == 0 && host.ssrc[0] == 1;
    int64_t blocks = ((quad ? host.total / 4 : host.total) + kBlock - 1) / kBlock;
    if (blocks > kLineBlocks) blocks = kLineBlocks;
    if (blocks < 1) blocks = 1;
#define DOP_LINE(KK)                                                                         \
    do {                                                                                     \
        if constexpr (!std::is_same<T, bf16>::value) {                                       \
            if (quad) {                                                                      \
                hipLaunchKernelGGL((k_apply_line4<T, ACC, KK>), dim3((unsigned)blocks),       \
                                   dim3(kBlock), 0, s, dev, (T*)du, (const T*)u, accumulate, a); \
                break;                                                                       \
            }                                                                                \
        }                                                                                    \
        hipLaunchKernelGGL((k_apply_line<T, ACC, KK>), dim3((unsigned)blocks), dim3(kBlock), 0, s, \
                           dev, (T*)du, (const T*)u, accumulate, a);                         \
    } while (0)
    switch (host.ax[a].K) {
    case 2: DOP_LINE(2); break;
    case 3: DOP_LINE(3); break;
    case 4: DOP_LINE(4); break;
    case 5: DOP_LINE(5); break;
    case 6: DOP_LINE(6); break;
    case 7: DOP_LINE(7); break;
    case 8: DOP_LINE(8); break;
    case 9: DOP_LINE(9); break;
    default: return hipErrorInvalidValue;
    }
#undef DOP_LINE
    return hipGetLastError();
}

// the single active axis of a plan the line kernel can take, or -1
static int line_axis(const PlanDev& h)
{
    if (h.total >= ((int64_t)1 << 31)) return -1;
    int a = -1;
    for (int b = 0; b < h.ndim; ++b) {
        if (!h.ax[b].active) continue;
        if (a >= 0) return -1;
        a = b;
    }
    if (a < 0 || h.ax[a].K < 2 || h.ax[a].K > 9 || h.ax[a].rhi <= h.ax[a].rlo) return -1;
    return a;
}

template <typename T, int ACC>
static hipError_t launch_generic(const PlanDev& host, const PlanDev* dev, void* du,
                                 const void* u, int accumulate, hipStream_t s, int line = 1)
{
    if (line) {
        const int a = line_axis(host);
        if (a >= 0) return launch_line<T, ACC>(host, dev, du, u, accumulate, a, s);
    }
    int64_t blocks = (host.total + kBlock - 1) / kBlock;
    if (blocks > (1 << 22)) blocks = 1 << 22;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_apply_generic<T, ACC>), dim3((unsigned)blocks), dim3(kBlock), 0, s,
                       dev, (T*)du, (const T*)u, accumulate);
    return hipGetLastError();
}

int apply_variants(const DevicePlan& p) { return p.fast.kahan ? 1 : fast_variant_count(p.fast); }

// copies a plan's values, passed by value, word by word into its device copy
__global__ __launch_bounds__(64) void k_set_plan(PlanDev* __restrict__ dst, const PlanDev v)
{
    static_assert(sizeof(PlanDev) % sizeof(uint64_t) == 0, "whole words");
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&v);
    uint64_t* out = reinterpret_cast<uint64_t*>(dst);
    for (int i = threadIdx.x; i < (int)(sizeof(PlanDev) / sizeof(uint64_t)); i += 64) out[i] = src[i];
}

hipError_t launch_set_plan(PlanDev* dst, const PlanDev& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_plan, dim3(1), dim3(64), 0, s, dst, v);
    return hipGetLastError();
}

hipError_t launch_tiled(const DevicePlan& p, void* du, const void* u, int accumulate, int variant,
                        hipStream_t s);

static hipError_t launch_generic_any(const PlanDev& host, const PlanDev* dev, void* du,
                                     const void* u, int accumulate, hipStream_t s, int line)
{
    switch (host.storage) {
    case kF64:
        return launch_generic<double, kAccF64>(host, dev, du, u, accumulate, s, line);
    case kF32:
        if (host.accum == kAccF64) return launch_generic<float, kAccF64>(host, dev, du, u, accumulate, s, line);
        if (host.accum == kAccF32Kahan)
            return launch_generic<float, kAccF32Kahan>(host, dev, du, u, accumulate, s, line);
        if (host.accum == kAccF32Wide)
            return launch_generic<float, kAccF32Wide>(host, dev, du, u, accumulate, s, line);
        return launch_generic<float, kAccF32>(host, dev, du, u, accumulate, s, line);
    case kBF16:
        if (host.accum == kAccF64) return launch_generic<bf16, kAccF64>(host, dev, du, u, accumulate, s, line);
        if (host.accum == kAccF32Kahan)
            return launch_generic<bf16, kAccF32Kahan>(host, dev, du, u, accumulate, s, line);
        if (host.accum == kAccF32Wide)
            return launch_generic<bf16, kAccF32Wide>(host, dev, du, u, accumulate, s, line);
        return launch_generic<bf16, kAccF32>(host, dev, du, u, accumulate, s, line);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_apply(const DevicePlan& p, void* du, const void* u, int accumulate,
                        int variant, int force_generic, hipStream_t s)
{
    // the tiled kernels use 16-byte lane loads/stores: both arrays must be 16-byte aligned
    const bool aligned = ((uintptr_t)du % 16 == 0) && ((uintptr_t)u % 16 == 0);
    if (p.fast.kind >= 2 && !p.fast.kahan && !force_generic && aligned)
        return launch_tiled(p, du, u, accumulate,
                            variant < 0 ? p.variant.load(std::memory_order_relaxed) : variant, s);
    return launch_generic_any(p.host, p.dev, du, u, accumulate, s,
                              !force_generic && !p.generic_only);
}

}  // namespace dop
