// dop_shell.hpp -- the device kernel of the 3-D shell: one work item (dop_strip.hpp) per
// thread, blocks handed out so that consecutive blocks of the segment table run on one XCD
// (items that read each other's rows meet in its L2; the rows of a face segment take turns
// block by block), each thread's 16 staged values of its line and its parked accumulators in
// LDS, vector loads and stores of an item's values and weights.
#pragma once
#include <hip/hip_runtime.h>

#include "dop_strip.hpp"

namespace dop {
namespace shell {

using strip::kSW;
using strip::kStage;
constexpr int kShellBlock = kStripBlock;

// Weights (plan tables, global memory): w (<= 4) consecutive values, one 16-byte load (f32)
// or two (f64) when `vec` (then the whole aligned group of 4 is loaded).
struct VecW {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef double d2 __attribute__((ext_vector_type(2)));
    __device__ static void ldw(const float* p, int w, bool vec, float* o)
    {
        if (vec) {
            const f4 v = tab(reinterpret_cast<const f4*>(p), 0);
            o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i) o[i] = i < w ? tab(p, i) : 0.0f;
    }
    __device__ static void ldw(const double* p, int w, bool vec, double* o)
    {
        if (vec) {
            const d2 a = tab(reinterpret_cast<const d2*>(p), 0);
            const d2 b = tab(reinterpret_cast<const d2*>(p), 1);
            o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i) o[i] = i < w ? tab(p, i) : 0.0;
    }
};

// w (<= 4) consecutive values at p, as C: when `vec` the aligned group of 4 starting at p in
// one vector load (16 bytes f32, two 16-byte loads f64, 8 bytes bf16), one load per value
// otherwise; st stores a whole group with one vector store when `vec` and w == 4
template <typename T>
struct VecLoad;

template <>
struct VecLoad<float> : VecW {
    __device__ static void ld(const float* p, int w, bool vec, float* o)
    {
        if (vec) {
            const float4 v = *reinterpret_cast<const float4*>(p);
            o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i) o[i] = i < w ? p[i] : 0.0f;
    }
    __device__ static void st(float* p, int w, bool vec, const double* v)
    {
        if (vec && w == 4) {
            *reinterpret_cast<float4*>(p) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i)
            if (i < w) p[i] = (float)v[i];
    }
};

template <>
struct VecLoad<double> : VecW {
    __device__ static void ld(const double* p, int w, bool vec, double* o)
    {
        if (vec) {
            const double2 a = *reinterpret_cast<const double2*>(p);
            const double2 b = *reinterpret_cast<const double2*>(p + 2);
            o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i) o[i] = i < w ? p[i] : 0.0;
    }
    __device__ static void st(double* p, int w, bool vec, const double* v)
    {
        if (vec && w == 4) {
            *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
            *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i)
            if (i < w) p[i] = v[i];
    }
};

template <>
struct VecLoad<bf16> : VecW {
    __device__ static void ld(const bf16* p, int w, bool vec, float* o)
    {
        if (vec) {
            const uint2 v = *reinterpret_cast<const uint2*>(p);
            o[0] = __builtin_bit_cast(float, v.x << 16);
            o[1] = __builtin_bit_cast(float, v.x & 0xffff0000u);
            o[2] = __builtin_bit_cast(float, v.y << 16);
            o[3] = __builtin_bit_cast(float, v.y & 0xffff0000u);
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i) o[i] = i < w ? bf16_to_f32(p[i].bits) : 0.0f;
    }
    __device__ static void st(bf16* p, int w, bool vec, const double* v)
    {
        if (vec && w == 4) {
            uint2 o;
            o.x = (uint32_t)f32_to_bf16((float)v[0]) | ((uint32_t)f32_to_bf16((float)v[1]) << 16);
            o.y = (uint32_t)f32_to_bf16((float)v[2]) | ((uint32_t)f32_to_bf16((float)v[3]) << 16);
            *reinterpret_cast<uint2*>(p) = o;
            return;
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i)
            if (i < w) p[i].bits = f32_to_bf16((float)v[i]);
    }
};

// Block i of the launch takes block strip_block(i) of work items: the hardware hands consecutive block
// indices to the 8 XCDs in turn, so XCD x receives the blocks i = x, x + 8, ... and this
// gives it one contiguous range of work items.
__device__ inline int64_t strip_block(int i, int nb)
{
    const int per = nb >> 3;
    if (i >= per * 8) return i;
    return (int64_t)(i & 7) * per + (i >> 3);
}

// Kind 4: the x-face strip x0 .. x0 + w - 1 of a tile of 16 x 16 lines (y, z), one line per
// thread.  The block first loads, with four lanes per line end, the 16 values at the end of
// each of its own lines (the x term's taps and ghosts) and the strip's 4 values of every line of
// the tile and its kXfHalo-line border (the y and z taps) into LDS; then each thread evaluates
// its points with the generic arithmetic: x term per point (row programs uniform over the
// block), y and z terms as regular rows (the plan's own windows) from LDS.
template <typename T, int ACC, int KM>
__device__ inline void xface_tile(const PlanDev& P, const StripArgs& SA, const ShellSeg& G, int tile,
                                  T (*full)[kStage + 1], T (*grp)[kSW], T* __restrict__ du,
                                  const T* __restrict__ src, int accumulate)
{
    using C = typename Store<T>::C;
    using W = typename Acc<ACC>::W;
    using LD = VecLoad<T>;
    constexpr int TL = kXfTile;                  // lines per tile side
    constexpr int EL = TL + 2 * kXfHalo;         // lines per side with the border
    const int tid = (int)threadIdx.x;
    const AxisDev& X0 = P.ax[0];
    const AxisDev& X1 = P.ax[1];
    const AxisDev& X2 = P.ax[2];
    const int n0 = X0.n, n1 = X1.n, n2 = X2.n;
    const int ylo = SA.flo[0][0], yhi = ylo + SA.fn[0][0];
    const int zlo = SA.flo[0][1], zhi = zlo + SA.fn[0][1];
    const int nty = (SA.fn[0][0] + TL - 1) / TL;
    const int yb = ylo + (tile % nty) * TL;
    const int zb = zlo + (tile / nty) * TL;
    const int x0 = G.r0;
    const int w = min(kSW, G.xend - x0);
    const int g0 = x0 < n0 / 2 ? 0 : n0 - kStage;
    const int64_t s1 = P.ssrc[1], s2 = P.ssrc[2];
    const int64_t xb = P.tshift[0] + g0;
    // the strip's group of every line of the tile and border
    for (int g = tid; g < EL * EL; g += kShellBlock) {
        const int yy = min(max(yb - kXfHalo + g % EL, 0), n1 - 1);
        const int zz = min(max(zb - kXfHalo + g / EL, 0), n2 - 1);
        C v[kSW];
        LD::ld(src + xb + (x0 - g0) + yy * s1 + zz * s2, kSW, true, v);
#pragma unroll
        for (int i = 0; i < kSW; ++i) Store<T>::st(grp[g], i, (double)v[i]);
    }
    // the 16 values at the end of each of the tile's lines
#pragma unroll
    for (int m = 0; m < (TL * TL * (kStage / kSW)) / kShellBlock; ++m) {
        const int g = tid + m * kShellBlock;
        const int c = g >> 2, part = g & 3;
        const int yy = min(yb + c % TL, n1 - 1);
        const int zz = min(zb + c / TL, n2 - 1);
        C v[kSW];
        LD::ld(src + xb + part * kSW + yy * s1 + zz * s2, kSW, true, v);
#pragma unroll
        for (int i = 0; i < kSW; ++i) Store<T>::st(full[c], part * kSW + i, (double)v[i]);
    }
    __syncthreads();
    const int ty = tid % TL, tz = tid / TL;
    const int y = yb + ty, z = zb + tz;
    if (y >= yhi || z >= zhi) return;
    Acc<ACC> acc[kSW];
    auto ldx = [&](int64_t j) -> double { return (double)Store<T>::ld(full[tid], j - g0); };
#pragma unroll
    for (int i = 0; i < kSW; ++i)
        if (i < w) strip::x_term_point<T, ACC, KM>(P, X0, x0 + i, 0, ldx, acc[i]);
    // y and z: regular rows, taps from the staged groups
    const int self = (tz + kXfHalo) * EL + ty + kXfHalo;
    auto regular = [&](const AxisDev& X, int r, int step) {
        const W* wt = Acc<ACC>::table(X) + (X.uniform ? 0 : r);
        const int64_t ws = X.uniform ? 1 : X.n;
        W wv[KM];
        C xv[KM][kSW];
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            if (k < X.K) {
                wv[k] = tab(wt, k * ws);
                const int L = self + (X.dmin + k) * step;
#pragma unroll
                for (int i = 0; i < kSW; ++i) xv[k][i] = (C)Store<T>::ld(grp[L], i);
            }
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            if (k < X.K) {
#pragma unroll
                for (int i = 0; i < kSW; ++i)
                    if (i < w) acc[i].add(wv[k], (W)xv[k][i]);
            }
        }
    };
    regular(X1, y, 1);
    regular(X2, z, EL);
    const int64_t idx[3] = {x0, y, z};
    strip::store_item<T, ACC, LD>(P, idx, w, true, acc, du, accumulate);
}

template <typename T, int ACC, int KM, bool VEC>
__global__ __launch_bounds__(kShellBlock) void k_shell_strips(const PlanDev* __restrict__ Pp,
                                                             const StripArgs SA,
                                                             T* __restrict__ du,
                                                             const T* __restrict__ src,
                                                             int accumulate)
{
    // per thread 16 staged values (general and face items; kind 4: the ends of the tile's
    // lines) and parked accumulators; kind 4: the strip's 4 values of every line of the tile and
    // its border
    __shared__ T stage[kShellBlock][kStage + 1];
    __shared__ strip::Slot<ACC> slots[kShellBlock][kSW + 1];
    __shared__ __attribute__((aligned(16))) T xgrp[(kXfTile + 2 * kXfHalo) * (kXfTile + 2 * kXfHalo)][kSW];
    // the block's segment and row: uniform, so their fields are scalar loads.  The blocks of
    // the x-face tiles come first in hardware order, spread over the XCDs; the rest in
    // XCD-contiguous order
    const int bi = (int)blockIdx.x;
    const int lb = bi < SA.nfirst ? bi : SA.nfirst + (int)strip_block(bi - SA.nfirst, (int)gridDim.x - SA.nfirst);
    int s = 0;
    while (s + 1 < SA.nseg && lb >= SA.seg[s + 1].first) ++s;
    const int rb = lb - SA.seg[s].first;
    const int nr = SA.seg[s].nr;
    const int j = rb % nr;
    const int l = (rb / nr) * kShellBlock + (int)threadIdx.x;
    if constexpr (VEC && KM / 2 <= kXfHalo) {
        if (SA.seg[s].kind == 4) {
            // one tile of lines per block (the block-wide branch precedes the block barrier)
            if (rb >= SA.seg[s].items / kShellBlock) return;
            xface_tile<T, ACC, KM>(*Pp, SA, SA.seg[s], rb, stage, xgrp, du, src, accumulate);
            return;
        }
    }
    strip::eval_item<T, ACC, KM, VEC, VecLoad<T>>(*Pp, SA, s, j, l, stage[threadIdx.x], slots[threadIdx.x],
                                                  du, src, accumulate);
}

}  // namespace shell
}  // namespace dop
