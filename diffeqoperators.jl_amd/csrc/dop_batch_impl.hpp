// dop_batch_impl.hpp -- device side of the persistent batched launches, instantiated per
// storage type by dop_batch_f64.hip / dop_batch_f32.hip / dop_batch_bf16.hip.
//
// The host turns a batch into a work list: every 2-D problem with a tiled path contributes
// row-chunk tiles (4 wave segments x `chunk` rows) plus chunks of its shell points; every
// other problem (1-D, 3-D, ...) contributes chunks of points for the generic evaluator.
// The list is sorted into groups of one kind, window width and accumulator; each group is
// one launch of a kernel specialised for it (registers sized for that case alone), whose
// blocks walk the group's items with a static stride.  A 2-D tile keeps one point per lane and
// takes the x window from neighbouring lanes with DPP wavefront shifts (wave_shr/wave_shl
// by one lane per step), so a wave's 64 lanes yield 64-2H outputs per row; the y window
// comes from a register ring with PF rows of prefetch, as in the single-plan kernel.
// Every path sums a point's taps in the generic order with the plan's accumulator (f64,
// f32 or compensated f32), so batched and single results agree bitwise.
//
// Epilogues: plain apply (du = L Q u, optionally += du), or one stage of classical RK4.
#pragma once
#include <hip/hip_runtime.h>

#include "dop_batch.hpp"
#include "dop_eval.hpp"
#include "dop_ring.hpp"

namespace dop {
namespace batch {

// lane i <- lane i-1 (wave_shr:1) / lane i <- lane i+1 (wave_shl:1); edge lanes get 0
__device__ inline float shr1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                  0x138, 0xf, 0xf, false));
}
__device__ inline float shl1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                  0x130, 0xf, 0xf, false));
}
__device__ inline double shr1(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ inline double shl1(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <typename T>
struct StageBufs {
    const T* X;   // stencil input
    T* out;       // apply: du
    T* u;
    T* acc;
    T* nxt;
};

// store of one point: plain apply, or RK4 stage `stage` (see dop_kernels.hip)
template <typename T>
__device__ inline void epilogue(const BProb& B, const StageBufs<T>& S, int stage, int accumulate,
                                int64_t p, double v)
{
    if (stage < 0) {
        if (accumulate) v += (double)Store<T>::ld(S.out, p);
        Store<T>::st(S.out, p, v);
        return;
    }
    const double k = B.scale * v;
    const double dt = B.dt;
    if (stage == 3) {
        Store<T>::st(S.u, p, (double)Store<T>::ld(S.acc, p) + (dt / 6.0) * k);
    } else {
        const double up = (double)Store<T>::ld(S.u, p);
        if (stage == 0) Store<T>::st(S.acc, p, up + (dt / 6.0) * k);
        else Store<T>::st(S.acc, p, (double)Store<T>::ld(S.acc, p) + (dt / 3.0) * k);
        const double h = (stage == 2) ? dt : 0.5 * dt;
        Store<T>::st(S.nxt, p, up + h * k);
    }
}

template <typename T>
__device__ inline StageBufs<T> stage_bufs(const BProb& B, int stage)
{
    StageBufs<T> S;
    S.out = (T*)B.out;
    S.u = (T*)B.u;
    S.acc = (T*)B.acc;
    if (stage < 0) {
        S.X = (const T*)B.src;
        S.nxt = nullptr;
    } else {
        S.X = stage == 0 ? (const T*)B.u : stage == 1 ? (const T*)B.s1
            : stage == 2 ? (const T*)B.s2 : (const T*)B.s1;
        S.nxt = stage == 1 ? (T*)B.s2 : (T*)B.s1;
    }
    return S;
}

template <int ACC>
__device__ inline const typename Acc<ACC>::W* btab(const BProb& B, int a);
template <>
__device__ inline const double* btab<kAccF64>(const BProb& B, int a) { return B.w64[a]; }
template <>
__device__ inline const float* btab<kAccF32>(const BProb& B, int a) { return B.w32[a]; }
template <>
__device__ inline const float* btab<kAccF32Kahan>(const BProb& B, int a) { return B.w32[a]; }
template <>
__device__ inline const float* btab<kAccF32Wide>(const BProb& B, int a) { return B.w32[a]; }

// What the epilogue of STAGE reads at its point besides the stencil (STAGE -2: apply with
// accumulate, -1: apply, 0..3: RK4 stage): side a = du (-2) or u (0..2), side b = acc (1..3).
template <int STAGE>
constexpr bool kSideA = STAGE == -2 || (STAGE >= 0 && STAGE <= 2);
template <int STAGE>
constexpr bool kSideB = STAGE >= 1;

// the epilogue with its side values already loaded; same arithmetic as epilogue()
template <typename T, int STAGE>
__device__ inline void epilogue_side(const BProb& B, const StageBufs<T>& S, int64_t p, double v,
                                     double sa, double sb)
{
    if constexpr (STAGE == -1) {
        Store<T>::st(S.out, p, v);
    } else if constexpr (STAGE == -2) {
        Store<T>::st(S.out, p, v + sa);
    } else {
        const double k = B.scale * v;
        const double dt = B.dt;
        if constexpr (STAGE == 3) {
            Store<T>::st(S.u, p, sb + (dt / 6.0) * k);
        } else {
            if constexpr (STAGE == 0) Store<T>::st(S.acc, p, sa + (dt / 6.0) * k);
            else Store<T>::st(S.acc, p, sb + (dt / 3.0) * k);
            const double h = (STAGE == 2) ? dt : 0.5 * dt;
            Store<T>::st(S.nxt, p, sa + h * k);
        }
    }
}

// ---------------------------------------------------------------------------------------
// 2-D tile: wave segment `seg` (64 lanes, 64-2H outputs), rows [y0, y1).  The stencil rows
// come through a register ring PF rows ahead; the epilogue's side values (u, acc or du) of
// row y + PF are loaded while row y is computed, so no step waits on a fresh load.
template <typename T, int ACC, int K, int STAGE>
__device__ void tile2d(const BProb& B, const StageBufs<T>& S, int xg, int ch)
{
    using W = typename Acc<ACC>::W;
    constexpr int H = K / 2;
    constexpr int PF = 4;
    constexpr int R = K + PF;
    constexpr int OUTW = 64 - 2 * H;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int seg = xg * 4 + wave;
    if (seg >= B.nseg) return;
    const int y0 = B.lo[1] + ch * B.chunk;
    const int y1 = min(y0 + B.chunk, B.hi[1]);
    const int64_t n0 = B.n0;
    const int n1 = (int)B.n1;
    const int64_t s1 = B.ys;
    const int64_t x = (int64_t)B.lo[0] + (int64_t)seg * OUTW + (lane - H);   // this lane's column
    const bool out_ok = lane >= H && lane < 64 - H && x < B.hi[0];
    const W* wxt = btab<ACC>(B, 0);
    const W* wyt = btab<ACC>(B, 1);
    W wx[K];
    {
        const int64_t xr = B.uniform[0] ? 0 : min(max(x, (int64_t)0), n0 - 1);
#pragma unroll
        for (int k = 0; k < K; ++k) wx[k] = wxt[xr * K + k];
    }
    using SC = typename Store<T>::C;
    const T* sidea = STAGE == -2 ? S.out : S.u;
    const T* sideb = S.acc;
    SC ea[R], eb[R];
    // rows past the chunk (the prefetch of its last steps) reload the chunk's last row or
    // last tap row, already in cache, instead of reading rows this wave never uses
    auto load_side = [&](int row, SC* a, SC* b) {
        row = min(max(row, 0), min(y1 - 1, n1 - 1));
        const int64_t p = (int64_t)row * s1 + min(max(x, (int64_t)0), n0 - 1);
        if constexpr (kSideA<STAGE>) *a = Store<T>::ld(sidea, p);
        if constexpr (kSideB<STAGE>) *b = Store<T>::ld(sideb, p);
    };
    W q[R];
    // clamped loads on every lane: a lane outside [0, n0) holds another column's value and
    // stores nothing; its neighbours' x windows never reach it (the box is H from the edges)
    const int64_t xc = min(max(x, (int64_t)0), n0 - 1);
    auto load_row = [&](int row) -> W {
        row = min(max(row, 0), min(y1 + H - 1, n1 - 1));
        return (W)Store<T>::ld(S.X, (int64_t)row * s1 + xc);
    };
#pragma unroll
    for (int s = 0; s < R - 1; ++s) q[s] = load_row(y0 - H + s);
#pragma unroll
    for (int s = 0; s < PF; ++s) load_side(y0 + s, &ea[s], &eb[s]);
    for (int yb = y0; yb < y1; yb += R) {
        auto step = [&](auto jc) -> bool {
            constexpr int j = decltype(jc)::value;
            const int y = yb + j;
            if (y >= y1) return false;
            q[(j + K - 1 + PF) % R] = load_row(y + H + PF);
            load_side(y + PF, &ea[(j + PF) % R], &eb[(j + PF) % R]);
            // x window of the centre row from the neighbouring lanes
            W xw[K];
            const W c = q[(j + H) % R];
            xw[H] = c;
            W t = c;
#pragma unroll
            for (int m = 1; m <= H; ++m) {
                t = shl1(t);
                xw[H + m] = t;
            }
            t = c;
#pragma unroll
            for (int m = 1; m <= H; ++m) {
                t = shr1(t);
                xw[H - m] = t;
            }
            const int64_t yr = B.uniform[1] ? 0 : y;
            Acc<ACC> acc;
#pragma unroll
            for (int k = 0; k < K; ++k) acc.add(wx[k], xw[k]);
#pragma unroll
            for (int k = 0; k < K; ++k) acc.add(wyt[yr * K + k], q[(j + k) % R]);
            if (out_ok)
                epilogue_side<T, STAGE>(B, S, (int64_t)y * s1 + x, acc.value(),
                                        kSideA<STAGE> ? (double)ea[j] : 0.0,
                                        kSideB<STAGE> ? (double)eb[j] : 0.0);
            return true;
        };
        if (!ring_steps<0, R>(step)) break;
    }
}

// 1-D tile: wave segment `seg` (64 lanes, 64-2H outputs) of the regular rows
template <typename T, int ACC, int K, int STAGE>
__device__ void tile1d(const BProb& B, const StageBufs<T>& S, int xg)
{
    using W = typename Acc<ACC>::W;
    using SC = typename Store<T>::C;
    constexpr int H = K / 2;
    constexpr int OUTW = 64 - 2 * H;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int seg = xg * 4 + wave;
    if (seg >= B.nseg) return;
    const int64_t n0 = B.n0;
    const int64_t x = (int64_t)B.lo[0] + (int64_t)seg * OUTW + (lane - H);
    const bool out_ok = lane >= H && lane < 64 - H && x < B.hi[0];
    // clamped loads: a lane outside [0, n0) stores nothing and no stored window reaches it
    const int64_t xc = min(max(x, (int64_t)0), n0 - 1);
    const W* wxt = btab<ACC>(B, 0);
    const int64_t xr = B.uniform[0] ? 0 : xc;
    const W c = (W)Store<T>::ld(S.X, xc);
    SC sa = SC(0), sb = SC(0);
    if constexpr (kSideA<STAGE>) sa = Store<T>::ld(STAGE == -2 ? S.out : S.u, xc);
    if constexpr (kSideB<STAGE>) sb = Store<T>::ld(S.acc, xc);
    W xw[K];
    xw[H] = c;
    W t = c;
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        t = shl1(t);
        xw[H + m] = t;
    }
    t = c;
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        t = shr1(t);
        xw[H - m] = t;
    }
    Acc<ACC> acc;
#pragma unroll
    for (int k = 0; k < K; ++k) acc.add(wxt[xr * K + k], xw[k]);
    if (out_ok) epilogue_side<T, STAGE>(B, S, x, acc.value(), (double)sa, (double)sb);
}

// generic chunk: points [c*kChunkPts, ...) of the problem's shell boxes
template <typename T, int ACC>
__device__ void shell_chunk(const BProb& B, const StageBufs<T>& S, int stage, int accumulate,
                            int64_t c)
{
    const PlanDev& P = *B.plan;
    const ShellArgs& SH = B.shell;
    const int64_t total = SH.start[SH.nbox];
    const int64_t first = c * kChunkPts;
    for (int j = threadIdx.x; j < kChunkPts; j += kBlock) {
        const int64_t p = first + j;
        if (p >= total) break;
        int b = 0;
        while (b + 1 < SH.nbox && p >= SH.start[b + 1]) ++b;
        int64_t l = p - SH.start[b];
        const int64_t e0 = SH.ext[b][0], e1 = SH.ext[b][1];
        int64_t idx[3];
        idx[0] = SH.org[b][0] + l % e0;
        l /= e0;
        idx[1] = SH.org[b][1] + l % e1;
        idx[2] = SH.org[b][2] + l / e1;
        const int64_t lin = idx[0] + idx[1] * P.sdu[1] + idx[2] * P.sdu[2];
        epilogue<T>(B, S, stage, accumulate, lin, eval_point<T, ACC>(P, S.X, idx));
    }
}

// one group of 2-D tiles with window K and accumulator ACC, for epilogue STAGE
template <typename T, int ACC, int K, int STAGE>
__global__ __launch_bounds__(kBlock) void k_batch_tiles(const BProb* __restrict__ probs,
                                                      const BItem* __restrict__ items, int nitems)
{
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BItem I = items[it];
        const BProb& B = probs[I.prob];
        tile2d<T, ACC, K, STAGE>(B, stage_bufs<T>(B, STAGE < 0 ? -1 : STAGE), I.a, I.b);
    }
}

// one group of 1-D tiles with window K and accumulator ACC, for epilogue STAGE
template <typename T, int ACC, int K, int STAGE>
__global__ __launch_bounds__(kBlock) void k_batch_tiles1d(const BProb* __restrict__ probs,
                                                        const BItem* __restrict__ items, int nitems)
{
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BItem I = items[it];
        const BProb& B = probs[I.prob];
        tile1d<T, ACC, K, STAGE>(B, stage_bufs<T>(B, STAGE < 0 ? -1 : STAGE), I.a);
    }
}

// one group of generic point chunks with accumulator ACC
template <typename T, int ACC>
__global__ __launch_bounds__(kBlock) void k_batch_points(const BProb* __restrict__ probs,
                                                       const BItem* __restrict__ items, int nitems,
                                                       int stage, int accumulate)
{
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BItem I = items[it];
        const BProb& B = probs[I.prob];
        shell_chunk<T, ACC>(B, stage_bufs<T>(B, stage), stage, accumulate, I.a);
    }
}

template <typename T, int ACC, int K>
hipError_t launch_tiles(int kind, int st, dim3 gr, const BProb* probs, const BItem* it, int n,
                        hipStream_t s)
{
    switch (st) {
#define DOP_BS(SS)                                                                               \
    case SS:                                                                                      \
        if (kind == 2)                                                                            \
            hipLaunchKernelGGL((k_batch_tiles1d<T, ACC, K, SS>), gr, dim3(kBlock), 0, s, probs,  \
                               it, n);                                                            \
        else                                                                                      \
            hipLaunchKernelGGL((k_batch_tiles<T, ACC, K, SS>), gr, dim3(kBlock), 0, s, probs, it, \
                               n);                                                                \
        return hipGetLastError();
        DOP_BS(-2) DOP_BS(-1) DOP_BS(0) DOP_BS(1) DOP_BS(2) DOP_BS(3)
#undef DOP_BS
    }
    return hipErrorInvalidValue;
}

template <typename T, int ACC>
hipError_t launch_group(const BGroup& g, const BProb* probs, const BItem* items, int stage,
                        int accumulate, int grid, hipStream_t s)
{
    const dim3 gr((unsigned)(g.count < grid ? g.count : grid));
    const BItem* it = items + g.first;
    if (g.kind == 1) {
        hipLaunchKernelGGL((k_batch_points<T, ACC>), gr, dim3(kBlock), 0, s, probs, it, g.count,
                           stage, accumulate);
        return hipGetLastError();
    }
    const int st = stage < 0 ? (accumulate ? -2 : -1) : stage;
    switch (g.K) {
    case 3: return launch_tiles<T, ACC, 3>(g.kind, st, gr, probs, it, g.count, s);
    case 5: return launch_tiles<T, ACC, 5>(g.kind, st, gr, probs, it, g.count, s);
    case 7: return launch_tiles<T, ACC, 7>(g.kind, st, gr, probs, it, g.count, s);
    case 9: return launch_tiles<T, ACC, 9>(g.kind, st, gr, probs, it, g.count, s);
    case 11: return launch_tiles<T, ACC, 11>(g.kind, st, gr, probs, it, g.count, s);
    }
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_group_any(const BGroup& g, const BProb* probs, const BItem* items, int stage,
                            int accumulate, int grid, hipStream_t s)
{
    if constexpr (sizeof(T) == 8) {
        if (g.acc != kAccF64) return hipErrorInvalidValue;
        return launch_group<T, kAccF64>(g, probs, items, stage, accumulate, grid, s);
    } else {
        // f32 / bf16 storage with f64 accumulation has point groups only (no tiles)
        if (g.acc == kAccF32Kahan)
            return launch_group<T, kAccF32Kahan>(g, probs, items, stage, accumulate, grid, s);
        if (g.acc == kAccF32Wide)
            return launch_group<T, kAccF32Wide>(g, probs, items, stage, accumulate, grid, s);
        if (g.acc == kAccF64) {
            if (g.kind != 1) return hipErrorInvalidValue;
            return launch_group<T, kAccF64>(g, probs, items, stage, accumulate, grid, s);
        }
        return launch_group<T, kAccF32>(g, probs, items, stage, accumulate, grid, s);
    }
}

}  // namespace batch
}  // namespace dop
