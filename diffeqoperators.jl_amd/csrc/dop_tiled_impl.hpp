// dop_tiled_impl.hpp -- tiled interior kernels for 2-D / 3-D Q*u plans, instantiated per
// storage type by dop_tiled_f64.hip / dop_tiled_f32.hip / dop_tiled_bf16.hip.
//
// 2-D: a wave loads an aligned 64*VX-point row segment and marches a chunk of rows.  The K
//      rows of the y stencil live in a register ring of K+PF rows (PF rows prefetched
//      ahead), so every row of u is loaded once per chunk; the x window of the centre row
//      comes from neighbouring lanes (DPP), segments overlapping by the x halo.
// 3-D: a block of W waves owns W consecutive y rows of one x segment and marches a chunk
//      of z planes with the same register ring along z.  Each step stages the centre plane
//      (own rows + 2H halo rows, each with H x-halo columns) in LDS, double-buffered, one
//      barrier per step; x and y neighbours are read from LDS, z neighbours from the ring.
// Both sum, per point, the x window, then y, then z, each in tap order: the same FMA
// sequence as the generic path, so the two agree bitwise on finite inputs.
//
// The points outside the interior box (the shell) are evaluated by the first blocks of the
// same launch with the generic per-point path, so they overlap the tiles.
//
// Blocks are dispatched round-robin over the 8 XCDs; with xcd_remap the logical tile of a
// block is chosen so that each XCD walks a contiguous band of tiles (halo rows of
// neighbouring tiles then hit the same L2).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dop_eval.hpp"
#include "dop_kernels.hpp"
#include "dop_ring.hpp"
#include "dop_shell.hpp"

namespace dop {
namespace tiled {


// ---------------------------------------------------------------------------------------
// VX consecutive elements per lane
template <typename T, int VX>
struct VecIO;

template <>
struct VecIO<double, 2> {
    __device__ static void ld(const double* p, double* o)
    {
        const double2 v = *reinterpret_cast<const double2*>(p);
        o[0] = v.x;
        o[1] = v.y;
    }
    __device__ static void st(double* p, const double* v)
    {
        *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    }
};

template <>
struct VecIO<float, 4> {
    __device__ static void ld(const float* p, float* o)
    {
        const float4 v = *reinterpret_cast<const float4*>(p);
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
        o[3] = v.w;
    }
    __device__ static void st(float* p, const float* v)
    {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    }
};

template <>
struct VecIO<float, 2> {
    __device__ static void ld(const float* p, float* o)
    {
        const float2 v = *reinterpret_cast<const float2*>(p);
        o[0] = v.x;
        o[1] = v.y;
    }
    __device__ static void st(float* p, const float* v)
    {
        *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    }
};

template <>
struct VecIO<bf16, 2> {
    __device__ static void ld(const bf16* p, float* o)
    {
        const uint32_t v = *reinterpret_cast<const uint32_t*>(p);
        o[0] = __builtin_bit_cast(float, v << 16);
        o[1] = __builtin_bit_cast(float, v & 0xffff0000u);
    }
    __device__ static void st(bf16* p, const float* v)
    {
        *reinterpret_cast<uint32_t*>(p) =
            (uint32_t)f32_to_bf16(v[0]) | ((uint32_t)f32_to_bf16(v[1]) << 16);
    }
};

template <>
struct VecIO<bf16, 4> {
    __device__ static void ld(const bf16* p, float* o)
    {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        o[0] = __builtin_bit_cast(float, v.x << 16);
        o[1] = __builtin_bit_cast(float, v.x & 0xffff0000u);
        o[2] = __builtin_bit_cast(float, v.y << 16);
        o[3] = __builtin_bit_cast(float, v.y & 0xffff0000u);
    }
    __device__ static void st(bf16* p, const float* v)
    {
        uint2 w;
        w.x = (uint32_t)f32_to_bf16(v[0]) | ((uint32_t)f32_to_bf16(v[1]) << 16);
        w.y = (uint32_t)f32_to_bf16(v[2]) | ((uint32_t)f32_to_bf16(v[3]) << 16);
        *reinterpret_cast<uint2*>(p) = w;
    }
};

// two 16-byte loads per lane
template <>
struct VecIO<double, 4> {
    __device__ static void ld(const double* p, double* o)
    {
        VecIO<double, 2>::ld(p, o);
        VecIO<double, 2>::ld(p + 2, o + 2);
    }
    __device__ static void st(double* p, const double* v)
    {
        VecIO<double, 2>::st(p, v);
        VecIO<double, 2>::st(p + 2, v + 2);
    }
};

template <>
struct VecIO<float, 8> {
    __device__ static void ld(const float* p, float* o)
    {
        VecIO<float, 4>::ld(p, o);
        VecIO<float, 4>::ld(p + 4, o + 4);
    }
    __device__ static void st(float* p, const float* v)
    {
        VecIO<float, 4>::st(p, v);
        VecIO<float, 4>::st(p + 4, v + 4);
    }
};

template <>
struct VecIO<bf16, 8> {
    __device__ static void ld(const bf16* p, float* o)
    {
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            o[2 * i] = __builtin_bit_cast(float, w[i] << 16);
            o[2 * i + 1] = __builtin_bit_cast(float, w[i] & 0xffff0000u);
        }
    }
    __device__ static void st(bf16* p, const float* v)
    {
        uint32_t w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            w[i] = (uint32_t)f32_to_bf16(v[2 * i]) | ((uint32_t)f32_to_bf16(v[2 * i + 1]) << 16);
        *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
    }
};

template <typename T>
struct VecIO1 {
    using C = typename Store<T>::C;
    __device__ static void ld(const T* p, C* o) { o[0] = (C)Store<T>::ld(p, 0); }
    __device__ static void st(T* p, const C* v) { Store<T>::st(p, 0, (double)v[0]); }
};
template <>
struct VecIO<double, 1> : VecIO1<double> {};
template <>
struct VecIO<float, 1> : VecIO1<float> {};
template <>
struct VecIO<bf16, 1> : VecIO1<bf16> {};

template <typename C>
__device__ inline C fmac(C a, C b, C c);
template <>
__device__ inline double fmac<double>(double a, double b, double c) { return fma(a, b, c); }
template <>
__device__ inline float fmac<float>(float a, float b, float c) { return fmaf(a, b, c); }

// The running sum of a point: the storage's compute type C, or (ACC = kAccF32Wide) f32 weights
// and values with products formed and summed in f64 -- the generic evaluator's Acc<ACC>
// arithmetic, so tiled and generic results stay bitwise equal.
template <typename C, int ACC>
using Sum = typename std::conditional<ACC == kAccF32Wide, double, C>::type;

template <typename C, int ACC>
__device__ inline Sum<C, ACC> fsum(C w, C x, Sum<C, ACC> s)
{
    return fmac<Sum<C, ACC>>((Sum<C, ACC>)w, (Sum<C, ACC>)x, s);
}

// VX sums stored as T (plus the old du when accumulating, added in f64 like the generic store)
template <typename T, int VX, int ACC>
__device__ inline void store_sums(T* po, const Sum<typename Store<T>::C, ACC>* sums, int accumulate)
{
    using C = typename Store<T>::C;
    C out[VX];
    if (accumulate) {
        C old[VX];
        VecIO<T, VX>::ld(po, old);
#pragma unroll
        for (int v = 0; v < VX; ++v) out[v] = (C)((double)sums[v] + (double)old[v]);
    } else {
#pragma unroll
        for (int v = 0; v < VX; ++v) out[v] = (C)sums[v];
    }
    VecIO<T, VX>::st(po, out);
}

template <typename C>
__device__ inline const C* wtab(const FastArgs& A, int a);
template <>
__device__ inline const double* wtab<double>(const FastArgs& A, int a) { return A.w64[a]; }
template <>
__device__ inline const float* wtab<float>(const FastArgs& A, int a) { return A.w32[a]; }

// lane i <- lane i-1 (wave_shr:1) / lane i <- lane i+1 (wave_shl:1); lanes without a source
// read 0
__device__ inline float dpp_shr1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                  0x138, 0xf, 0xf, false));
}
__device__ inline float dpp_shl1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                  0x130, 0xf, 0xf, false));
}
__device__ inline double dpp_shr1(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ inline double dpp_shl1(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Role of block i of a launch with `ntile` tile blocks and `nshell` shell blocks: the two are
// interleaved in groups of 8 (one group of each per round over the 8 XCDs), so the
// latency-bound shell runs beside the bandwidth-bound tiles and tile block k still lands on
// XCD k mod 8.  Returns true for a tile (its index in *idx), false for a shell chunk.
__device__ inline bool block_role(int i, int ntile, int nshell, int* idx)
{
    const int m = min(ntile, nshell) / 8 * 8;
    if (i < 2 * m) {
        *idx = (i >> 4) * 8 + (i & 7);
        return (i & 15) < 8;
    }
    const int rest = i - 2 * m;
    if (rest < ntile - m) {
        *idx = m + rest;
        return true;
    }
    *idx = m + (rest - (ntile - m));
    return false;
}

// logical tile of this block: each XCD (blocks b, b+8, b+16, ...) gets a contiguous band
__device__ inline int tile_of(int b, int nb, int remap)
{
    if (!remap) return b;
    const int per = nb >> 3;
    if (b >= per * 8) return b;
    return (b & 7) * per + (b >> 3);
}

template <typename T>
constexpr int acc_of() { return std::is_same<T, double>::value ? kAccF64 : kAccF32; }

// generic evaluation of shell points [first, first + count)
template <typename T, int ACC>
__device__ inline void shell_points(const PlanDev& P, const ShellArgs& S, T* __restrict__ du,
                                    const T* __restrict__ src, int accumulate, int64_t first,
                                    int64_t count)
{
    const int64_t total = S.start[S.nbox];
    const int64_t end = min(first + count, total);
    for (int64_t p = first + threadIdx.x; p < end; p += blockDim.x) {
        int b = 0;
        while (b + 1 < S.nbox && p >= S.start[b + 1]) ++b;
        int64_t l = p - S.start[b];
        const int64_t e0 = S.ext[b][0], e1 = S.ext[b][1];
        int64_t idx[3];
        if (S.start[b + 1] - S.start[b] < ((int64_t)1 << 31)) {   // 32-bit divisions
            const uint32_t l32 = (uint32_t)l, f0 = (uint32_t)e0, f1 = (uint32_t)e1;
            const uint32_t q0 = l32 / f0;
            idx[0] = S.org[b][0] + (l32 - q0 * f0);
            const uint32_t q1 = q0 / f1;
            idx[1] = S.org[b][1] + (q0 - q1 * f1);
            idx[2] = S.org[b][2] + q1;
        } else {
            idx[0] = S.org[b][0] + l % e0;
            l /= e0;
            idx[1] = S.org[b][1] + l % e1;
            idx[2] = S.org[b][2] + l / e1;
        }
        const int64_t lin = idx[0] + idx[1] * P.sdu[1] + idx[2] * P.sdu[2];
        double v = eval_point<T, ACC>(P, src, idx);
        if (accumulate) v += (double)Store<T>::ld(du, lin);
        Store<T>::st(du, lin, v);
    }
}

constexpr int kShellPtsPerThread = 4;

// x window of a lane's VX consecutive points from its own registers and its neighbours'
// through DPP shifts: xw[e] = u(xs - H + e), e < VX + 2H.  Lanes within ceil(H/VX) of the
// wave's ends get zeros for what lies beyond the wave; those lanes only load.
template <typename T, int VX, int H>
__device__ inline void x_window_dpp(const typename Store<T>::C* own, typename Store<T>::C* xw)
{
    using C = typename Store<T>::C;
    constexpr int DL = (H + VX - 1) / VX;   // lanes reached on each side
#pragma unroll
    for (int v = 0; v < VX; ++v) xw[H + v] = own[v];
    C l[VX], r[VX];
#pragma unroll
    for (int v = 0; v < VX; ++v) {
        l[v] = own[v];
        r[v] = own[v];
    }
#pragma unroll
    for (int d = 1; d <= DL; ++d) {
#pragma unroll
        for (int v = 0; v < VX; ++v) {
            l[v] = dpp_shr1(l[v]);   // now holds lane (lane - d)'s value v
            r[v] = dpp_shl1(r[v]);   // lane (lane + d)'s value v
        }
#pragma unroll
        for (int v = 0; v < VX; ++v) {
            const int el = H - d * VX + v;              // u(xs - d*VX + v)
            if (el >= 0 && el < H) xw[el] = l[v];
            const int er = H + VX + (d - 1) * VX + v;   // u(xs + d*VX + v)
            if (er < VX + 2 * H) xw[er] = r[v];
        }
    }
}

// lane i <- lane i-1 (lane 0 keeps `fill`) / lane i <- lane i+1 (lane 63 keeps `fill`)
__device__ inline float dpp_shr1_fill(float fill, float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                                  __builtin_bit_cast(int, v),
                                                                  0x138, 0xf, 0xf, false));
}
__device__ inline float dpp_shl1_fill(float fill, float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                                  __builtin_bit_cast(int, v),
                                                                  0x130, 0xf, 0xf, false));
}
__device__ inline double dpp2(double fill, double v, int ctrl_is_shl)
{
    const uint64_t f = __builtin_bit_cast(uint64_t, fill), b = __builtin_bit_cast(uint64_t, v);
    int lo, hi;
    if (ctrl_is_shl) {
        lo = __builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)b, 0x130, 0xf, 0xf, false);
        hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32), (int)(uint32_t)(b >> 32), 0x130,
                                         0xf, 0xf, false);
    } else {
        lo = __builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)b, 0x138, 0xf, 0xf, false);
        hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32), (int)(uint32_t)(b >> 32), 0x138,
                                         0xf, 0xf, false);
    }
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ inline double dpp_shr1_fill(double fill, double v) { return dpp2(fill, v, 0); }
__device__ inline double dpp_shl1_fill(double fill, double v) { return dpp2(fill, v, 1); }

// x window xw[e] = u(xs - H + e), e < VX + 2H, of a lane's VX points: its own values, its
// neighbours' through DPP, and at the wave's ends the edge groups eg[g] (lane 0: group
// -(g+1) of the segment, lane 63: group 64+g).
template <typename C, int VX, int H>
__device__ inline void x_window_edges(const C* own, const C (*eg)[VX], C* xw)
{
    constexpr int NE = (H + VX - 1) / VX;
#pragma unroll
    for (int v = 0; v < VX; ++v) xw[H + v] = own[v];
    C l[VX], r[VX];
#pragma unroll
    for (int v = 0; v < VX; ++v) {
        l[v] = own[v];
        r[v] = own[v];
    }
#pragma unroll
    for (int d = 1; d <= NE; ++d) {
#pragma unroll
        for (int v = 0; v < VX; ++v) {
            const int el = H - d * VX + v;              // u(xs - d*VX + v)
            const int er = H + VX + (d - 1) * VX + v;   // u(xs + d*VX + v)
            // value v of the group at distance d lands in the window (and then so does every
            // nearer one, so the shift chain of v runs up to the farthest such d)
            if (el >= 0) {
                l[v] = dpp_shr1_fill(eg[d - 1][v], l[v]);
                xw[el] = l[v];
            }
            if (er < VX + 2 * H) {
                r[v] = dpp_shl1_fill(eg[d - 1][v], r[v]);
                xw[er] = r[v];
            }
        }
    }
}

// H consecutive values from an element-aligned address, in the compute type
template <typename T, int H>
struct EdgeIO {
    using C = typename Store<T>::C;
    __device__ static void ld(const T* p, C* o)
    {
        if constexpr (!std::is_same<T, bf16>::value && (H == 2 || H == 4)) {
            typedef T V __attribute__((ext_vector_type(H), aligned(sizeof(T))));
            const V v = *reinterpret_cast<const V*>(p);
#pragma unroll
            for (int h = 0; h < H; ++h) o[h] = v[h];
        } else {
#pragma unroll
            for (int h = 0; h < H; ++h) o[h] = Store<T>::ld(p, h);
        }
    }
};

// x window of the edge-loaded 2-D form (H <= VX): own values, the neighbours' through DPP, and
// at the wave's ends the H values beyond the segment, e[] (lane 0: u(xs0 - H + j), lane 63:
// u(xs0 + 64*VX + j)).  xw[e] = u(xs - H + e), e < VX + 2H.
template <typename C, int VX, int H>
__device__ inline void x_window_edgeh(const C* own, const C* e, C* xw)
{
    static_assert(H <= VX, "the window reaches one lane on each side");
#pragma unroll
    for (int v = 0; v < VX; ++v) xw[H + v] = own[v];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        xw[j] = dpp_shr1_fill(e[j], own[VX - H + j]);
        xw[H + VX + j] = dpp_shl1_fill(e[j], own[j]);
    }
}

// ---------------------------------------------------------------------------------------
// 2-D: a wave loads an aligned 64*VX-point row segment and marches a chunk of rows with a
// register ring of K + PF rows (PF rows prefetched ahead).  Overlapping form (EDG false): the
// segment's first and last PADL points (whole lanes) are the x halo, segments overlap by
// 2*PADL, the x window comes from neighbouring lanes through DPP, and only the middle lanes
// store.  Edge-loaded form (EDG true, H <= VX): segments start at multiples of 64*VX points
// and do not overlap (every row read is whole 128-byte lines and no point is read twice by
// neighbouring segments); each lane also loads the H values beside the segment on one side per
// row -- lanes 0-31 those left of it, lanes 32-63 those right of it, two addresses per wave --
// and lanes 0 and 63 feed theirs into the DPP window.  Loads are clamped into the array and issued
// on every path, stores are one predicated vector store, so the in-order load counter waits
// stay precise and the ring stays in flight.
template <typename T, int VX, int K, bool UNI, int PF, int ACC, bool EDG = false>
__global__ __launch_bounds__(256) void k_tiled2d(const FastArgs A, const PlanDev* __restrict__ Pp,
                                                const ShellArgs SH, int nshell,
                                                T* __restrict__ du, const T* __restrict__ u,
                                                int accumulate)
{
    using C = typename Store<T>::C;
    constexpr int H = K / 2;
    constexpr int R = K + PF;
    constexpr int NWIN = VX + 2 * H;
    constexpr int PADL = (H + VX - 1) / VX * VX;   // x halo, whole lanes
    constexpr int LEDGE = PADL / VX;               // load-only lanes at each end
    int bi;
    if (!block_role((int)blockIdx.x, A.nblocks, nshell, &bi)) {
        const int64_t per = 256 * kShellPtsPerThread;
        shell_points<T, ACC>(*Pp, SH, du, u, accumulate, (int64_t)bi * per, per);
        return;
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = tile_of(bi, A.nblocks, A.xcd_remap);
    const int xg = b % A.gx;
    const int ch = b / A.gx;
    const int xseg = xg * 4 + wave;
    if (xseg >= A.nxseg) return;
    const int y0 = A.lo[1] + ch * A.chunk;
    const int y1 = min(y0 + A.chunk, A.hi[1]);
    const int64_t n0 = A.n[0];
    const int n1 = (int)A.n[1];
    const int64_t s1 = A.s1;
    static_assert(!EDG || H <= VX, "one edge group per side");
    const int64_t xs0 = (int64_t)A.xbase + (int64_t)xseg * (EDG ? 64 * VX : 64 * VX - 2 * PADL);
    const int64_t xs = xs0 + lane * VX;
    // A lane whose group lies outside [0, n0) loads a clamped (other) group: only outputs in
    // the box are stored and their windows stay within H of the box, inside the array.
    const int64_t xl = min(max(xs, (int64_t)0), n0 - VX);
    const bool st_ok = (EDG || (lane >= LEDGE && lane < 64 - LEDGE)) && xs >= A.lo[0] && xs < A.hi[0];
    // the lane's H edge values (EDG): left of the segment for lanes 0-31, right of it for 32-63
    const int64_t xe = min(max(lane < 32 ? xs0 - H : xs0 + 64 * VX, (int64_t)0), n0 - H);

    const C* wxt = wtab<C>(A, 0);
    const C* wyt = wtab<C>(A, 1);
    C wx[VX][K];
#pragma unroll
    for (int v = 0; v < VX; ++v) {
        int64_t xr = 0;
        if (!UNI) xr = min(max(xs + v, (int64_t)0), n0 - 1);
#pragma unroll
        for (int k = 0; k < K; ++k) wx[v][k] = wxt[xr * K + k];
    }
    C wyu[K];
    if (UNI) {
#pragma unroll
        for (int k = 0; k < K; ++k) wyu[k] = wyt[k];
    }

    const T* ucol = u + xl;
    const T* ecol = u + xe;
    // rows outside the chunk's tap rows (the prefetch of its last steps) reload its last tap
    // row, already in cache, instead of reading rows this wave never uses
    const int ylo = max(y0 - H, 0);
    const int ylast = min(y1 + H - 1, n1 - 1);
    auto load_row = [&](int row, C* dst, C* edst) {
        row = min(max(row, ylo), ylast);
        VecIO<T, VX>::ld(ucol + (int64_t)row * s1, dst);
        if constexpr (EDG) EdgeIO<T, H>::ld(ecol + (int64_t)row * s1, edst);
    };
    const int nrows = y1 - y0;
    // The march over the chunk's rows, downwards (DIR = 1) or upwards (DIR = -1).  Ring slot
    // of a row: (its distance from the first output row along the march + H) mod R, so the
    // slots of a step are compile-time; the taps are summed in ascending row order either way.
    auto march = [&](auto dirc) {
        constexpr int DIR = decltype(dirc)::value;
        const int ystart = DIR > 0 ? y0 : y1 - 1;
        C q[R][VX];
        C qe[EDG ? R : 1][H];   // the row's edge values (EDG)
#pragma unroll
        for (int s = 0; s < R - 1; ++s) load_row(ystart + DIR * (s - H), q[s], qe[EDG ? s : 0]);
        for (int t0 = 0; t0 < nrows; t0 += R) {
            auto step = [&](auto jc) -> bool {
                constexpr int j = decltype(jc)::value;
                const int t = t0 + j;
                if (t >= nrows) return false;
                const int y = ystart + DIR * t;
                constexpr int sl = (j + K - 1 + PF) % R;
                load_row(y + DIR * (H + PF), q[sl], qe[EDG ? sl : 0]);
                C xw[NWIN];
                if constexpr (EDG)
                    x_window_edgeh<C, VX, H>(q[(j + H) % R], qe[(j + H) % R], xw);
                else
                    x_window_dpp<T, VX, H>(q[(j + H) % R], xw);
                C wy[K];
#pragma unroll
                for (int k = 0; k < K; ++k) wy[k] = UNI ? wyu[k] : wyt[(int64_t)y * K + k];
                Sum<C, ACC> out[VX];
#pragma unroll
                for (int v = 0; v < VX; ++v) {
                    Sum<C, ACC> s = 0;
#pragma unroll
                    for (int k = 0; k < K; ++k) s = fsum<C, ACC>(wx[v][k], xw[v + k], s);
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        s = fsum<C, ACC>(wy[k], q[(DIR > 0 ? j + k : j + 2 * H - k) % R][v], s);
                    out[v] = s;
                }
                if (st_ok) store_sums<T, VX, ACC>(du + (int64_t)y * s1 + xs, out, accumulate);
                return true;
            };
            if (!ring_steps<0, R>(step)) break;
        }
    };
    // alternate chunks march towards each other's shared edge: both read the halo rows there at
    // the same time, and neighbouring chunks on one XCD (the remapped order) share them in L2
    if (A.alt && (ch & 1)) march(std::integral_constant<int, -1>{});
    else march(std::integral_constant<int, 1>{});
}

// ---------------------------------------------------------------------------------------
// 3-D: a block of W waves stages W consecutive y rows of one x segment and computes the
// middle W - 2H of them (the first and last H waves load halo rows only).  Every wave walks
// a chunk of z planes with a register ring of K + PF planes of its row, loaded PF + H
// planes ahead, and each step copies its centre plane into LDS (double-buffered, one
// barrier per step), so every staged value was loaded several steps before it is used.
// A wave loads an aligned 64*VX-point row segment whose first and last PADL points (whole
// lanes) are the x halo: segments overlap by 2*PADL and only the middle lanes store.  The
// memory operations of a step are the same on every path (addresses are clamped into the
// array and out-of-range values selected to zero), which keeps the hardware's in-order
// load counters precise so that a wait never drains the prefetched ring.  x and y taps come
// from LDS, z taps from the ring; non-uniform x weights (per column) and z weights (per
// plane of the chunk) are staged in LDS once.
constexpr int kMaxChunk3D = 64;
// planes prefetched beyond the deepest z tap (the ring holds K + PF planes), as deep as the
// registers allow without spills
template <typename T, int K>
constexpr int kPF3D = sizeof(T) == 8 ? 1 : K <= 3 ? 4 : K <= 7 ? 2 : 1;

template <typename T, int VX, int K, bool UNI, int PF, int W, int VY, int ACC>
__global__ __launch_bounds__(64 * W) void k_tiled3d(const FastArgs A, const PlanDev* __restrict__ Pp,
                                                     const ShellArgs SH, int nshell,
                                                     T* __restrict__ du, const T* __restrict__ u,
                                                     int accumulate)
{
    using C = typename Store<T>::C;
    constexpr int H = K / 2;
    constexpr int R = K + PF;
    constexpr int SR = W * VY;                            // staged rows per block
    constexpr int ROWW = 64 * VX;
    constexpr int PADL = (H + VX - 1) / VX * VX;          // x halo, whole lanes
    constexpr int LEDGE = PADL / VX;                      // load-only lanes at each end
    constexpr int NRD = (PADL + H + VX + VX - 1) / VX;    // VX-wide reads covering a window
    static_assert(2 * H < SR, "a block computes at least one row");
    static_assert((64 - 2 * LEDGE - 1) * VX + NRD * VX <= ROWW, "x window inside the row");
    __shared__ __attribute__((aligned(16))) C lds[2][SR][ROWW];
    __shared__ C wzs[UNI ? 1 : kMaxChunk3D][K];
    __shared__ __attribute__((aligned(16))) C wxs[UNI ? 1 : K][UNI ? 1 : ROWW];   // per column
    int bi;
    if (!block_role((int)blockIdx.x, A.nblocks, nshell, &bi)) {   // whole block: no barrier
        const int64_t per = 64 * W * kShellPtsPerThread;
        shell_points<T, ACC>(*Pp, SH, du, u, accumulate, (int64_t)bi * per, per);
        return;
    }

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = tile_of(bi, A.nblocks, A.xcd_remap);
    const int xseg = b % A.nxseg;
    const int t = b / A.nxseg;
    const int yblk = t % A.nyblk;
    const int zch = t / A.nyblk;
    // staged row r of the block is y = yb0 + r; wave w stages rows w + i*W (i < VY) and
    // computes those in [H, SR - H)
    const int yb0 = A.lo[1] + yblk * (SR - 2 * H) - H;
    const int z0 = A.lo[2] + zch * A.chunk;
    const int z1 = min(z0 + A.chunk, A.hi[2]);
    const int64_t n0 = A.n[0];
    const int n1 = (int)A.n[1], n2 = (int)A.n[2];
    const int64_t s1 = A.s1, s2 = A.s2;
    const int64_t xs0 = (int64_t)A.xbase + (int64_t)xseg * (ROWW - 2 * PADL);
    const int64_t xs = xs0 + lane * VX;
    // Loads are clamped into the array.  A lane whose segment lies outside [0, n0) (or a
    // row / plane past the edge) then holds other points' values: only outputs in the box
    // [lo, hi) are stored, and their x / y / z windows reach no further than H points past the
    // box, i.e. never outside the array, so those values are never used.
    const int64_t xl = min(max(xs, (int64_t)0), n0 - VX);  // clamped load column (rows are
                                                           // whole VX groups)
    const bool xst = lane >= LEDGE && lane < 64 - LEDGE && xs >= A.lo[0] && xs < A.hi[0];
    const int lr = min(max(lane, LEDGE), 63 - LEDGE);      // window reads of load-only lanes
    const int wcol = (lr - LEDGE) * VX;                     // first column of the window reads

    int row[VY], y[VY], yc[VY];
    bool computes[VY], st_ok[VY];
#pragma unroll
    for (int i = 0; i < VY; ++i) {
        row[i] = wave + i * W;
        y[i] = yb0 + row[i];
        yc[i] = min(max(y[i], 0), n1 - 1);
        computes[i] = row[i] >= H && row[i] < SR - H;      // wave-uniform
        st_ok[i] = computes[i] && y[i] < A.hi[1] && xst;
    }

    const C* wxt = wtab<C>(A, 0);
    const C* wyt = wtab<C>(A, 1);
    const C* wzt = wtab<C>(A, 2);
    C wxu[K], wzu[K];
    C wy[VY][K];
#pragma unroll
    for (int i = 0; i < VY; ++i) {
        const int64_t yr = UNI ? 0 : yc[i];
#pragma unroll
        for (int k = 0; k < K; ++k) wy[i][k] = wyt[yr * K + k];
    }
    if constexpr (UNI) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            wxu[k] = wxt[k];
            wzu[k] = wzt[k];
        }
    } else {
        // visible to every wave after the first step's barrier
        for (int i = threadIdx.x; i < (z1 - z0) * K; i += 64 * W)
            wzs[i / K][i % K] = wzt[(int64_t)(z0 + i / K) * K + i % K];
        for (int i = threadIdx.x; i < ROWW * K; i += 64 * W) {
            const int c = i / K, k = i % K;
            const int64_t x = min(max(xs0 + c, (int64_t)0), n0 - 1);
            wxs[k][c] = wxt[x * K + k];
        }
    }

    // one VX-wide load per call, on every path, straight into its ring slot
    const T* urow[VY];
#pragma unroll
    for (int i = 0; i < VY; ++i) urow[i] = u + (int64_t)yc[i] * s1 + xl;
    // planes past the chunk's last tap plane reload that plane (cached) instead of reading
    // planes this wave never uses
    const int zlast = min(z1 + H - 1, n2 - 1);
    auto load_plane = [&](int i, int zz, C* dst) {
        zz = min(max(zz, 0), zlast);
        VecIO<T, VX>::ld(urow[i] + (int64_t)zz * s2, dst);
    };

    C q[VY][R][VX];
#pragma unroll
    for (int s = 0; s < R - 1; ++s) {
#pragma unroll
        for (int i = 0; i < VY; ++i) load_plane(i, z0 - H + s, q[i][s]);
    }

    int buf = 0;
    for (int zb = z0; zb < z1; zb += R) {
        auto step = [&](auto jc) -> bool {
            constexpr int j = decltype(jc)::value;
            const int z = zb + j;
            if (z >= z1) return false;   // block-uniform
#pragma unroll
            for (int i = 0; i < VY; ++i) load_plane(i, z + H + PF, q[i][(j + K - 1 + PF) % R]);
#pragma unroll
            for (int i = 0; i < VY; ++i) {
#pragma unroll
                for (int v = 0; v < VX; ++v) lds[buf][row[i]][lane * VX + v] = q[i][(j + H) % R][v];
            }
            __syncthreads();
            C wz[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if constexpr (UNI) wz[k] = wzu[k];
                else wz[k] = wzs[z - z0][k];
            }
#pragma unroll
            for (int i = 0; i < VY; ++i) {
                if (!computes[i]) continue;   // wave-uniform
                C xw[NRD * VX];
#pragma unroll
                for (int r = 0; r < NRD; ++r) {
#pragma unroll
                    for (int v = 0; v < VX; ++v) xw[r * VX + v] = lds[buf][row[i]][wcol + r * VX + v];
                }
                Sum<C, ACC> out[VX];
#pragma unroll
                for (int v = 0; v < VX; ++v) {
                    Sum<C, ACC> s = 0;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        C w;
                        if constexpr (UNI) w = wxu[k];
                        else w = wxs[k][lane * VX + v];
                        s = fsum<C, ACC>(w, xw[PADL - H + v + k], s);
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        s = fsum<C, ACC>(wy[i][k], lds[buf][row[i] - H + k][lane * VX + v], s);
#pragma unroll
                    for (int k = 0; k < K; ++k) s = fsum<C, ACC>(wz[k], q[i][(j + k) % R][v], s);
                    out[v] = s;
                }
                if (st_ok[i])
                    store_sums<T, VX, ACC>(du + (int64_t)z * s2 + (int64_t)y[i] * s1 + xs, out, accumulate);
            }
            buf ^= 1;
            return true;
        };
        if (!ring_steps<0, R>(step)) break;
    }
}

// ---------------------------------------------------------------------------------------
// 3-D, edge-loaded form.  Segments do not overlap: a wave owns an aligned 64*VX-point x
// segment of VY consecutive rows and every lane computes.  The x window comes from the
// neighbouring lanes through DPP; the outermost lanes take the NE groups beyond the segment
// from LDS (lane 0 those on the left, lane 63 those on the right, as the DPP's fill value), where
// a few lanes of every wave stage them next to the segment each step.  The block's W*VY rows
// plus H halo rows above and below are staged in LDS (double-buffered, one barrier per
// step); y taps are LDS reads, z taps come from a register ring of K + PF planes per
// computed row.  Halo rows and edge groups are loaded one plane ahead into two-slot queues
// (the ring length K + PF is even, so the slot of a plane is its parity within the ring).
// Weights along y are wave-uniform (scalar registers), along z block-uniform (LDS), along x
// per column (registers, or LDS when VX * K values would not fit).  The per-point FMA
// sequence is the generic evaluator's: x taps, then y, then z, each in tap order.
template <typename C>
__device__ inline C uniform_val(C v);
template <>
__device__ inline float uniform_val<float>(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
template <>
__device__ inline double uniform_val<double>(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <typename T, int VX, int K, bool UNI, int PF, int W, int VY, int ACC>
__global__ __launch_bounds__(64 * W) void k_tiled3e(const FastArgs A, T* __restrict__ du,
                                                     const T* __restrict__ u, int accumulate)
{
    using C = typename Store<T>::C;
    constexpr int H = K / 2;
    constexpr int R = K + PF;
    static_assert(R % 2 == 0, "two-slot queues advance with an even ring");
    static_assert(2 * H <= W, "one halo row per wave at most");
    constexpr int TY = W * VY;                          // computed rows per block
    constexpr int SR = TY + 2 * H;                      // staged rows
    constexpr int NE = (H + VX - 1) / VX;               // edge groups on each side
    constexpr int ROWE = (64 + 2 * NE) * VX;            // staged row: edges + segment + edges
    constexpr int C0 = NE * VX;                         // column of the segment's first point
    constexpr bool XWL = !UNI && VX * K * sizeof(C) > 64;   // x weights in LDS
    static_assert(2 * NE * VY <= 64, "edge groups of a wave's rows fit its lanes");
    __shared__ __attribute__((aligned(16))) C lds[2][SR][ROWE];
    __shared__ C wzs[UNI ? 1 : kMaxChunk3D][K];
    __shared__ __attribute__((aligned(16))) C wxs[XWL ? K : 1][XWL ? 64 * VX : 1];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = tile_of((int)blockIdx.x, A.nblocks, A.xcd_remap);
    // logical tile: y block fastest (neighbouring y blocks, which read each other's rows as
    // halo, land on one XCD), then x segment, then z chunk
    const int yblk = b % A.nyblk;
    const int t = b / A.nyblk;
    const int xseg = t % A.nxseg;
    const int zch = t / A.nxseg;
    const int y0 = A.lo[1] + yblk * TY;
    const int z0 = A.lo[2] + zch * A.chunk;
    const int z1 = min(z0 + A.chunk, A.hi[2]);
    const int64_t n0 = A.n[0];
    const int n1 = (int)A.n[1], n2 = (int)A.n[2];
    const int64_t s1 = A.s1, s2 = A.s2;
    const int64_t xs0 = (int64_t)A.lo[0] + (int64_t)xseg * (64 * VX);
    const int64_t xs = xs0 + lane * VX;
    // Loads are clamped into the array: values of clamped lanes, rows or planes are only used
    // by outputs that are not stored (an output in the box [lo, hi) reads no further than H
    // points past the box, inside the array).
    const int64_t xl = min(xs, n0 - VX);
    const bool xst = xs < A.hi[0];
    // edge duty: lane e < 2*NE*VY stages group g of side sd of computed row ei
    const int ei = lane / (2 * NE), sd = (lane / NE) & 1, eg_g = lane % NE;
    const bool edge_lane = lane < 2 * NE * VY;
    int64_t xe = sd == 0 ? xs0 - (int64_t)(NE - eg_g) * VX : xs0 + (int64_t)(64 + eg_g) * VX;
    xe = edge_lane ? min(max(xe, (int64_t)0), n0 - VX) : xl;
    const int ecol = sd == 0 ? eg_g * VX : C0 + 64 * VX + eg_g * VX;
    // fill values of the DPP window: lane 0 reads groups -(g+1), lane 63 groups 64+g
    const int fcol0 = lane == 0 ? (NE - 1) * VX : C0 + 64 * VX;   // column of g = 0
    const int fstep = lane == 0 ? -VX : VX;

    int rr[VY], y[VY], yc[VY];
    bool st_ok[VY], live[VY];
#pragma unroll
    for (int i = 0; i < VY; ++i) {
        rr[i] = H + wave * VY + i;
        y[i] = y0 + wave * VY + i;
        yc[i] = min(y[i], n1 - 1);
        live[i] = y[i] < A.hi[1];                       // wave-uniform
        st_ok[i] = live[i] && xst;
    }
    const int jh = wave;                                 // halo row of this wave, if < 2H
    const bool has_halo = jh < 2 * H;
    const int hrow = jh < H ? jh : TY + jh;
    const int yh = min(max(y0 - H + hrow, 0), n1 - 1);
    const int erow = H + wave * VY + min(ei, VY - 1);    // staged row of the edge group
    const int ye = min(y0 + wave * VY + min(ei, VY - 1), n1 - 1);

    const C* wxt = wtab<C>(A, 0);
    const C* wyt = wtab<C>(A, 1);
    const C* wzt = wtab<C>(A, 2);
    C wx[XWL ? 1 : VX][K];
    C wy[VY][K];
    C wxu[K], wzu[K];
#pragma unroll
    for (int i = 0; i < VY; ++i) {
        const int64_t yr = UNI ? 0 : yc[i];
#pragma unroll
        for (int k = 0; k < K; ++k) wy[i][k] = uniform_val<C>(wyt[yr * K + k]);
    }
    if constexpr (UNI) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            wxu[k] = uniform_val<C>(wxt[k]);
            wzu[k] = uniform_val<C>(wzt[k]);
        }
    } else {
        for (int i = threadIdx.x; i < (z1 - z0) * K; i += 64 * W)
            wzs[i / K][i % K] = wzt[(int64_t)(z0 + i / K) * K + i % K];
        if constexpr (XWL) {
            for (int i = threadIdx.x; i < 64 * VX * K; i += 64 * W) {
                const int c = i / K, k = i % K;
                const int64_t x = min(xs0 + c, n0 - 1);
                wxs[k][c] = wxt[x * K + k];
            }
        } else {
#pragma unroll
            for (int v = 0; v < VX; ++v) {
                const int64_t xr = min(xs + v, n0 - 1);
#pragma unroll
                for (int k = 0; k < K; ++k) wx[v][k] = wxt[xr * K + k];
            }
        }
    }

    const T* urow[VY];
#pragma unroll
    for (int i = 0; i < VY; ++i) urow[i] = u + (int64_t)yc[i] * s1 + xl;
    const T* uhalo = u + (int64_t)yh * s1 + xl;
    const T* uedge = u + (int64_t)ye * s1 + xe;
    const int zlast = min(z1 + H - 1, n2 - 1);
    auto zc = [&](int zz) { return (int64_t)min(max(zz, 0), zlast) * s2; };

    // halo row and edge group: a ring of D planes (D divides R, so a slot is a compile-time
    // function of the step), loaded D - 1 planes ahead
    constexpr int D = (R % 4 == 0) ? 4 : 2;
    C q[VY][R][VX];
    C hq[D][VX];   // halo row, planes z .. z + D - 1
    C eq[D][VX];   // edge group, planes z .. z + D - 1
#pragma unroll
    for (int s = 0; s < R - 1; ++s) {
#pragma unroll
        for (int i = 0; i < VY; ++i) VecIO<T, VX>::ld(urow[i] + zc(z0 - H + s), q[i][s]);
    }
#pragma unroll
    for (int s = 0; s < D - 1; ++s) {
        VecIO<T, VX>::ld(uhalo + zc(z0 + s), hq[s]);
        VecIO<T, VX>::ld(uedge + zc(z0 + s), eq[s]);
    }

    int buf = 0;
    for (int zb = z0; zb < z1; zb += R) {
        auto step = [&](auto jc) -> bool {
            constexpr int j = decltype(jc)::value;
            constexpr int cur = j % D, nxt = (j + D - 1) % D;
            const int z = zb + j;
            if (z >= z1) return false;   // block-uniform
#pragma unroll
            for (int i = 0; i < VY; ++i)
                VecIO<T, VX>::ld(urow[i] + zc(z + H + PF), q[i][(j + K - 1 + PF) % R]);
            VecIO<T, VX>::ld(uhalo + zc(z + D - 1), hq[nxt]);
            VecIO<T, VX>::ld(uedge + zc(z + D - 1), eq[nxt]);
            // stage plane z: computed rows, the halo row, the edge groups
#pragma unroll
            for (int i = 0; i < VY; ++i) {
#pragma unroll
                for (int v = 0; v < VX; ++v) lds[buf][rr[i]][C0 + lane * VX + v] = q[i][(j + H) % R][v];
            }
            if (has_halo) {
#pragma unroll
                for (int v = 0; v < VX; ++v) lds[buf][hrow][C0 + lane * VX + v] = hq[cur][v];
            }
            if (edge_lane) {
#pragma unroll
                for (int v = 0; v < VX; ++v) lds[buf][erow][ecol + v] = eq[cur][v];
            }
            __syncthreads();
            C wz[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if constexpr (UNI) wz[k] = wzu[k];
                else wz[k] = wzs[z - z0][k];
            }
            C wxl[XWL ? K : 1][XWL ? VX : 1];
            if constexpr (XWL) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
#pragma unroll
                    for (int v = 0; v < VX; ++v) wxl[k][v] = wxs[k][lane * VX + v];
                }
            }
#pragma unroll
            for (int i = 0; i < VY; ++i) {
                if (!live[i]) continue;   // wave-uniform
                const C* c = q[i][(j + H) % R];
                C eg[NE][VX];
#pragma unroll
                for (int g = 0; g < NE; ++g) {
#pragma unroll
                    for (int v = 0; v < VX; ++v) eg[g][v] = C(0);
                }
                if (lane == 0 || lane == 63) {
#pragma unroll
                    for (int g = 0; g < NE; ++g) {
#pragma unroll
                        for (int v = 0; v < VX; ++v) eg[g][v] = lds[buf][rr[i]][fcol0 + g * fstep + v];
                    }
                }
                C xw[VX + 2 * H];
                x_window_edges<C, VX, H>(c, eg, xw);
                Sum<C, ACC> out[VX];
#pragma unroll
                for (int v = 0; v < VX; ++v) {
                    Sum<C, ACC> s = 0;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        C w;
                        if constexpr (UNI) w = wxu[k];
                        else if constexpr (XWL) w = wxl[k][v];
                        else w = wx[v][k];
                        s = fsum<C, ACC>(w, xw[v + k], s);
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const C val = k == H ? c[v] : lds[buf][rr[i] - H + k][C0 + lane * VX + v];
                        s = fsum<C, ACC>(wy[i][k], val, s);
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k) s = fsum<C, ACC>(wz[k], q[i][(j + k) % R][v], s);
                    out[v] = s;
                }
                if (st_ok[i])
                    store_sums<T, VX, ACC>(du + (int64_t)z * s2 + (int64_t)y[i] * s1 + xs, out, accumulate);
            }
            buf ^= 1;
            return true;
        };
        if (!ring_steps<0, R>(step)) break;
    }
}

// ---------------------------------------------------------------------------------------
template <typename T, int VX, int K, int ACC>
hipError_t launch_interior_tiles(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                                 int accumulate, hipStream_t s)
{
    const FastPlan& fp = p.fast;
    const bool uni = A.uniform[0] && A.uniform[1] && (fp.kind == 2 || A.uniform[2]);
    // per-row indexing needs a row of weights for every point of every axis
    for (int a = 0; a < fp.kind; ++a)
        if (fp.w64[a].size() != (size_t)(uni ? 1 : A.n[a]) * K) return hipErrorInvalidValue;
    const int64_t per = (int64_t)256 * kShellPtsPerThread;
    // 2-D: the (thin) shell rides in the tile launch; 3-D: a launch of its own first
    const int nshell = fp.kind == 2 ? (int)((fp.shell_total + per - 1) / per) : 0;
    if (fp.kind == 3 && A.gen == 2) {
        if constexpr (K <= 9 && VX <= 16 / (int)sizeof(T) && VX <= 4) {
#define DOP_L3E(UNI, W, VY)                                                                  \
    hipLaunchKernelGGL((k_tiled3e<T, VX, K, UNI, 1, W, VY, ACC>), dim3((unsigned)A.nblocks),   \
                       dim3(64 * W), 0, s, A, (T*)du, (const T*)u, accumulate)
#define DOP_L3EU(W, VY)                                                                      \
    do {                                                                                     \
        if (uni) DOP_L3E(true, W, VY); else DOP_L3E(false, W, VY);                           \
    } while (0)
            const int shape = A.W * 8 + A.VY;
            if (shape == 8 * 8 + 1) DOP_L3EU(8, 1);
            else if (shape == 8 * 8 + 2) DOP_L3EU(8, 2);
            else if (shape == 16 * 8 + 1) DOP_L3EU(16, 1);
            else if (shape == 8 * 8 + 3) DOP_L3EU(8, 3);
            else return hipErrorInvalidValue;
#undef DOP_L3EU
#undef DOP_L3E
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
    const dim3 grid((unsigned)(A.nblocks + nshell));
    if (fp.kind == 2) {
#define DOP_L2E(UNI, PF, EDG)                                                                \
    hipLaunchKernelGGL((k_tiled2d<T, VX, K, UNI, PF, ACC, EDG>), grid, dim3(256), 0, s, A, p.dev, \
                       fp.shell, nshell, (T*)du, (const T*)u, accumulate)
#define DOP_L2(UNI)                                                                          \
    do {                                                                                     \
        if (A.gen == 2) {                                                                    \
            if constexpr (K / 2 <= VX) {                                                     \
                if constexpr (sizeof(T) == 8) {                                              \
                    if (A.PF == 16) DOP_L2E(UNI, 16, true); else DOP_L2E(UNI, 8, true);      \
                } else {                                                                     \
                    if (A.PF == 8) DOP_L2E(UNI, 8, true); else DOP_L2E(UNI, 4, true);        \
                }                                                                            \
                break;                                                                       \
            }                                                                                \
            return hipErrorInvalidValue;                                                     \
        }                                                                                    \
        if (A.PF == 4) DOP_L2E(UNI, 4, false); else DOP_L2E(UNI, 2, false);                  \
    } while (0)
        if (uni) DOP_L2(true); else DOP_L2(false);
#undef DOP_L2
#undef DOP_L2E
    } else {
        if constexpr (K <= 9 && VX <= 16 / (int)sizeof(T) && VX <= 4) {
#define DOP_L3(UNI, W, VY)                                                                   \
    hipLaunchKernelGGL((k_tiled3d<T, VX, K, UNI, kPF3D<T, K>, W, VY, ACC>), grid, dim3(64 * W), 0, s, \
                       A, p.dev, fp.shell, nshell, (T*)du, (const T*)u, accumulate)
            constexpr bool w8 = 2 * (K / 2) < 8;   // 8-wave blocks of one row compute 8 - 2H rows
#define DOP_L3U(W, VY)                                                                       \
    do {                                                                                     \
        if (uni) DOP_L3(true, W, VY); else DOP_L3(false, W, VY);                             \
    } while (0)
            // (waves per block, rows per wave): the shapes fast_variant produces
            const int shape = A.W * 8 + A.VY;
            if (shape == 8 * 8 + 1) {
                if constexpr (w8) DOP_L3U(8, 1);
                else return hipErrorInvalidValue;
            } else if (shape == 16 * 8 + 1) {
                DOP_L3U(16, 1);
            } else if (shape == 8 * 8 + 2) {
                DOP_L3U(8, 2);
            } else if (shape == 16 * 8 + 2) {
                DOP_L3U(16, 2);
            } else if (shape == 8 * 8 + 3) {
                DOP_L3U(8, 3);
            } else {
                return hipErrorInvalidValue;
            }
#undef DOP_L3U
#undef DOP_L3
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

// The 3-D shell (its own launch, on s2; may be s)
template <typename T, int K, int ACC>
hipError_t launch_shell3d(const DevicePlan& p, void* du, const void* u, int accumulate, hipStream_t s2)
{
    const FastPlan& fp = p.fast;
    if constexpr (K <= 9) {
        if (fp.kind == 3 && fp.strips.nblocks > 0) {
            const dim3 sg((unsigned)fp.strips.nblocks);
            const StripArgs& SA = fp.strips;
            if (SA.vec_ok)
                hipLaunchKernelGGL((shell::k_shell_strips<T, ACC, K, true>), sg,
                                   dim3(shell::kShellBlock), 0, s2, p.dev, SA, (T*)du, (const T*)u,
                                   accumulate);
            else
                hipLaunchKernelGGL((shell::k_shell_strips<T, ACC, K, false>), sg,
                                   dim3(shell::kShellBlock), 0, s2, p.dev, SA, (T*)du, (const T*)u,
                                   accumulate);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// The tiles on s, then (3-D) the shell on s2: the long bandwidth-bound launch is dispatched first
template <typename T, int VX, int K, int ACC>
hipError_t launch_interior(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                           int accumulate, hipStream_t s, hipStream_t s2)
{
    const hipError_t e = launch_interior_tiles<T, VX, K, ACC>(p, A, du, u, accumulate, s);
    if (e != hipSuccess || p.fast.kind != 3) return e;
    return launch_shell3d<T, K, ACC>(p, du, u, accumulate, s2);
}

template <typename T, int VX, int ACC>
hipError_t launch_interior_acc(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                               int accumulate, hipStream_t s, hipStream_t s2)
{
    switch (A.K) {
    case 3: return launch_interior<T, VX, 3, ACC>(p, A, du, u, accumulate, s, s2);
    case 5: return launch_interior<T, VX, 5, ACC>(p, A, du, u, accumulate, s, s2);
    case 7: return launch_interior<T, VX, 7, ACC>(p, A, du, u, accumulate, s, s2);
    case 9: return launch_interior<T, VX, 9, ACC>(p, A, du, u, accumulate, s, s2);
    case 11: return launch_interior<T, VX, 11, ACC>(p, A, du, u, accumulate, s, s2);
    case 13: return launch_interior<T, VX, 13, ACC>(p, A, du, u, accumulate, s, s2);
    }
    return hipErrorInvalidValue;
}

// the plan's accumulation: the storage's own (f64 / f32) or, for f32 and bf16 storage,
// f32 values summed in f64 (DOP_ACC_F32_WIDE, what DOP_ACC_AUTO picks for derivative orders
// above 2)
template <typename T, int VX>
hipError_t launch_interior_k(const DevicePlan& p, const FastArgs& A, void* du, const void* u,
                             int accumulate, hipStream_t s, hipStream_t s2)
{
    if constexpr (!std::is_same<T, double>::value) {
        if (p.host.accum == kAccF32Wide)
            return launch_interior_acc<T, VX, kAccF32Wide>(p, A, du, u, accumulate, s, s2);
    }
    return launch_interior_acc<T, VX, acc_of<T>()>(p, A, du, u, accumulate, s, s2);
}

}  // namespace tiled
}  // namespace dop
