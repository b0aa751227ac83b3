// dop_shell.hpp -- the boundary shell of a 3-D tiled plan (every point outside the tiles'
// interior box), evaluated in strips of up to four consecutive x points of one row.
//
// The arithmetic of every point is the generic evaluator's (x term, then y, then z; regular
// rows as FMA chains in tap order, boundary rows in FP64 with their ghosts formed first), so the
// shell agrees with k_apply_generic bit for bit.  What changes is how the values arrive:
//  - y and z terms: the strip's points share their row along y and z, so every tap is one
//    load of the strip's four values (one 16-byte vector load for an aligned f32 strip) and the
//    row's weights, ghost taps and boundary program are the same for the four points;
//  - x term: on the faces that hold the x boundary columns, each thread first stages the end
//    of its line (kStage values) in LDS with vector loads, and the boundary rows read from
//    there; elsewhere the x taps are loads of neighbouring points, which the row's other
//    strips share in L1.
#pragma once
#include <hip/hip_runtime.h>

#include "dop_eval.hpp"
#include "dop_fast.hpp"

namespace dop {
namespace shell {

constexpr int kSW = kStripPts;  // points per strip
constexpr int kStage = 16;      // staged values at an x end of a line
constexpr int kShellBlock = 256;

// w (<= 4) consecutive values at p, as C; one vector load when `vec` (16-byte aligned f32,
// two 16-byte loads f64, one 8-byte load bf16)
template <typename T>
__device__ inline void ld_strip(const T* p, int w, bool vec, typename Store<T>::C* o);

template <>
__device__ inline void ld_strip<float>(const float* p, int w, bool vec, float* o)
{
    if (vec && w == 4) {
        const float4 v = *reinterpret_cast<const float4*>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
        return;
    }
#pragma unroll
    for (int i = 0; i < kSW; ++i) o[i] = i < w ? p[i] : 0.0f;
}

template <>
__device__ inline void ld_strip<double>(const double* p, int w, bool vec, double* o)
{
    if (vec && w == 4) {
        const double2 a = *reinterpret_cast<const double2*>(p);
        const double2 b = *reinterpret_cast<const double2*>(p + 2);
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
        return;
    }
#pragma unroll
    for (int i = 0; i < kSW; ++i) o[i] = i < w ? p[i] : 0.0;
}

template <>
__device__ inline void ld_strip<bf16>(const bf16* p, int w, bool vec, float* o)
{
    if (vec && w == 4) {
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

// The term of axis X (y or z: the strip's points share row r) added to acc[0..w): point i's
// line starts at base0 + i.  Same arithmetic per point as axis_term.
template <typename T, int ACC>
__device__ inline void strip_axis_term(const PlanDev& P, const AxisDev& X, const T* src, int r,
                                       int64_t base0, int64_t S, int w, bool vec, const int64_t* idx,
                                       int a, Acc<ACC>* acc)
{
    using C = typename Store<T>::C;
    using W = typename Acc<ACC>::W;
    if (r >= X.rlo && r < X.rhi) {
        const W* wt = Acc<ACC>::table(X) + (X.uniform ? 0 : r);
        const int64_t ws = X.uniform ? 1 : X.n;
        const int64_t p0 = base0 + (int64_t)(r + X.dmin + (P.ghost ? 0 : 1)) * S;
        if (X.K <= kEvalUnroll) {
            W wv[kEvalUnroll];
            C xv[kEvalUnroll][kSW];
#pragma unroll
            for (int k = 0; k < kEvalUnroll; ++k) {
                if (k < X.K) {
                    wv[k] = wt[k * ws];
                    ld_strip<T>(src + p0 + k * S, w, vec, xv[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < kEvalUnroll; ++k) {
                if (k < X.K) {
#pragma unroll
                    for (int i = 0; i < kSW; ++i)
                        if (i < w) acc[i].add(wv[k], (W)xv[k][i]);
                }
            }
        } else {
            for (int k = 0; k < X.K; ++k) {
                C xv[kSW];
                ld_strip<T>(src + p0 + k * S, w, vec, xv);
                const W wk = wt[k * ws];
#pragma unroll
                for (int i = 0; i < kSW; ++i)
                    if (i < w) acc[i].add(wk, (W)xv[i]);
            }
        }
        return;
    }
    if (X.lbc) {
        // one BC per line: each point's own ghosts
        for (int i = 0; i < w; ++i) {
            int64_t id[3] = {idx[0] + i, idx[1], idx[2]};
            acc[i].add_sum(boundary_row<T>(P, X, src, r, base0 + i, S, line_of(X, id)));
        }
        return;
    }
    // boundary row: the row program, ghost taps and constants are the strip's; the values are
    // the strip's four lines, each summed in the generic order
    const int b = r < X.rlo ? r : X.rlo + (r - X.rhi);
    const int J0 = X.bJ0[b];
    const int nt = X.bn[b];
    const double* wr = X.bw + (int64_t)b * X.KB;
    const int joff = P.ghost ? -1 : 0;
    const bool gl_used = P.ghost && J0 == 0;
    const bool gr_used = P.ghost && J0 + nt - 1 == X.n + 1;
    double gl[kSW], gr[kSW];
    if (gl_used) {
        if (X.periodic) {
            C v[kSW];
            ld_strip<T>(src + base0 + (X.periodic == 1 ? (int64_t)(X.n - 1) * S : 0), w, vec, v);
#pragma unroll
            for (int i = 0; i < kSW; ++i) gl[i] = (double)v[i];
        } else {
#pragma unroll
            for (int i = 0; i < kSW; ++i) gl[i] = 0.0;
            for (int k = 0; k < X.nl; ++k) {
                C v[kSW];
                ld_strip<T>(src + base0 + (int64_t)k * S, w, vec, v);
#pragma unroll
                for (int i = 0; i < kSW; ++i) gl[i] = fma(X.al[k], (double)v[i], gl[i]);
            }
#pragma unroll
            for (int i = 0; i < kSW; ++i) gl[i] = gl[i] + X.bl;
        }
    }
    if (gr_used) {
        if (X.periodic) {
            C v[kSW];
            ld_strip<T>(src + base0 + (X.periodic == 2 ? (int64_t)(X.n - 1) * S : 0), w, vec, v);
#pragma unroll
            for (int i = 0; i < kSW; ++i) gr[i] = (double)v[i];
        } else {
#pragma unroll
            for (int i = 0; i < kSW; ++i) gr[i] = 0.0;
            const int64_t o = (int64_t)X.n - X.nr;
            for (int k = 0; k < X.nr; ++k) {
                C v[kSW];
                ld_strip<T>(src + base0 + (o + k) * S, w, vec, v);
#pragma unroll
                for (int i = 0; i < kSW; ++i) gr[i] = fma(X.ar[k], (double)v[i], gr[i]);
            }
#pragma unroll
            for (int i = 0; i < kSW; ++i) gr[i] = gr[i] + X.br;
        }
    }
    double s[kSW];
#pragma unroll
    for (int i = 0; i < kSW; ++i) s[i] = 0.0;
    for (int k = 0; k < nt; ++k) {
        const int J = J0 + k;
        double v[kSW];
        if (gl_used && J == 0) {
#pragma unroll
            for (int i = 0; i < kSW; ++i) v[i] = gl[i];
        } else if (gr_used && J == X.n + 1) {
#pragma unroll
            for (int i = 0; i < kSW; ++i) v[i] = gr[i];
        } else {
            C c[kSW];
            ld_strip<T>(src + base0 + (int64_t)(J + joff) * S, w, vec, c);
#pragma unroll
            for (int i = 0; i < kSW; ++i) v[i] = (double)c[i];
        }
#pragma unroll
        for (int i = 0; i < kSW; ++i) s[i] = fma(wr[k], v[i], s[i]);
    }
#pragma unroll
    for (int i = 0; i < kSW; ++i)
        if (i < w) acc[i].add_sum(s[i]);
}

// The shell of a 3-D plan, one strip per thread.  Strips of one row are consecutive threads, so
// the faces that span whole rows load coalesced; on the x boundary faces consecutive threads
// are consecutive lines.
template <typename T, int ACC>
__global__ __launch_bounds__(kShellBlock) void k_shell_strips(const PlanDev* __restrict__ Pp,
                                                             const StripArgs SA,
                                                             T* __restrict__ du,
                                                             const T* __restrict__ src,
                                                             int accumulate)
{
    __shared__ T stage[kShellBlock][kStage + 1];
    const PlanDev& P = *Pp;
    const int64_t item = (int64_t)blockIdx.x * kShellBlock + threadIdx.x;
    if (item >= SA.start[SA.nbox]) return;
    int b = 0;
    while (b + 1 < SA.nbox && item >= SA.start[b + 1]) ++b;
    int64_t l = item - SA.start[b];
    const int64_t sx = l % SA.nsx[b];
    l /= SA.nsx[b];
    int64_t idx[3];
    idx[0] = SA.org[b][0] + sx * kSW;
    idx[1] = SA.org[b][1] + l % SA.ext[b][1];
    idx[2] = SA.org[b][2] + l / SA.ext[b][1];
    const int w = (int)min((int64_t)kSW, SA.org[b][0] + SA.ext[b][0] - idx[0]);
    const bool vec = SA.vec_ok && (idx[0] % kSW) == 0;
    Acc<ACC> acc[kSW];

    // x term
    const AxisDev& X0 = P.ax[0];
    if (X0.active) {
        int64_t base = P.tshift[0];
        for (int a = 1; a < P.ndim; ++a) base += idx[a] * P.ssrc[a];
        // on an x boundary face, the end of the line this strip lies in is staged in LDS
        const bool staged = SA.xcol[b] && P.ghost && X0.n >= kStage && (X0.n % kSW) == 0;
        int64_t sx0 = 0;
        if (staged) {
            sx0 = idx[0] < X0.n / 2 ? 0 : X0.n - kStage;
            T* st = stage[threadIdx.x];
#pragma unroll
            for (int g = 0; g < kStage / kSW; ++g) {
                typename Store<T>::C v[kSW];
                ld_strip<T>(src + base + sx0 + g * kSW, kSW, SA.vec_ok != 0, v);
#pragma unroll
                for (int i = 0; i < kSW; ++i) Store<T>::st(st, g * kSW + i, (double)v[i]);
            }
        }
        const T* st = stage[threadIdx.x];
        auto ldx = [&](int64_t j) -> double {
            if (staged && j >= sx0 && j < sx0 + kStage) return (double)Store<T>::ld(st, j - sx0);
            return (double)Store<T>::ld(src, base + j);
        };
        for (int i = 0; i < w; ++i) {
            const int r = (int)(idx[0] + i);
            if (r >= X0.rlo && r < X0.rhi) {
                using W = typename Acc<ACC>::W;
                const W* wt = Acc<ACC>::table(X0) + (X0.uniform ? 0 : r);
                const int64_t ws = X0.uniform ? 1 : X0.n;
                const int64_t j0 = (int64_t)(r + X0.dmin + (P.ghost ? 0 : 1));
                for (int k = 0; k < X0.K; ++k) acc[i].add(wt[k * ws], (W)ldx(j0 + k));
            } else {
                int64_t id[3] = {idx[0] + i, idx[1], idx[2]};
                acc[i].add_sum(boundary_row_f(P, X0, r, line_of(X0, id), ldx));
            }
        }
    }
    // y and z terms
    for (int a = 1; a < P.ndim; ++a) {
        const AxisDev& X = P.ax[a];
        if (!X.active) continue;
        int64_t base0 = P.tshift[a] + idx[0] * P.ssrc[0];
        for (int c = 1; c < P.ndim; ++c)
            if (c != a) base0 += idx[c] * P.ssrc[c];
        strip_axis_term<T, ACC>(P, X, src, (int)idx[a], base0, P.ssrc[a], w, vec, idx, a, acc);
    }
    const int64_t lin = idx[0] + idx[1] * P.sdu[1] + idx[2] * P.sdu[2];
    for (int i = 0; i < w; ++i) {
        double v = acc[i].value();
        if (accumulate) v += (double)Store<T>::ld(du, lin + i);
        Store<T>::st(du, lin + i, v);
    }
}

}  // namespace shell
}  // namespace dop
