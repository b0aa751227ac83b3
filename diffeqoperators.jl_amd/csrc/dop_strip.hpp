// dop_strip.hpp -- the shell of a 3-D tiled plan (every point outside the tiles' interior box)
// in work items of up to four consecutive x points of one row: the per-item evaluation, written
// for host and device so that the CPU build of the plan path (tests) runs the same code as
// k_shell_strips (dop_shell.hpp).
//
// The arithmetic of every point is the generic evaluator's (x term, then y, then z; regular
// rows as FMA chains in tap order, boundary rows in FP64 with their ghosts formed first), so the
// shell agrees with k_apply_generic bit for bit.  What changes is how the values arrive:
//  - x term: an item first stages 16 consecutive values of its line (4 aligned vector loads)
//    -- the window of four regular rows, or the end of the line for an item that touches x
//    boundary rows -- and reads its taps, ghosts and boundary rows from there;
//  - y and z terms: the item's points share their row along y and z, so every tap is one
//    load of the item's four values (one 16-byte vector load for an aligned f32 item) and the
//    row's weights, ghost taps and boundary program are the same for the four points.
// Face items (eval_face) have the same shell row for the whole block and regular rows on the
// other two axes, so the boundary program is uniform and every load of a term is issued before
// its first FMA; general items (eval_strip) cover the rest (edges, ends of lines, per-line BCs).
// Regular windows are unrolled to the tiles' compile-time width KM; the rarer paths (per-point
// x boundary rows, per-line BCs, ghosts) are loops over the points, so the kernel stays small.
#pragma once
#include <cstdint>

#include "dop_eval.hpp"
#include "dop_fast.hpp"

namespace dop {
namespace strip {

constexpr int kSW = kStripPts;        // points per strip
constexpr int kStage = kStripStage;   // staged values of a strip's line

// w (<= kSW) consecutive values at p, one load each (the device kernel uses vector loads);
// ldw: the same for weights; st: w values stored
template <typename T>
struct ScalarLoad {
    using C = typename Store<T>::C;
    DOP_HD static void ld(const T* p, int w, bool, C* o)
    {
#pragma unroll
        for (int i = 0; i < kSW; ++i) o[i] = i < w ? (C)Store<T>::ld(p, i) : C(0);
    }
    template <typename W>
    DOP_HD static void ldw(const W* p, int w, bool, W* o)
    {
#pragma unroll
        for (int i = 0; i < kSW; ++i) o[i] = i < w ? p[i] : W(0);
    }
    DOP_HD static void st(T* p, int w, bool, const double* v)
    {
#pragma unroll
        for (int i = 0; i < kSW; ++i)
            if (i < w) Store<T>::st(p, i, v[i]);
    }
};

// An accumulator parked in memory (LDS on the device): the loops that take the points of a
// strip one at a time (x boundary rows, per-line BCs) index points at run time, which the
// registers holding acc[] cannot be.  The shell runs the f64, f32 and f32-values-in-f64 sums,
// whose state is one value.
template <int ACC>
using Slot = decltype(Acc<ACC>::s);

template <int ACC>
DOP_HD inline void slot_put(Slot<ACC>* sl, int i, const Acc<ACC>& a)
{
    static_assert(ACC == kAccF64 || ACC == kAccF32 || ACC == kAccF32Wide, "one-value accumulators only");
    sl[i] = a.s;
}

template <int ACC>
DOP_HD inline Acc<ACC> slot_get(const Slot<ACC>* sl, int i)
{
    Acc<ACC> a;
    a.s = sl[i];
    return a;
}

// Ghost values of the strip's four lines along X (left or right end), each a . u + b in FP64
// as ghost_left_f / ghost_right_f form it (the axis has one BC for all lines here).
template <typename T, typename LD>
DOP_HD inline void strip_ghost(const AxisDev& X, const T* src, int64_t base0, int64_t S, int w,
                               bool vec, bool left, double* g)
{
    using C = typename Store<T>::C;
    if (X.periodic) {
        const bool last = left ? X.periodic == 1 : X.periodic == 2;
        C v[kSW];
        LD::ld(src + base0 + (last ? (int64_t)(X.n - 1) * S : 0), w, vec, v);
#pragma unroll
        for (int i = 0; i < kSW; ++i) g[i] = (double)v[i];
        return;
    }
    const int nb = left ? X.nl : X.nr;
    const double* a = left ? X.al : X.ar;
    const int64_t o = left ? 0 : (int64_t)X.n - nb;
#pragma unroll
    for (int i = 0; i < kSW; ++i) g[i] = 0.0;
    for (int k = 0; k < nb; ++k) {
        C v[kSW];
        LD::ld(src + base0 + (o + k) * S, w, vec, v);
#pragma unroll
        for (int i = 0; i < kSW; ++i) g[i] = fma(tab(a, k), (double)v[i], g[i]);
    }
    const double c = left ? X.bl : X.br;
#pragma unroll
    for (int i = 0; i < kSW; ++i) g[i] = g[i] + c;
}

// A regular row r of axis X (y or z: the item's points share the row) added to acc[0..w):
// K <= KM taps, all loads before the first FMA.  Point i's line starts at base0 + i.
template <typename T, int ACC, int KM, typename LD>
DOP_HD inline void strip_regular(const AxisDev& X, const T* src, int r, int64_t base0, int64_t S,
                                 int w, bool vec, Acc<ACC>* acc)
{
    using C = typename Store<T>::C;
    using W = typename Acc<ACC>::W;
    const W* wt = Acc<ACC>::table(X) + (X.uniform ? 0 : r);
    const int64_t ws = X.uniform ? 1 : X.n;
    const int64_t p0 = base0 + (int64_t)(r + X.dmin) * S;
    W wv[KM];
    C xv[KM][kSW];
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        if (k < X.K) {
            wv[k] = tab(wt, k * ws);
            LD::ld(src + p0 + k * S, w, vec, xv[k]);
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
}

// A boundary row r of axis X (one BC for all lines): the row program, ghost taps and constants
// are the item's; the values are its four lines, each summed in the generic order.
template <typename T, int ACC, typename LD>
DOP_HD inline void strip_boundary(const AxisDev& X, const T* src, int r, int64_t base0, int64_t S,
                                  int w, bool vec, Acc<ACC>* acc)
{
    using C = typename Store<T>::C;
    const int b = r < X.rlo ? r : X.rlo + (r - X.rhi);
    const int J0 = tab(X.bJ0, b);
    const int nt = tab(X.bn, b);
    const double* wr = X.bw + (int64_t)b * X.KB;
    const bool gl_used = J0 == 0;
    const bool gr_used = J0 + nt - 1 == X.n + 1;
    double gl[kSW], gr[kSW];
#pragma unroll
    for (int i = 0; i < kSW; ++i) gl[i] = gr[i] = 0.0;
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
        if (!(side == 0 ? gl_used : gr_used)) continue;
        double g[kSW];
        strip_ghost<T, LD>(X, src, base0, S, w, vec, side == 0, g);
#pragma unroll
        for (int i = 0; i < kSW; ++i) {
            if (side == 0) gl[i] = g[i];
            else gr[i] = g[i];
        }
    }
    double s[kSW];
#pragma unroll
    for (int i = 0; i < kSW; ++i) s[i] = 0.0;
    if (nt <= kEvalUnroll) {
        C c[kEvalUnroll][kSW];
        double wk[kEvalUnroll];
#pragma unroll
        for (int k = 0; k < kEvalUnroll; ++k) {
            const int J = J0 + k;
            if (k < nt) {
                wk[k] = tab(wr, k);
                if (!(gl_used && J == 0) && !(gr_used && J == X.n + 1))
                    LD::ld(src + base0 + (int64_t)(J - 1) * S, w, vec, c[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < kEvalUnroll; ++k) {
            if (k < nt) {
                const int J = J0 + k;
#pragma unroll
                for (int i = 0; i < kSW; ++i) {
                    const double v = (gl_used && J == 0) ? gl[i]
                                     : (gr_used && J == X.n + 1) ? gr[i] : (double)c[k][i];
                    s[i] = fma(wk[k], v, s[i]);
                }
            }
        }
    } else {
        for (int k = 0; k < nt; ++k) {
            const int J = J0 + k;
            C c[kSW];
            if (!(gl_used && J == 0) && !(gr_used && J == X.n + 1))
                LD::ld(src + base0 + (int64_t)(J - 1) * S, w, vec, c);
#pragma unroll
            for (int i = 0; i < kSW; ++i) {
                const double v = (gl_used && J == 0) ? gl[i]
                                 : (gr_used && J == X.n + 1) ? gr[i] : (double)c[i];
                s[i] = fma(tab(wr, k), v, s[i]);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kSW; ++i)
        if (i < w) acc[i].add_sum(s[i]);
}

// The term of axis X (y or z) at row r for the item's points: axis_term's arithmetic per point.
template <typename T, int ACC, int KM, typename LD>
DOP_HD inline void strip_axis_term(const PlanDev& P, const AxisDev& X, const T* src, int r,
                                   int64_t base0, int64_t S, int w, bool vec, const int64_t* idx,
                                   Acc<ACC>* acc, Slot<ACC>* slots)
{
    if (r >= X.rlo && r < X.rhi) {
        strip_regular<T, ACC, KM, LD>(X, src, r, base0, S, w, vec, acc);
        return;
    }
    if (X.lbc) {
        // one BC per line: each point's own ghosts
#pragma unroll
        for (int j = 0; j < kSW; ++j) slot_put(slots, j, acc[j]);
#pragma unroll 1
        for (int i = 0; i < w; ++i) {
            int64_t id[3] = {idx[0] + i, idx[1], idx[2]};
            Acc<ACC> a = slot_get<ACC>(slots, i);
            a.add_sum(boundary_row<T>(P, X, src, r, base0 + i, S, line_of(X, id)));
            slot_put(slots, i, a);
        }
#pragma unroll
        for (int j = 0; j < kSW; ++j) acc[j] = slot_get<ACC>(slots, j);
        return;
    }
    strip_boundary<T, ACC, LD>(X, src, r, base0, S, w, vec, acc);
}

// The x term of one point (row r of the x axis, line `line`), its line read through ld(j):
// a regular row's taps (all loads before the first FMA), or a boundary row's FP64 value --
// axis_term's arithmetic.
template <typename T, int ACC, int KM, typename F>
DOP_HD inline void x_term_point(const PlanDev& P, const AxisDev& X, int r, int64_t line,
                                const F& ld, Acc<ACC>& acc)
{
    using W = typename Acc<ACC>::W;
    if (r >= X.rlo && r < X.rhi) {
        const W* wt = Acc<ACC>::table(X) + (X.uniform ? 0 : r);
        const int64_t ws = X.uniform ? 1 : X.n;
        const int64_t j0 = (int64_t)(r + X.dmin);
        W wv[KM], xv[KM];
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            if (k < X.K) {
                wv[k] = tab(wt, k * ws);
                xv[k] = (W)ld(j0 + k);
            }
        }
#pragma unroll
        for (int k = 0; k < KM; ++k)
            if (k < X.K) acc.add(wv[k], xv[k]);
    } else {
        acc.add_sum(boundary_row_f(P, X, r, line, ld));
    }
}

// The term of axis X at a row r that is the same for the whole block (the shell row of a face):
// regular or boundary (one BC for all lines), with vector loads.
template <typename T, int ACC, int KM, typename LD>
DOP_HD inline void strip_boundary_or_regular(const AxisDev& X, const T* src, int r, int64_t base0,
                                             int64_t S, int w, Acc<ACC>* acc)
{
    if (r >= X.rlo && r < X.rhi) strip_regular<T, ACC, KM, LD>(X, src, r, base0, S, w, true, acc);
    else strip_boundary<T, ACC, LD>(X, src, r, base0, S, w, true, acc);
}

// The x line's 16 values from g0 (g0 + 16 <= n0, a multiple of 4 when vec) into stage
template <typename T, typename LD>
DOP_HD inline void x_stage(const T* src, int64_t base, int64_t g0, bool vec, T* stage)
{
    using C = typename Store<T>::C;
    C v[kStage / kSW][kSW];
#pragma unroll
    for (int g = 0; g < kStage / kSW; ++g) LD::ld(src + base + g0 + g * kSW, kSW, vec, v[g]);
#pragma unroll
    for (int g = 0; g < kStage / kSW; ++g) {
#pragma unroll
        for (int i = 0; i < kSW; ++i) Store<T>::st(stage, g * kSW + i, (double)v[g][i]);
    }
}

// The x term of w points x0 .. x0 + w - 1 that are all regular rows (the first x term, so
// acc[] is zero before): the window (K + w - 1 values) staged from the aligned group at or below
// x0 + dmin, the weights loaded with it, then each point's taps in order.  n0 >= 16.
template <typename T, int ACC, int KM, typename LD>
DOP_HD inline void x_regular(const AxisDev& X0, const T* src, int64_t base, int x0, int w, bool vec,
                             T* stage, Acc<ACC>* acc)
{
    using W = typename Acc<ACC>::W;
    const int64_t n0 = X0.n;
    const int64_t j = x0 + X0.dmin;
    int64_t g0 = j - (j & (kSW - 1));
    if (g0 > n0 - kStage) g0 = n0 - kStage;
    W wv[KM][kSW];
    const W* wt = Acc<ACC>::table(X0);
    if (X0.uniform) {
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            if (k < X0.K) {
                const W c = tab(wt, k);
#pragma unroll
                for (int i = 0; i < kSW; ++i) wv[k][i] = c;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < KM; ++k)
            if (k < X0.K) LD::ldw(wt + k * n0 + x0, w, vec, wv[k]);
    }
    x_stage<T, LD>(src, base, g0, vec, stage);
    const int64_t o = j - g0;
    W xv[KM + kSW - 1];
#pragma unroll
    for (int m = 0; m < KM + kSW - 1; ++m)
        if (m < X0.K + w - 1) xv[m] = (W)Store<T>::ld(stage, o + m);
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        if (k < X0.K) {
#pragma unroll
            for (int i = 0; i < kSW; ++i)
                if (i < w) acc[i].add(wv[k][i], xv[i + k]);
        }
    }
}

// The x term of w points one at a time (x boundary rows among them), each line read through
// ld(j); the first x term, so each point starts from zero.
template <typename T, int ACC, int KM, typename F>
DOP_HD inline void x_points(const PlanDev& P, const AxisDev& X0, const int64_t* idx, int w,
                            const F& ld, Acc<ACC>* acc, Slot<ACC>* slots)
{
#pragma unroll 1
    for (int i = 0; i < w; ++i) {
        int64_t id[3] = {idx[0] + i, idx[1], idx[2]};
        Acc<ACC> a;
        x_term_point<T, ACC, KM>(P, X0, (int)idx[0] + i, line_of(X0, id), ld, a);
        slot_put(slots, i, a);
    }
#pragma unroll
    for (int j = 0; j < kSW; ++j)
        if (j < w) acc[j] = slot_get<ACC>(slots, j);
}

// du at the item's w points (+ du when accumulating)
template <typename T, int ACC, typename LD>
DOP_HD inline void store_item(const PlanDev& P, const int64_t* idx, int w, bool vec, const Acc<ACC>* acc,
                              T* __restrict__ du, int accumulate)
{
    using C = typename Store<T>::C;
    const int64_t lin = idx[0] + idx[1] * P.sdu[1] + idx[2] * P.sdu[2];
    double out[kSW];
#pragma unroll
    for (int i = 0; i < kSW; ++i) out[i] = acc[i].value();
    if (accumulate) {
        C d[kSW];
        LD::ld(du + lin, w, vec, d);
#pragma unroll
        for (int i = 0; i < kSW; ++i) out[i] += (double)d[i];
    }
    LD::st(du + lin, w, vec, out);
}

// One general item (strip l of box b of SA): its points' values written to du.  `stage` holds
// kStage values and `slots` kSW accumulators of this thread (LDS on the device); LD loads up
// to kSW consecutive values.  KM >= the window of every regular row (the tiles' window).  VEC:
// SA.vec_ok (every item starts a 16-byte group of u and du that lies in the array, so whole
// groups are loaded even for items of fewer than 4 points).  Plans with ghost-point input only
// (the tiled path's condition).
template <typename T, int ACC, int KM, bool VEC, typename LD>
DOP_HD inline void eval_strip(const PlanDev& P, const StripArgs& SA, int b, int l, T* stage,
                              Slot<ACC>* slots, T* __restrict__ du, const T* __restrict__ src,
                              int accumulate)
{
    // strip l of box b: x fastest, then the box's thin axis
    const int nsx = SA.nsx[b];
    const int sx = l % nsx;
    l /= nsx;
    int64_t idx[3];
    idx[0] = SA.org[b][0] + (int64_t)sx * kSW;
    if (SA.zfirst[b]) {
        const int e2 = SA.ext[b][2];
        idx[2] = SA.org[b][2] + l % e2;
        idx[1] = SA.org[b][1] + l / e2;
    } else {
        const int e1 = SA.ext[b][1];
        idx[1] = SA.org[b][1] + l % e1;
        idx[2] = SA.org[b][2] + l / e1;
    }
    const int64_t wl = SA.org[b][0] + SA.ext[b][0] - idx[0];
    const int w = wl < kSW ? (int)wl : kSW;
    constexpr bool vec = VEC;
    Acc<ACC> acc[kSW];

    // x term
    const AxisDev& X0 = P.ax[0];
    if (X0.active) {
        int64_t base = P.tshift[0];
        for (int a = 1; a < P.ndim; ++a) base += idx[a] * P.ssrc[a];
        const int x0 = (int)idx[0];
        const int64_t n0 = X0.n;
        const bool regular = x0 >= X0.rlo && x0 + w <= X0.rhi;   // no x boundary row
        if (regular && n0 >= kStage) {
            x_regular<T, ACC, KM, LD>(X0, src, base, x0, w, vec, stage, acc);
        } else if (!regular && SA.stage_ok && n0 >= kStage) {
            // the end of the line staged (the host checked that every tap and ghost of an item
            // touching x boundary rows lies there)
            const int64_t g0 = x0 < n0 / 2 ? 0 : n0 - kStage;
            x_stage<T, LD>(src, base, g0, vec, stage);
            auto ldx = [&](int64_t j) -> double { return (double)Store<T>::ld(stage, j - g0); };
            x_points<T, ACC, KM>(P, X0, idx, w, ldx, acc, slots);
        } else {
            auto ldx = [&](int64_t j) -> double { return (double)Store<T>::ld(src, base + j); };
            x_points<T, ACC, KM>(P, X0, idx, w, ldx, acc, slots);
        }
    }
    // y and z terms
    for (int a = 1; a < P.ndim; ++a) {
        const AxisDev& X = P.ax[a];
        if (!X.active) continue;
        int64_t base0 = P.tshift[a] + idx[0] * P.ssrc[0];
        for (int c = 1; c < P.ndim; ++c)
            if (c != a) base0 += idx[c] * P.ssrc[c];
        strip_axis_term<T, ACC, KM, LD>(P, X, src, (int)idx[a], base0, P.ssrc[a], w, vec, idx, acc,
                                        slots);
    }
    store_item<T, ACC, LD>(P, idx, w, vec, acc, du, accumulate);
}

// One face item of kind A (item l of shell row `row` along axis A; kind 0: the item's strip
// starts at x0 = row and ends before xend): the other two axes are regular rows at every point,
// the row along A is the same for the whole block.  The host builds face segments only when
// SA.vec_ok, n0 >= 16, the A axis has one BC for all lines and (kind 0) SA.stage_ok.
template <typename T, int ACC, int KM, int A, typename LD>
DOP_HD inline void eval_face(const PlanDev& P, const StripArgs& SA, int row, int xend, int l, T* stage,
                             Slot<ACC>* slots, T* __restrict__ du, const T* __restrict__ src,
                             int accumulate)
{
    const int nu = SA.fn[A][0];
    const int u = l % nu, v = l / nu;
    int64_t idx[3];
    int w = kSW;
    if constexpr (A == 2) {
        idx[0] = SA.flo[2][0] + (int64_t)u * kSW;
        idx[1] = SA.flo[2][1] + v;
        idx[2] = row;
    } else if constexpr (A == 1) {
        idx[0] = SA.flo[1][0] + (int64_t)u * kSW;
        idx[1] = row;
        idx[2] = SA.flo[1][1] + v;
    } else {
        idx[0] = row;
        idx[1] = SA.flo[0][0] + u;
        idx[2] = SA.flo[0][1] + v;
        w = xend - row < kSW ? xend - row : kSW;
    }
    Acc<ACC> acc[kSW];
    const AxisDev& X0 = P.ax[0];
    const AxisDev& X1 = P.ax[1];
    const AxisDev& X2 = P.ax[2];
    const int64_t base = P.tshift[0] + idx[1] * P.ssrc[1] + idx[2] * P.ssrc[2];
    if constexpr (A == 0) {
        const int64_t n0 = X0.n;
        const int64_t g0 = row < n0 / 2 ? 0 : n0 - kStage;
        x_stage<T, LD>(src, base, g0, true, stage);
        auto ldx = [&](int64_t j) -> double { return (double)Store<T>::ld(stage, j - g0); };
        x_points<T, ACC, KM>(P, X0, idx, w, ldx, acc, slots);
    } else {
        x_regular<T, ACC, KM, LD>(X0, src, base, (int)idx[0], w, true, stage, acc);
    }
    const int64_t b1 = P.tshift[1] + idx[0] * P.ssrc[0] + idx[2] * P.ssrc[2];
    const int64_t b2 = P.tshift[2] + idx[0] * P.ssrc[0] + idx[1] * P.ssrc[1];
    if constexpr (A == 1) strip_boundary_or_regular<T, ACC, KM, LD>(X1, src, row, b1, P.ssrc[1], w, acc);
    else strip_regular<T, ACC, KM, LD>(X1, src, (int)idx[1], b1, P.ssrc[1], w, true, acc);
    if constexpr (A == 2) strip_boundary_or_regular<T, ACC, KM, LD>(X2, src, row, b2, P.ssrc[2], w, acc);
    else strip_regular<T, ACC, KM, LD>(X2, src, (int)idx[2], b2, P.ssrc[2], w, true, acc);
    store_item<T, ACC, LD>(P, idx, w, true, acc, du, accumulate);
}

// Work item `l` of row j of segment `s`: the kernel's (and the CPU check's) dispatch.
template <typename T, int ACC, int KM, bool VEC, typename LD>
DOP_HD inline void eval_item(const PlanDev& P, const StripArgs& SA, int s, int j, int l, T* stage,
                             Slot<ACC>* slots, T* __restrict__ du, const T* __restrict__ src,
                             int accumulate)
{
    const ShellSeg& G = SA.seg[s];
    if (l >= G.items || G.kind == 4) return;   // kind 4: the device's z-march
    if (G.kind == 3) {
        eval_strip<T, ACC, KM, VEC, LD>(P, SA, G.r0, l, stage, slots, du, src, accumulate);
        return;
    }
    if constexpr (VEC) {
        if (G.kind == 2)
            eval_face<T, ACC, KM, 2, LD>(P, SA, G.r0 + j, 0, l, stage, slots, du, src, accumulate);
        else if (G.kind == 1)
            eval_face<T, ACC, KM, 1, LD>(P, SA, G.r0 + j, 0, l, stage, slots, du, src, accumulate);
        else
            eval_face<T, ACC, KM, 0, LD>(P, SA, G.r0 + j * kSW, G.xend, l, stage, slots, du, src,
                                         accumulate);
    }
}

}  // namespace strip
}  // namespace dop
