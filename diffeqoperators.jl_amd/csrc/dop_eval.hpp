// dop_eval.hpp -- storage conversions, accumulators and the per-point evaluation of a
// plan (the generic path: any ndim, any row program).  __host__ __device__ so that the
// index arithmetic can be checked on the CPU against the FP64 oracle.
#pragma once
#include <cmath>
#include <cstdint>

#include "dop_plan.hpp"

namespace dop {

// ---------------------------------------------------------------------------------------
// storage types: double, float, and bf16 held as uint16_t
struct bf16 {
    uint16_t bits;
};

DOP_HD inline float bf16_to_f32(uint16_t b)
{
    return __builtin_bit_cast(float, (uint32_t)b << 16);
}

// round to nearest even; NaN stays NaN (quiet)
DOP_HD inline uint16_t f32_to_bf16(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename T>
struct Store;
template <>
struct Store<double> {
    using C = double;
    DOP_HD static double ld(const double* p, int64_t i) { return p[i]; }
    DOP_HD static void st(double* p, int64_t i, double v) { p[i] = v; }
};
template <>
struct Store<float> {
    using C = float;
    DOP_HD static float ld(const float* p, int64_t i) { return p[i]; }
    DOP_HD static void st(float* p, int64_t i, double v) { p[i] = (float)v; }
};
template <>
struct Store<bf16> {
    using C = float;
    DOP_HD static float ld(const bf16* p, int64_t i) { return bf16_to_f32(p[i].bits); }
    DOP_HD static void st(bf16* p, int64_t i, double v) { p[i].bits = f32_to_bf16((float)v); }
};

// ---------------------------------------------------------------------------------------
// accumulators
template <int ACC>
struct Acc;

template <>
struct Acc<kAccF64> {
    using W = double;   // weight type of the regular rows
    double s = 0.0;
    DOP_HD void add(double w, double x) { s = fma(w, x, s); }
    DOP_HD void add_sum(double v) { s += v; }
    DOP_HD double value() const { return s; }
    DOP_HD static const double* table(const AxisDev& X) { return X.w64; }
};

template <>
struct Acc<kAccF32> {
    using W = float;
    float s = 0.0f;
    DOP_HD void add(float w, float x) { s = fmaf(w, x, s); }
    DOP_HD void add_sum(double v) { s += (float)v; }
    DOP_HD double value() const { return s; }
    DOP_HD static const float* table(const AxisDev& X) { return X.w32; }
};

// Compensated dot product in f32 (Ogita, Rump & Oishi 2005, "Dot2"): the product error
// comes from one FMA (TwoProduct), the addition error from Knuth's TwoSum.  Contraction
// is switched off so that the error terms are formed exactly as written.
template <>
struct Acc<kAccF32Kahan> {
    using W = float;
    float s = 0.0f, c = 0.0f;
    DOP_HD void two_sum_add(float p)
    {
#pragma clang fp contract(off)
        const float t = s + p;
        const float z = t - s;
        const float e = (s - (t - z)) + (p - z);
        s = t;
        c = c + e;
    }
    DOP_HD void add(float w, float x)
    {
#pragma clang fp contract(off)
        const float p = w * x;
        const float pe = fmaf(w, x, -p);
        two_sum_add(p);
        c = c + pe;
    }
    DOP_HD void add_sum(double v)
    {
#pragma clang fp contract(off)
        const float hi = (float)v;
        const float lo = (float)(v - (double)hi);
        two_sum_add(hi);
        c = c + lo;
    }
    DOP_HD double value() const
    {
#pragma clang fp contract(off)
        return (double)s + (double)c;
    }
    DOP_HD static const float* table(const AxisDev& X) { return X.w32; }
};

// f32 weights and values, f64 sum: every product of two f32 numbers is exact in f64, so the
// only roundings are the f64 additions (error ~K * 2^-53 of the sum of |terms|).
template <>
struct Acc<kAccF32Wide> {
    using W = float;
    double s = 0.0;
    DOP_HD void add(float w, float x) { s = fma((double)w, (double)x, s); }
    DOP_HD void add_sum(double v) { s += v; }
    DOP_HD double value() const { return s; }
    DOP_HD static const float* table(const AxisDev& X) { return X.w32; }
};

// ---------------------------------------------------------------------------------------
// ghost values of one line: a . u[0:nl] + b (bc_operators.jl:188-191), with the line's own
// BC when the axis has one per line (multi_dim_bc_operators.jl:34, 93-95); a periodic line
// wraps (u[n-1] on the left, u[0] on the right, bc_operators.jl:192).  ld(j) is element j of
// the line, as a double.
template <typename F>
DOP_HD inline double ghost_left_f(const AxisDev& X, int64_t line, const F& ld)
{
    if (X.periodic) return ld(X.periodic == 1 ? (int64_t)(X.n - 1) : 0);
    const double* e = X.lbc ? X.lbc + line * kLineBC : nullptr;
    if (e && tab(e, 0) < 0.0) return ld((int64_t)(X.n - 1));
    const int nl = e ? (int)tab(e, 0) : X.nl;
    const double* al = e ? e + 2 : X.al;
    double g = 0.0;
    for (int k = 0; k < nl; ++k) g = fma(tab(al, k), ld(k), g);
    return g + (e ? tab(e, 2 + kMaxBCTaps) : X.bl);
}

template <typename F>
DOP_HD inline double ghost_right_f(const AxisDev& X, int64_t line, const F& ld)
{
    if (X.periodic) return ld(X.periodic == 2 ? (int64_t)(X.n - 1) : 0);
    const double* e = X.lbc ? X.lbc + line * kLineBC : nullptr;
    if (e && tab(e, 1) < 0.0) return ld(0);
    const int nr = e ? (int)tab(e, 1) : X.nr;
    const double* ar = e ? e + 3 + kMaxBCTaps : X.ar;
    double g = 0.0;
    const int64_t o = (int64_t)X.n - nr;
    for (int k = 0; k < nr; ++k) g = fma(tab(ar, k), ld(o + k), g);
    return g + (e ? tab(e, 3 + 2 * kMaxBCTaps) : X.br);
}

template <typename T>
DOP_HD inline double ghost_left(const AxisDev& X, const T* src, int64_t base, int64_t S,
                                int64_t line)
{
    return ghost_left_f(X, line, [&](int64_t j) { return (double)Store<T>::ld(src, base + j * S); });
}

template <typename T>
DOP_HD inline double ghost_right(const AxisDev& X, const T* src, int64_t base, int64_t S,
                                 int64_t line)
{
    return ghost_right_f(X, line, [&](int64_t j) { return (double)Store<T>::ld(src, base + j * S); });
}

// line index of output point idx along axis a (per-line BCs only)
DOP_HD inline int64_t line_of(const AxisDev& X, const int64_t* idx)
{
    return X.lbc ? idx[0] * X.lstride[0] + idx[1] * X.lstride[1] + idx[2] * X.lstride[2] : 0;
}

constexpr int kEvalUnroll = 8;

// FP64 value of a boundary row (always evaluated in double).  The ghosts a row touches are
// formed once, and every tap's load is issued before the first FMA.  ld(j) is element j of
// the source line (u, or the padded M), as a double.
template <typename F>
DOP_HD inline double boundary_row_f(const PlanDev& P, const AxisDev& X, int r, int64_t line,
                                    const F& ld)
{
    const int b = r < X.rlo ? r : X.rlo + (r - X.rhi);
    const int J0 = tab(X.bJ0, b);
    const int nt = tab(X.bn, b);
    const double* w = X.bw + (int64_t)b * X.KB;
    const int joff = P.ghost ? -1 : 0;
    const bool gl_used = P.ghost && J0 == 0;
    const bool gr_used = P.ghost && J0 + nt - 1 == X.n + 1;
    const double gl = gl_used ? ghost_left_f(X, line, ld) : 0.0;
    const double gr = gr_used ? ghost_right_f(X, line, ld) : 0.0;
    double s = 0.0;
    if (nt <= kEvalUnroll) {
        double v[kEvalUnroll];
#pragma unroll
        for (int k = 0; k < kEvalUnroll; ++k) {
            if (k < nt) {
                const int J = J0 + k;
                if (gl_used && J == 0) v[k] = gl;
                else if (gr_used && J == X.n + 1) v[k] = gr;
                else v[k] = ld((int64_t)(J + joff));
            }
        }
#pragma unroll
        for (int k = 0; k < kEvalUnroll; ++k)
            if (k < nt) s = fma(tab(w, k), v[k], s);
        return s;
    }
    for (int k = 0; k < nt; ++k) {
        const int J = J0 + k;
        double v;
        if (gl_used && J == 0) v = gl;
        else if (gr_used && J == X.n + 1) v = gr;
        else v = ld((int64_t)(J + joff));
        s = fma(tab(w, k), v, s);
    }
    return s;
}

template <typename T>
DOP_HD inline double boundary_row(const PlanDev& P, const AxisDev& X, const T* src, int r,
                                  int64_t base, int64_t S, int64_t line)
{
    return boundary_row_f(P, X, r, line,
                          [&](int64_t j) { return (double)Store<T>::ld(src, base + j * S); });
}

// The terms of axis X (row r, line origin `base`, stride S) added to acc: a regular row's K
// taps in order, or a boundary row's FP64 value.  Shared by eval_point and the kernels that
// stage a line elsewhere (same arithmetic, same order).
template <typename T, int ACC>
DOP_HD inline void axis_term(const PlanDev& P, const AxisDev& X, const T* src, int r, int64_t base,
                             int64_t S, int64_t line, Acc<ACC>& acc)
{
    if (r >= X.rlo && r < X.rhi) {
        // weights of row r: tap k at w[k * ws] (tap-major table; one row when uniform)
        const auto* w = Acc<ACC>::table(X) + (X.uniform ? 0 : r);
        const int64_t ws = X.uniform ? 1 : X.n;
        const int64_t p0 = base + (int64_t)(r + X.dmin + (P.ghost ? 0 : 1)) * S;
        // all taps' loads are issued before the first FMA (K <= kEvalUnroll covers every
        // centered/upwind window of approximation order <= 8 without a loop)
        if (X.K <= kEvalUnroll) {
            typename Acc<ACC>::W wv[kEvalUnroll], xv[kEvalUnroll];
#pragma unroll
            for (int k = 0; k < kEvalUnroll; ++k) {
                if (k < X.K) {
                    wv[k] = tab(w, k * ws);
                    xv[k] = (typename Acc<ACC>::W)Store<T>::ld(src, p0 + k * S);
                }
            }
#pragma unroll
            for (int k = 0; k < kEvalUnroll; ++k)
                if (k < X.K) acc.add(wv[k], xv[k]);
        } else {
            for (int k = 0; k < X.K; ++k)
                acc.add(tab(w, k * ws), (typename Acc<ACC>::W)Store<T>::ld(src, p0 + k * S));
        }
    } else {
        acc.add_sum(boundary_row<T>(P, X, src, r, base, S, line));
    }
}

// du value at multi-index idx (without the accumulate term)
template <typename T, int ACC>
DOP_HD inline double eval_point(const PlanDev& P, const T* src, const int64_t* idx)
{
    Acc<ACC> acc;
    for (int a = 0; a < P.ndim; ++a) {
        const AxisDev& X = P.ax[a];
        if (!X.active) continue;
        int64_t base = P.tshift[a];
        for (int b = 0; b < P.ndim; ++b)
            if (b != a) base += idx[b] * P.ssrc[b];
        axis_term<T, ACC>(P, X, src, (int)idx[a], base, P.ssrc[a], line_of(X, idx), acc);
    }
    return acc.value();
}

}  // namespace dop
