// dop_concrete.hpp -- the matrix form of a plan.  A plan applies the affine map
//     du = A * vec(src) + b        (accumulate = 0)
// where src is u (ghost input) or the padded array M (padded input), both column-major, and
// b collects the BCs' affine parts.  For one output point, row_entries() enumerates the
// (source index, weight) pairs of its row of A and the terms of b, with the same index
// arithmetic as eval_point() in dop_eval.hpp; RowList merges them into one sorted row.
// __host__ __device__ so that the CPU test build evaluates exactly what the kernels export.
#pragma once
#include <cstdint>

#include "dop_plan.hpp"

namespace dop {

// Entries one output row can emit before merging: per active axis, the regular window or the
// widest boundary row with both ghosts expanded into their taps.
DOP_HD inline int row_entry_bound(const PlanDev& P)
{
    int s = 0;
    for (int a = 0; a < P.ndim; ++a) {
        const AxisDev& X = P.ax[a];
        if (!X.active) continue;
        const int nl = X.periodic ? 1 : X.nl, nr = X.periodic ? 1 : X.nr;
        const int bnd = X.KB + (P.ghost ? nl + nr : 0);
        s += X.K > bnd ? X.K : bnd;
    }
    return s;
}

// The ghost value a*u[..] + b of one line expands into its taps and its constant; weight w is
// the row's weight on the ghost (bc_operators.jl:188-192; periodic: multi_dim_bc_operators.jl:
// 221-252 for the N-D form).  col < 0 marks a constant term.
template <typename F>
DOP_HD inline void ghost_entries(const AxisDev& X, bool left, int64_t base, int64_t S, double w,
                                 int64_t line, F& emit)
{
#pragma clang fp contract(off)
    if (X.periodic) {
        const bool last = left ? X.periodic == 1 : X.periodic == 2;
        emit(base + (last ? (int64_t)(X.n - 1) * S : 0), w);
        return;
    }
    const double* e = X.lbc ? X.lbc + line * kLineBC : nullptr;
    if (e && e[0] < 0.0) {   // a periodic line: u[n-1] on the left, u[0] on the right
        emit(base + (left ? (int64_t)(X.n - 1) * S : 0), w);
        return;
    }
    if (left) {
        const int nl = e ? (int)e[0] : X.nl;
        const double* al = e ? e + 2 : X.al;
        for (int m = 0; m < nl; ++m) emit(base + (int64_t)m * S, w * al[m]);
        emit(-1, w * (e ? e[2 + kMaxBCTaps] : X.bl));
    } else {
        const int nr = e ? (int)e[1] : X.nr;
        const double* ar = e ? e + 3 + kMaxBCTaps : X.ar;
        const int64_t o = (int64_t)X.n - nr;
        for (int m = 0; m < nr; ++m) emit(base + (o + m) * S, w * ar[m]);
        emit(-1, w * (e ? e[3 + 2 * kMaxBCTaps] : X.br));
    }
}

// Row of output point idx: emit(col, w) for every tap, emit(-1, c) for every constant.
template <typename F>
DOP_HD inline void row_entries(const PlanDev& P, const int64_t* idx, F& emit)
{
    for (int a = 0; a < P.ndim; ++a) {
        const AxisDev& X = P.ax[a];
        if (!X.active) continue;
        const int r = (int)idx[a];
        const int64_t S = P.ssrc[a];
        int64_t base = P.tshift[a];
        for (int b = 0; b < P.ndim; ++b)
            if (b != a) base += idx[b] * P.ssrc[b];
        if (r >= X.rlo && r < X.rhi) {
            const double* w = X.w64 + (X.uniform ? 0 : r);   // tap-major table
            const int64_t ws = X.uniform ? 1 : X.n;
            const int64_t p0 = base + (int64_t)(r + X.dmin + (P.ghost ? 0 : 1)) * S;
            for (int k = 0; k < X.K; ++k) emit(p0 + k * S, w[k * ws]);
        } else {
            const int b = r < X.rlo ? r : X.rlo + (r - X.rhi);
            const int J0 = X.bJ0[b];
            const int nt = X.bn[b];
            const double* w = X.bw + (int64_t)b * X.KB;
            const int joff = P.ghost ? -1 : 0;
            const int64_t line = X.lbc ? idx[0] * X.lstride[0] + idx[1] * X.lstride[1] +
                                             idx[2] * X.lstride[2]
                                       : 0;
            for (int k = 0; k < nt; ++k) {
                const int J = J0 + k;
                if (P.ghost && J == 0) ghost_entries(X, true, base, S, w[k], line, emit);
                else if (P.ghost && J == X.n + 1) ghost_entries(X, false, base, S, w[k], line, emit);
                else emit(base + (int64_t)(J + joff) * S, w[k]);
            }
        }
    }
}

// One row, merged: columns ascending, weights of a repeated column summed in emission order
// (for one axis: left ghost, line taps, right ghost -- the order of the padded index of
// L * Q_L, concretization.jl:866-870), constants summed into b.
template <int CAP>
struct RowList {
    int n = 0;
    double b = 0.0;
    int64_t col[CAP];
    double val[CAP];
    // contraction off: the products emitted above are rounded before they are summed, on the
    // device as in the CPU build
    DOP_HD void operator()(int64_t c, double v)
    {
#pragma clang fp contract(off)
        if (c < 0) {
            b += v;
            return;
        }
        int i = n;
        while (i > 0 && col[i - 1] > c) --i;
        if (i > 0 && col[i - 1] == c) {
            val[i - 1] += v;
            return;
        }
        for (int j = n; j > i; --j) {
            col[j] = col[j - 1];
            val[j] = val[j - 1];
        }
        col[i] = c;
        val[i] = v;
        ++n;
    }
    // stored entries: exact zeros (zero taps of merged windows, homogeneous Dirichlet ghosts)
    // are not stored
    DOP_HD int nnz() const
    {
        int c = 0;
        for (int i = 0; i < n; ++i) c += val[i] != 0.0;
        return c;
    }
};

DOP_HD inline void point_index(const PlanDev& P, int64_t p, int64_t* idx)
{
    idx[0] = p % P.n[0];
    const int64_t q = p / P.n[0];
    idx[1] = q % P.n[1];
    idx[2] = q / P.n[1];
}

}  // namespace dop
