// dop_rows.cpp -- row programs of DiffEqOperators.jl's 1-D convolution kernels.
//
// For each output row the reference evaluates  x_temp[i] = sum_idx c * w[idx] * x[j(idx)]
// over a contiguous window of the padded vector (convolutions.jl).  The window start, the
// table row and the coefficient index depend on the kernel (centered / upwind, uniform /
// non-uniform, AbstractVector / BoundaryPaddedVector) and, for upwind operators, on the
// sign of the row's coefficient.  This file turns those choices into one program per row:
// start J0 (0-based padded index) and weights w[k] = c * w_table[k] (the reference's
// `cur_coeff * cur_stencil[idx]`, formed in FP64 exactly as it does).
//
// Indices below: i is the reference's 1-based row, Jref its 1-based padded index; the
// stored J0 is Jref - 1.
#include "dop_rows.hpp"

#include <algorithm>
#include <cstring>

namespace dop {

namespace {

struct Tab {
    const dop_operator& A;
    int nin() const { return A.len - 2 * A.boundary_point_count; }
    int nhi() const { return A.boundary_point_count + A.offside; }
    // centered
    const double* stencil_row(int k) const {   // k 1-based (non-uniform) ; uniform: the stencil
        return A.uniform_grid ? A.stencil_coefs
                              : A.stencil_coefs + (size_t)(k - 1) * A.stencil_length;
    }
    const double* low(int k) const {
        return A.low_boundary_coefs + (size_t)(k - 1) * A.boundary_stencil_length;
    }
    const double* high(int k) const {
        return A.high_boundary_coefs + (size_t)(k - 1) * A.boundary_stencil_length;
    }
    // upwind non-uniform: [2][nin][SL], [2][BPC][BSL], [2][BPC+off][BSL]
    const double* up(int k) const { return A.stencil_coefs + (size_t)(k - 1) * A.stencil_length; }
    const double* dn(int k) const {
        return A.stencil_coefs + (size_t)(nin() + k - 1) * A.stencil_length;
    }
    const double* low2(int side, int k) const {
        return A.low_boundary_coefs +
               ((size_t)(side - 1) * A.boundary_point_count + (size_t)(k - 1)) *
                   A.boundary_stencil_length;
    }
    const double* high2(int side, int k) const {
        return A.high_boundary_coefs +
               ((size_t)(side - 1) * nhi() + (size_t)(k - 1)) * A.boundary_stencil_length;
    }
};

// program = c * table[0..n-1] starting at 1-based padded Jref
void set_prog(RowProgram* p, int Jref, double c, const double* tab, int n)
{
    p->J0 = Jref - 1;
    p->w.resize(n);
    for (int k = 0; k < n; ++k) p->w[k] = c * tab[k];
}

// program = c * (sgn * reverse(table)) -- the reference's `reverse(stencil)` /
// `-1 * reverse(stencil)` / `(-1)^d * reverse(...)` (convolutions.jl:134-136, 184, 193)
void set_prog_rev(RowProgram* p, int Jref, double c, double sgn, const double* tab, int n)
{
    p->J0 = Jref - 1;
    p->w.resize(n);
    for (int k = 0; k < n; ++k) p->w[k] = c * (sgn * tab[n - 1 - k]);
}

}  // namespace

bool validate_operator(const dop_operator& A, std::string* err)
{
    const int d = A.derivative_order, a = A.approximation_order;
    if (A.len < 1 || d < 0 || a < 1) { *err = "operator: bad len/order"; return false; }
    if (!A.coefficients || !A.stencil_coefs) { *err = "operator: missing tables"; return false; }
    if (A.wind == 0) {
        const int SL = d + a - 1 + (d + a) % 2;
        if (A.stencil_length != SL || A.boundary_stencil_length != d + a ||
            A.boundary_point_count != SL / 2 - 1 || A.offside != 0) {
            *err = "centered operator: stencil sizes do not match CenteredDifference(d, a)";
            return false;
        }
        if (A.len < 2 * A.boundary_point_count + 2) {
            *err = "centered operator: len must be >= 2*boundary_point_count+2";
            return false;
        }
    } else if (A.wind == 1) {
        if (A.offside < 0 || A.offside > (d + a - 1) / 2) {
            *err = "upwind operator: offside out of range (derivative_operator.jl:458-459)";
            return false;
        }
        if (A.stencil_length != d + a || A.boundary_stencil_length != d + a ||
            A.boundary_point_count != d + a - 2 - A.offside || A.boundary_point_count < 0) {
            *err = "upwind operator: stencil sizes do not match UpwindDifference(d, a)";
            return false;
        }
        if (A.len < 2 * A.boundary_point_count + A.offside + 1) {
            *err = "upwind operator: len too small for its boundary rows";
            return false;
        }
    } else {
        *err = "operator: wind must be 0 or 1";
        return false;
    }
    const int nhigh = A.boundary_point_count + (A.wind ? A.offside : 0);
    if ((A.boundary_point_count > 0 && !A.low_boundary_coefs) ||
        (nhigh > 0 && !A.high_boundary_coefs)) {
        *err = "operator: missing boundary tables";
        return false;
    }
    if (A.boundary_stencil_length > A.len + 2 || A.stencil_length > A.len + 2) {
        *err = "operator: stencil longer than the padded vector";
        return false;
    }
    return true;
}

void interior_rows(const dop_operator& A, int sem, int* lo, int* hi)
{
    const int len = A.len, bpc = A.boundary_point_count;
    if (A.wind == 0) {
        if (sem == DOP_SEM_BPV) { *lo = bpc + 1; *hi = len - bpc - 1; }   // rows 2+BPC : len-BPC-1
        else { *lo = bpc; *hi = len - bpc; }                              // rows 1+BPC : len-BPC
    } else {
        *lo = bpc;
        *hi = len - bpc - A.offside;                                      // 1+BPC : len-BPC-off
    }
    if (*hi < *lo) *hi = *lo;
}

bool row_program(const dop_operator& A, int sem, int r, RowProgram* p, std::string* err)
{
    Tab T{A};
    const int len = A.len, SL = A.stencil_length, BSL = A.boundary_stencil_length;
    const int BPC = A.boundary_point_count, off = A.offside;
    const int xlen = len + 2;
    const int i = r + 1;
    const double* coeff = A.coefficients;
    if (A.wind == 0) {
        const int mid = SL / 2;
        if (i <= BPC) {
            // convolve_BC_left!, convolutions.jl:76-95 (and :412-420)
            set_prog(p, 1, coeff[i - 1], T.low(i), BSL);
        } else if (i <= len - BPC) {
            // convolve_interior!, :40-60 (coeff[i]); BoundaryPaddedVector :381-398, 421-433,
            // 447-459 use coeff[i - BPC].  Same window either way.
            const double c = (sem == DOP_SEM_BPV) ? coeff[i - BPC - 1] : coeff[i - 1];
            set_prog(p, i - mid + 1, c, T.stencil_row(i - BPC), SL);
        } else {
            // convolve_BC_right!: plain :106-117 uses coeff[k]; BoundaryPaddedVector
            // :460-468 uses coeff[len - BPC + k]
            const int k = i - (len - BPC);
            const double c = (sem == DOP_SEM_BPV) ? coeff[i - 1] : coeff[k - 1];
            set_prog(p, xlen - BSL + 1, c, T.high(k), BSL);
        }
    } else if (A.uniform_grid) {
        const double* s = A.stencil_coefs;
        const double dsgn = (A.derivative_order % 2 == 0) ? 1.0 : -1.0;
        if (i <= BPC) {
            // convolve_BC_left!, :146-168
            const double c = coeff[i - 1];
            if (c >= 0 && i + SL <= xlen && i >= off) set_prog(p, i + 1 - off, c, s, SL);
            else set_prog(p, 1, c, T.low(i), BSL);
        } else if (i <= len - BPC - off) {
            // convolve_interior!, :130-143
            const double c = coeff[i - 1];
            if (c >= 0) set_prog(p, i + 1 - off, c, s, SL);
            else set_prog_rev(p, i - SL + 2 + off, c, dsgn, s, SL);
        } else {
            // convolve_BC_right!, :178-214 ; row i = len - BPC + k - off
            const int k = i - (len - BPC - off);
            const double c = coeff[i - 1];
            if (c < 0 && xlen - SL - BPC + k >= 1 && k <= BPC + 1)
                set_prog_rev(p, xlen - SL - BPC + k, c, dsgn, s, SL);
            else if (c < 0 && xlen - SL - BPC + k >= 1 && k > BPC + 1)
                set_prog_rev(p, xlen - BSL + 1, c, dsgn, T.high(BPC + off + 1 - k), BSL);
            else if (c >= 0 && k < off + 1)
                set_prog(p, xlen - SL + k + 1 - off, c, s, SL);
            else
                set_prog(p, xlen - BSL + 1, c, T.high(k), BSL);
        }
    } else {
        const double c = coeff[i - 1];
        if (i <= BPC) {
            // convolve_BC_left!, :285-316
            if (c >= 0 && off == 0) set_prog(p, i + 1, c, T.low2(1, i), SL);
            else if (c >= 0 && i < off) set_prog(p, 1, c, T.low2(1, i), SL);
            else if (c >= 0 && i >= off) set_prog(p, i + 1 - off, c, T.low2(1, i), SL);
            else set_prog(p, 1, c, T.low2(2, i), BSL);
        } else if (i <= len - BPC - off) {
            // convolve_interior!, :234-274 (AbstractVector, branches on coeff_func) and
            // :609-630 (BoundaryPaddedVector, always the row's sign)
            bool fwd;
            if (sem == DOP_SEM_LINES && A.coeff_kind == DOP_COEFF_NUM_POS) fwd = true;
            else if (sem == DOP_SEM_LINES && A.coeff_kind == DOP_COEFF_NUM_NEG) fwd = false;
            else fwd = (c >= 0);
            if (fwd) set_prog(p, i + 1 - off, c, T.up(i - BPC), SL);
            else set_prog(p, i - SL + 2 + off, c, T.dn(i - BPC), SL);
        } else {
            // convolve_BC_right!, :329-361
            const int k = i - len + BPC + off;
            if (c < 0) {
                if (i <= len - off) set_prog(p, i - SL + 2 + off, c, T.high2(2, k), SL);
                else set_prog(p, len - SL + 3, c, T.high2(2, k), SL);
            } else {
                if (i <= len - BPC) set_prog(p, i - SL + 2 + off, c, T.up(i - BPC), SL);
                else set_prog(p, len - BSL + 3, c, T.high2(1, k), BSL);
            }
        }
    }
    const int last = p->J0 + (int)p->w.size() - 1;
    if (p->J0 < 0 || last > len + 1) {
        *err = "operator row reads outside the padded vector (len too small for this stencil)";
        return false;
    }
    return true;
}

bool build_axis(const std::vector<const dop_operator*>& ops, int sem, int n, AxisTables* out,
                std::string* err)
{
    AxisTables& T = *out;
    T = AxisTables{};
    AxisDev& m = T.meta;
    m.n = n;
    if (ops.empty()) return true;
    m.active = 1;
    // merged program of every row: union window, weights summed in operator order
    std::vector<int> J0(n), J1(n);   // inclusive padded window
    std::vector<std::vector<double>> W(n);
    RowProgram p;
    int lo = 0, hi = n;
    for (size_t t = 0; t < ops.size(); ++t) {
        int l, h;
        interior_rows(*ops[t], sem, &l, &h);
        lo = std::max(lo, l);
        hi = std::min(hi, h);
    }
    if (hi < lo) hi = lo;
    for (int r = 0; r < n; ++r) {
        int a = 1 << 30, b = -(1 << 30);
        std::vector<RowProgram> progs(ops.size());
        for (size_t t = 0; t < ops.size(); ++t) {
            if (!row_program(*ops[t], sem, r, &progs[t], err)) return false;
            a = std::min(a, progs[t].J0);
            b = std::max(b, progs[t].J0 + (int)progs[t].w.size() - 1);
        }
        J0[r] = a;
        J1[r] = b;
        W[r].assign(b - a + 1, 0.0);
        for (auto& q : progs)
            for (size_t k = 0; k < q.w.size(); ++k) W[r][q.J0 - a + k] += q.w[k];
    }
    // regular window over the interior rows, relative to the row in u-index terms
    // (u index = J - 1): rel = J - 1 - r
    int dmin = 1 << 30, dmax = -(1 << 30);
    for (int r = lo; r < hi; ++r) {
        dmin = std::min(dmin, J0[r] - 1 - r);
        dmax = std::max(dmax, J1[r] - 1 - r);
    }
    int rlo = 0, rhi = 0;
    if (hi > lo) {
        rlo = std::max(lo, -dmin);
        rhi = std::min(hi, n - dmax);
        if (rhi < rlo) rlo = rhi = 0;
    }
    const int K = (rhi > rlo) ? dmax - dmin + 1 : 0;
    if (K > 32) { *err = "stencil window wider than 32 taps"; return false; }
    m.K = K;
    m.dmin = (rhi > rlo) ? dmin : 0;
    m.rlo = rlo;
    m.rhi = rhi;
    // regular weights
    if (rhi > rlo) {
        bool uni = true;
        std::vector<double> row0(K, 0.0), rowr(K, 0.0);
        auto fill = [&](int r, std::vector<double>& dst) {
            std::fill(dst.begin(), dst.end(), 0.0);
            for (int J = J0[r]; J <= J1[r]; ++J) dst[(J - 1 - r) - dmin] = W[r][J - J0[r]];
        };
        fill(rlo, row0);
        for (int r = rlo + 1; r < rhi && uni; ++r) {
            fill(r, rowr);
            uni = std::memcmp(row0.data(), rowr.data(), sizeof(double) * K) == 0;
        }
        m.uniform = uni ? 1 : 0;
        if (uni) {
            T.w64 = row0;
        } else {
            // tap-major, w[k * n + r]: the threads of neighbouring rows read neighbouring
            // words for every tap
            T.w64.assign((size_t)n * K, 0.0);
            for (int r = rlo; r < rhi; ++r) {
                fill(r, rowr);
                for (int k = 0; k < K; ++k) T.w64[(size_t)k * n + r] = rowr[k];
            }
        }
        T.w32.resize(T.w64.size());
        for (size_t k = 0; k < T.w64.size(); ++k) T.w32[k] = (float)T.w64[k];
    }
    // boundary rows
    const int nb = rlo + (n - rhi);
    int KB = 0;
    for (int r = 0; r < n; ++r)
        if (r < rlo || r >= rhi) KB = std::max(KB, J1[r] - J0[r] + 1);
    m.nb = nb;
    m.KB = KB;
    T.bJ0.assign(nb, 0);
    T.bn.assign(nb, 0);
    T.bw.assign((size_t)nb * std::max(KB, 1), 0.0);
    for (int r = 0, b = 0; r < n; ++r) {
        if (r >= rlo && r < rhi) continue;
        T.bJ0[b] = J0[r];
        T.bn[b] = J1[r] - J0[r] + 1;
        for (int k = 0; k < T.bn[b]; ++k) T.bw[(size_t)b * KB + k] = W[r][k];
        ++b;
    }
    return true;
}

}  // namespace dop
