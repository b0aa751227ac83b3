// dop_build.cpp -- dop_desc -> HostPlan.
//
// Shape rules follow the reference's assertions: every operator's axis must be padded in
// the source (derivative_operator_functions.jl:40), other axes either equal or padded by
// two (:33-35), len must match the axis extent, and an operator applied through Q needs
// a BC on its axis (multi_dim_bc_operators.jl:230-234).
#include "dop_build.hpp"

#include <algorithm>

namespace dop {

int build_host_plan(const dop_desc& D, HostPlan* out, std::string* err)
{
    HostPlan& H = *out;
    H = HostPlan{};
    PlanDev& P = H.meta;
    if (D.ndim < 1 || D.ndim > kMaxDim) { *err = "ndim must be 1, 2 or 3"; return DOP_EINVAL; }
    if (D.storage != DOP_F64 && D.storage != DOP_F32 && D.storage != DOP_BF16) {
        *err = "unknown storage type";
        return DOP_EINVAL;
    }
    if (D.input != DOP_INPUT_GHOST && D.input != DOP_INPUT_PADDED) {
        *err = "unknown input kind";
        return DOP_EINVAL;
    }
    if (D.semantics != DOP_SEM_LINES && D.semantics != DOP_SEM_BPV) {
        *err = "unknown semantics";
        return DOP_EINVAL;
    }
    if (D.semantics == DOP_SEM_BPV && (D.input != DOP_INPUT_GHOST || D.ndim != 1)) {
        *err = "BoundaryPaddedVector semantics only arise from 1-D Q*u";
        return DOP_EINVAL;
    }
    if (D.n_ops < 1 || !D.ops) { *err = "no operators"; return DOP_EINVAL; }
    P.ndim = D.ndim;
    P.storage = D.storage;
    P.ghost = D.input == DOP_INPUT_GHOST ? 1 : 0;
    int64_t msz[kMaxDim];
    for (int k = 0; k < kMaxDim; ++k) {
        if (k < D.ndim) {
            if (D.size[k] < 1 || D.size[k] > (1LL << 31) - 4) {
                *err = "extent out of range";
                return DOP_EINVAL;
            }
            P.n[k] = D.size[k];
            const int pad = (D.input == DOP_INPUT_PADDED) ? (D.padded[k] ? 2 : 0) : 0;
            if (D.input == DOP_INPUT_PADDED && D.padded[k] != 0 && D.padded[k] != 1) {
                *err = "padded[] entries must be 0 or 1";
                return DOP_EINVAL;
            }
            msz[k] = P.n[k] + pad;
        } else {
            P.n[k] = 1;
            msz[k] = 1;
        }
    }
    P.sdu[0] = 1;
    P.ssrc[0] = 1;
    for (int k = 1; k < kMaxDim; ++k) {
        P.sdu[k] = P.sdu[k - 1] * P.n[k - 1];
        P.ssrc[k] = P.ssrc[k - 1] * msz[k - 1];
    }
    P.total = P.n[0] * P.n[1] * P.n[2];
    H.out_elems = P.total;
    H.in_elems = msz[0] * msz[1] * msz[2];

    // operators, grouped by axis in their given order
    std::vector<const dop_operator*> by_axis[kMaxDim];
    for (int t = 0; t < D.n_ops; ++t) {
        const dop_operator& A = D.ops[t];
        if (A.axis < 0 || A.axis >= D.ndim) {
            *err = "operator axis outside the array's dimensions";
            return DOP_EINVAL;
        }
        if (A.len != D.size[A.axis]) {
            *err = "operator len does not match the array extent along its axis";
            return DOP_EDIM;
        }
        if (!validate_operator(A, err)) return DOP_EINVAL;
        if (A.coeff_kind < DOP_COEFF_VECTOR || A.coeff_kind > DOP_COEFF_NUM_NEG) {
            *err = "unknown coeff_kind";
            return DOP_EINVAL;
        }
        if (D.input == DOP_INPUT_GHOST && !D.bc[A.axis].present) {
            *err = "operator axis has no boundary condition";
            return DOP_EINVAL;
        }
        if (D.input == DOP_INPUT_PADDED && !D.padded[A.axis]) {
            *err = "differentiated dimension must be padded";
            return DOP_EINVAL;
        }
        H.max_derivative_order = std::max(H.max_derivative_order, (int)A.derivative_order);
        by_axis[A.axis].push_back(&A);
    }
    // accumulation
    int acc = D.accum;
    if (acc == DOP_ACC_AUTO) {
        if (D.storage == DOP_F64) acc = DOP_ACC_F64;
        // high derivative orders cancel large alternating weights: exact f32 products
        // summed in f64
        else acc = (H.max_derivative_order >= 3) ? DOP_ACC_F32_WIDE : DOP_ACC_F32;
    }
    if (acc != DOP_ACC_F64 && acc != DOP_ACC_F32 && acc != DOP_ACC_F32_KAHAN &&
        acc != DOP_ACC_F32_WIDE) {
        *err = "unknown accumulation mode";
        return DOP_EINVAL;
    }
    if (D.storage == DOP_F64 && acc != DOP_ACC_F64) {
        *err = "f64 storage requires f64 accumulation";
        return DOP_EINVAL;
    }
    P.accum = acc;
    // per-axis tables, ghosts and source offsets
    for (int a = 0; a < kMaxDim; ++a) {
        AxisTables& T = H.axis[a];
        if (a < D.ndim) {
            if (!build_axis(by_axis[a], D.semantics, (int)P.n[a], &T, err)) return DOP_EINVAL;
        } else {
            T.meta = AxisDev{};
            T.meta.n = 1;
        }
        AxisDev& X = T.meta;
        if (a < D.ndim && D.input == DOP_INPUT_GHOST && D.bc[a].present) {
            const dop_affine_bc& q = D.bc[a];
            if (q.kind != DOP_BC_AFFINE && q.kind != DOP_BC_PERIODIC && q.kind != DOP_BC_PERIODIC_ND) {
                *err = "unknown boundary condition kind";
                return DOP_EINVAL;
            }
            X.periodic = q.kind;
            if (q.kind != DOP_BC_AFFINE) {
                X.nl = X.nr = 0;
                X.bl = X.br = 0.0;
                for (int k = 0; k < DOP_MAX_BC_TAPS; ++k) X.al[k] = X.ar[k] = 0.0;
            } else if (q.n_l < 0 || q.n_r < 0 || q.n_l > DOP_MAX_BC_TAPS || q.n_r > DOP_MAX_BC_TAPS ||
                q.n_l > P.n[a] || q.n_r > P.n[a]) {
                *err = "boundary condition has more taps than the axis has points";
                return DOP_EINVAL;
            } else {
                X.nl = q.n_l;
                X.nr = q.n_r;
                for (int k = 0; k < DOP_MAX_BC_TAPS; ++k) {
                    X.al[k] = k < q.n_l ? q.a_l[k] : 0.0;
                    X.ar[k] = k < q.n_r ? q.a_r[k] : 0.0;
                }
                X.bl = q.b_l;
                X.br = q.b_r;
            }
            if (const dop_affine_bc* lb = D.line_bc[a]) {
                // one BC per line: the other axes' extents, column-major
                if (q.kind != DOP_BC_AFFINE) {
                    *err = "per-line boundary conditions need an affine axis BC";
                    return DOP_EINVAL;
                }
                int64_t nlines = 1;
                for (int b = 0; b < D.ndim; ++b) {
                    X.lstride[b] = b == a ? 0 : nlines;
                    if (b != a) nlines *= P.n[b];
                }
                T.lbc.assign((size_t)nlines * kLineBC, 0.0);
                int mnl = 0, mnr = 0;
                for (int64_t l = 0; l < nlines; ++l) {
                    const dop_affine_bc& e = lb[l];
                    double* o = T.lbc.data() + (size_t)l * kLineBC;
                    if (e.kind == DOP_BC_PERIODIC) {
                        // the line wraps: ghosts u[end], u[1] (one tap each)
                        o[0] = o[1] = kLinePeriodic;
                        mnl = std::max(mnl, 1);
                        mnr = std::max(mnr, 1);
                        continue;
                    }
                    if (e.kind != DOP_BC_AFFINE || e.n_l < 0 || e.n_r < 0 ||
                        e.n_l > DOP_MAX_BC_TAPS || e.n_r > DOP_MAX_BC_TAPS || e.n_l > P.n[a] ||
                        e.n_r > P.n[a]) {
                        *err = "per-line boundary condition " + std::to_string(l) +
                               ": not affine or periodic, or more taps than the axis has points";
                        return DOP_EINVAL;
                    }
                    o[0] = e.n_l;
                    o[1] = e.n_r;
                    for (int k = 0; k < e.n_l; ++k) o[2 + k] = e.a_l[k];
                    o[2 + kMaxBCTaps] = e.b_l;
                    for (int k = 0; k < e.n_r; ++k) o[3 + kMaxBCTaps + k] = e.a_r[k];
                    o[3 + 2 * kMaxBCTaps] = e.b_r;
                    mnl = std::max(mnl, (int)e.n_l);
                    mnr = std::max(mnr, (int)e.n_r);
                }
                // the widest line bounds the entries of a matrix row (concretization)
                X.nl = mnl;
                X.nr = mnr;
            }
        } else if (a < D.ndim && D.line_bc[a]) {
            *err = "per-line boundary conditions need GHOST input and a present axis BC";
            return DOP_EINVAL;
        }
        int64_t ts = 0;
        if (D.input == DOP_INPUT_PADDED) {
            for (int b = 0; b < D.ndim; ++b)
                if (b != a && D.padded[b] && !D.transverse_view) ts += P.ssrc[b];
        }
        P.tshift[a] = ts;
        P.ax[a] = X;
    }
    return DOP_OK;
}

bool same_structure(const PlanDev& a, const PlanDev& b)
{
    if (a.ndim != b.ndim || a.storage != b.storage || a.accum != b.accum || a.ghost != b.ghost ||
        a.total != b.total)
        return false;
    for (int k = 0; k < kMaxDim; ++k) {
        if (a.n[k] != b.n[k] || a.sdu[k] != b.sdu[k] || a.ssrc[k] != b.ssrc[k] ||
            a.tshift[k] != b.tshift[k])
            return false;
        const AxisDev &x = a.ax[k], &y = b.ax[k];
        if (x.n != y.n || x.active != y.active || x.K != y.K || x.dmin != y.dmin ||
            x.rlo != y.rlo || x.rhi != y.rhi || x.uniform != y.uniform || x.KB != y.KB ||
            x.nb != y.nb || x.nl != y.nl || x.nr != y.nr || x.periodic != y.periodic ||
            (x.lbc == nullptr) != (y.lbc == nullptr))
            return false;
        for (int m = 0; m < kMaxDim; ++m)
            if (x.lstride[m] != y.lstride[m]) return false;
    }
    return true;
}

}  // namespace dop
