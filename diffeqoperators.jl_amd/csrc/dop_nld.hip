// dop_nld.hip -- nonlinear_diffusion! (derivative_operator.jl:34-66) on the device: the
// Leibniz expansion of d^s/dx^s (p * d^f q / dx^f) on padded vectors,
//     du = sum_{l=0..s} C(s, l) (D^{l+f} q) .* (D^{s-l} p),   D^0 = the interior 2:n+1,
// every D^k a CenteredDifference plan on the padded vector, all terms of a point evaluated by
// one thread in one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>

#include "../../include/dop.h"
#include "dop_eval.hpp"
#include "dop_objects.hpp"

namespace dop {
namespace {

constexpr int kNldMax = 9;   // s + 1 terms, s <= 8

struct NldArgs {
    const PlanDev* q[kNldMax];   // D^{l+f} on q (nullptr: the interior of q)
    const PlanDev* p[kNldMax];   // D^{s-l} on p (nullptr: the interior of p)
    double c[kNldMax];           // C(s, l)
    int nterms;
    int64_t n;
};

template <typename T>
__device__ inline double term(const PlanDev* P, const T* v, int64_t i)
{
    if (!P) return (double)Store<T>::ld(v, i + 1);
    int64_t idx[3] = {i, 0, 0};
    return eval_point<T, kAccF64>(*P, v, idx);
}

// the reference's order: du = (D^f q) .* (D^s p); du = du + (C(s,l) * D^{l+f} q) .* D^{s-l} p
// for l = 1 .. s-1; du = du + (D^{f+s} q) .* p[2:n+1]
template <typename T>
__global__ __launch_bounds__(256) void k_nld(const NldArgs A, T* __restrict__ du,
                                             const T* __restrict__ p, const T* __restrict__ q)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < A.n;
         i += (int64_t)gridDim.x * 256) {
        double acc = term<T>(A.q[0], q, i) * term<T>(A.p[0], p, i);
        for (int l = 1; l < A.nterms - 1; ++l)
            acc = acc + (A.c[l] * term<T>(A.q[l], q, i)) * term<T>(A.p[l], p, i);
        if (A.nterms > 1) {
            const int l = A.nterms - 1;
            acc = acc + term<T>(A.q[l], q, i) * term<T>(A.p[l], p, i);
        }
        Store<T>::st(du, i, acc);
    }
}

// Uniform grids (every term's operator one row of weights): four points per thread.  The
// union of the points' windows of q and of p is read once, as four-value loads, and every
// term of a point is formed from those registers with the arithmetic of k_nld (taps in order
// with fma, the products and sums in the same order), so the two kernels agree bit for bit.
// Groups of four points with a boundary row in any term, or whose windows' four-value loads
// would pass the end of the array, and the last points of a length not a multiple of four, use
// k_nld's per-point form.
constexpr int kNldTaps = 9;

struct NldTerm {
    int on;                   // 0: the interior value v[i + 1]
    int off;                  // first tap of row i at v[i + off]
    int K;
    double w[kNldTaps];
};

struct NldFast {
    NldTerm q[kNldMax], p[kNldMax];
    int qlo, plo;             // windows start at v[i + lo]
    int64_t ilo, ihi;         // points whose every term is a regular row
    int64_t qend, pend;       // groups starting below these read in bounds
};

// value of one term at point i of a group whose window (from v[i0 + lo]) is in xv
template <int NW>
__device__ inline double fast_term(const NldTerm& t, int lo, int i, const double* xv)
{
    double acc = 0.0;
    if (!t.on) {
#pragma unroll
        for (int e = 0; e < NW; ++e)
            if (e == i + 1 - lo) acc = xv[e];
        return acc;
    }
    Acc<kAccF64> a;
#pragma unroll
    for (int e = 0; e < NW; ++e) {
        const int k = e - (i + t.off - lo);
        if (k >= 0 && k < t.K) a.add(t.w[k], xv[e]);
    }
    return a.value();
}

template <typename T, int NGQ, int NGP>
__global__ __launch_bounds__(256) void k_nld4(const NldArgs A, const NldFast F,
                                              T* __restrict__ du, const T* __restrict__ p,
                                              const T* __restrict__ q)
{
    constexpr int NWQ = 4 * NGQ, NWP = 4 * NGP;
    const int64_t ng = (A.n + 3) / 4;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < ng; g += (int64_t)gridDim.x * 256) {
        const int64_t i0 = 4 * g;
        double v[4];
        const bool fast = i0 >= F.ilo && i0 + 4 <= F.ihi && i0 < F.qend && i0 < F.pend;
        if (fast) {
            double xq[NWQ], xp[NWP];
#pragma unroll
            for (int m = 0; m < NGQ; ++m) {
                T t[4];
                Quad<T>::ld(q + i0 + F.qlo + 4 * m, t);
#pragma unroll
                for (int e = 0; e < 4; ++e) xq[4 * m + e] = (double)t[e];
            }
#pragma unroll
            for (int m = 0; m < NGP; ++m) {
                T t[4];
                Quad<T>::ld(p + i0 + F.plo + 4 * m, t);
#pragma unroll
                for (int e = 0; e < 4; ++e) xp[4 * m + e] = (double)t[e];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double acc = fast_term<NWQ>(F.q[0], F.qlo, i, xq) * fast_term<NWP>(F.p[0], F.plo, i, xp);
                for (int l = 1; l < A.nterms - 1; ++l)
                    acc = acc + (A.c[l] * fast_term<NWQ>(F.q[l], F.qlo, i, xq)) *
                                    fast_term<NWP>(F.p[l], F.plo, i, xp);
                if (A.nterms > 1) {
                    const int l = A.nterms - 1;
                    acc = acc + fast_term<NWQ>(F.q[l], F.qlo, i, xq) * fast_term<NWP>(F.p[l], F.plo, i, xp);
                }
                v[i] = acc;
            }
            Quad<T>::st(du + i0, v);
        } else {
            for (int i = 0; i < 4; ++i) {
                const int64_t j = i0 + i;
                if (j >= A.n) break;
                double acc = term<T>(A.q[0], q, j) * term<T>(A.p[0], p, j);
                for (int l = 1; l < A.nterms - 1; ++l)
                    acc = acc + (A.c[l] * term<T>(A.q[l], q, j)) * term<T>(A.p[l], p, j);
                if (A.nterms > 1) {
                    const int l = A.nterms - 1;
                    acc = acc + term<T>(A.q[l], q, j) * term<T>(A.p[l], p, j);
                }
                Store<T>::st(du, j, acc);
            }
        }
    }
}

// The fast kernel's description of the terms, or false when a term's operator has per-row
// weights, more than kNldTaps taps, or was built for the generic kernel, or the windows are
// wider than the widest instantiation.
bool nld_fast_args(const dop_plan* const* dq, const dop_plan* const* dp, int nterms, int64_t n,
                   NldFast* F, int* ngq, int* ngp)
{
    *F = NldFast{};
    F->ilo = 0;
    F->ihi = n;
    int lo[2] = {INT32_MAX, INT32_MAX}, hi[2] = {INT32_MIN, INT32_MIN};
    for (int l = 0; l < nterms; ++l) {
        for (int side = 0; side < 2; ++side) {
            const dop_plan* P = side ? dp[l] : dq[l];
            NldTerm& t = side ? F->p[l] : F->q[l];
            if (!P) {
                t.on = 0;
                lo[side] = std::min(lo[side], 1);
                hi[side] = std::max(hi[side], 2);
                continue;
            }
            const PlanDev& H = P->d.host;
            const AxisDev& X = H.ax[0];
            if (P->d.generic_only || !X.active || !X.uniform || X.K < 1 || X.K > kNldTaps ||
                (int64_t)P->H.axis[0].w64.size() < X.K || X.rhi <= X.rlo)
                return false;
            t.on = 1;
            t.off = (int)H.tshift[0] + X.dmin + 1;
            t.K = X.K;
            for (int k = 0; k < X.K; ++k) t.w[k] = P->H.axis[0].w64[k];
            F->ilo = std::max<int64_t>(F->ilo, X.rlo);
            F->ihi = std::min<int64_t>(F->ihi, X.rhi);
            lo[side] = std::min(lo[side], t.off);
            hi[side] = std::max(hi[side], t.off + X.K);
        }
    }
    // a group's window: from i0 + lo to i0 + 3 + hi - 1, read as whole groups of four
    int ng[2];
    for (int side = 0; side < 2; ++side) {
        if (lo[side] < 0) return false;
        ng[side] = (hi[side] - lo[side] + 3 + 3) / 4;
        if (ng[side] > 5) return false;
    }
    F->qlo = lo[0];
    F->plo = lo[1];
    // the last element a group reads is i0 + lo + 4*ng - 1 <= n + 1
    F->qend = n + 2 - (lo[0] + 4 * ng[0]) + 1;
    F->pend = n + 2 - (lo[1] + 4 * ng[1]) + 1;
    *ngq = ng[0];
    *ngp = ng[1];
    return true;
}

template <typename T>
hipError_t launch_nld4(const NldArgs& A, const NldFast& F, int ngq, int ngp, void* du,
                       const void* p, const void* q, hipStream_t st)
{
    const int64_t groups = (A.n + 3) / 4;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((groups + 255) / 256, 8192));
#define DOP_NLD4(GQ, GP)                                                                     \
    hipLaunchKernelGGL((k_nld4<T, GQ, GP>), dim3(g), dim3(256), 0, st, A, F, (T*)du,           \
                       (const T*)p, (const T*)q)
    const int key = ngq * 8 + ngp;
    switch (key) {
    case 1 * 8 + 1: DOP_NLD4(1, 1); break;
    case 2 * 8 + 1: DOP_NLD4(2, 1); break;
    case 1 * 8 + 2: DOP_NLD4(1, 2); break;
    case 2 * 8 + 2: DOP_NLD4(2, 2); break;
    case 3 * 8 + 1: DOP_NLD4(3, 1); break;
    case 3 * 8 + 2: DOP_NLD4(3, 2); break;
    case 3 * 8 + 3: DOP_NLD4(3, 3); break;
    case 4 * 8 + 1: DOP_NLD4(4, 1); break;
    case 4 * 8 + 2: DOP_NLD4(4, 2); break;
    case 4 * 8 + 3: DOP_NLD4(4, 3); break;
    case 4 * 8 + 4: DOP_NLD4(4, 4); break;
    case 5 * 8 + 1: DOP_NLD4(5, 1); break;
    case 5 * 8 + 2: DOP_NLD4(5, 2); break;
    case 5 * 8 + 3: DOP_NLD4(5, 3); break;
    case 5 * 8 + 4: DOP_NLD4(5, 4); break;
    case 5 * 8 + 5: DOP_NLD4(5, 5); break;
    default: return hipErrorInvalidValue;
    }
#undef DOP_NLD4
    return hipGetLastError();
}

}  // namespace
}  // namespace dop

extern "C" int dop_nonlinear_diffusion(const dop_plan* const* dq, const dop_plan* const* dp,
                                       int nterms, void* du, const void* p, const void* q,
                                       void* stream)
{
    using namespace dop;
    if (!dq || !dp || !du || !p || !q || nterms < 1 || nterms > kNldMax)
        return set_error(DOP_EINVAL, "nonlinear_diffusion: bad arguments (1..9 terms)");
    NldArgs A{};
    A.nterms = nterms;
    int64_t n = -1;
    int storage = -1;
    int64_t c = 1;
    const int s = nterms - 1;
    for (int l = 0; l < nterms; ++l) {
        for (const dop_plan* P : {dq[l], dp[l]}) {
            if (!P) continue;
            const PlanDev& H = P->d.host;
            if (H.ndim != 1 || H.ghost || P->in_elems != P->out_elems + 2)
                return set_error(DOP_EINVAL, "nonlinear_diffusion: every operator must be a 1-D "
                                             "operator on a padded vector");
            if (n >= 0 && (H.total != n || H.storage != storage))
                return set_error(DOP_EDIM, "nonlinear_diffusion: operators of different lengths "
                                           "or storage types");
            n = H.total;
            storage = H.storage;
        }
        A.q[l] = dq[l] ? dq[l]->dev : nullptr;
        A.p[l] = dp[l] ? dp[l]->dev : nullptr;
        A.c[l] = (double)c;
        c = c * (s - l) / (l + 1);   // C(s, l+1)
    }
    if (n < 0) return set_error(DOP_EINVAL, "nonlinear_diffusion: no operator given");
    if (storage != DOP_F64 && storage != DOP_F32)
        return set_error(DOP_EUNSUPPORTED, "nonlinear_diffusion: f64 or f32 storage");
    A.n = n;
    const hipStream_t st = (hipStream_t)stream;
    NldFast F;
    int ngq = 0, ngp = 0;
    if (nld_fast_args(dq, dp, nterms, n, &F, &ngq, &ngp)) {
        const hipError_t e = storage == DOP_F64 ? launch_nld4<double>(A, F, ngq, ngp, du, p, q, st)
                                                : launch_nld4<float>(A, F, ngq, ngp, du, p, q, st);
        if (e != hipErrorInvalidValue) {
            if (e != hipSuccess)
                return set_error(DOP_EHIP, std::string("nonlinear_diffusion: ") + hipGetErrorString(e));
            return DOP_OK;
        }
    }
    const unsigned g = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    if (storage == DOP_F64)
        hipLaunchKernelGGL((k_nld<double>), dim3(g > 0 ? g : 1), dim3(256), 0, st, A, (double*)du,
                           (const double*)p, (const double*)q);
    else
        hipLaunchKernelGGL((k_nld<float>), dim3(g > 0 ? g : 1), dim3(256), 0, st, A, (float*)du,
                           (const float*)p, (const float*)q);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DOP_EHIP, std::string("nonlinear_diffusion: ") + hipGetErrorString(e));
    return DOP_OK;
}
