// dop_nld.hip -- nonlinear_diffusion! (derivative_operator.jl:34-66) on the device: the
// Leibniz expansion of d^s/dx^s (p * d^f q / dx^f) on padded vectors,
//     du = sum_{l=0..s} C(s, l) (D^{l+f} q) .* (D^{s-l} p),   D^0 = the interior 2:n+1,
// every D^k a CenteredDifference plan on the padded vector, all terms of a point evaluated by
// one thread in one launch.
#include <hip/hip_runtime.h>

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
