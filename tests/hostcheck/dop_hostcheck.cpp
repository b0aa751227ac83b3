// dop_hostcheck.cpp -- TEST INFRASTRUCTURE.  Evaluates a plan on the CPU with the same
// plan builder (dop_build.cpp, dop_rows.cpp) and the same per-point arithmetic
// (dop_eval.hpp) that the gfx950 kernels run, so that the index math of the device path
// can be compared with the FP64 oracle without a GPU.  Never part of libdop.so.
#include <cstdint>
#include <string>

#include "../../diffeqoperators.jl_amd/csrc/dop_build.hpp"
#include "../../diffeqoperators.jl_amd/csrc/dop_concrete.hpp"
#include "../../diffeqoperators.jl_amd/csrc/dop_eval.hpp"

namespace {
thread_local std::string g_err;

template <typename T, int ACC>
void run(const dop::PlanDev& P, T* du, const T* src, int accumulate)
{
    for (int64_t p = 0; p < P.total; ++p) {
        int64_t idx[3];
        idx[0] = p % P.n[0];
        const int64_t q = p / P.n[0];
        idx[1] = q % P.n[1];
        idx[2] = q / P.n[1];
        double v = dop::eval_point<T, ACC>(P, src, idx);
        if (accumulate) v += (double)dop::Store<T>::ld(du, p);
        dop::Store<T>::st(du, p, v);
    }
}

// the host plan with its table pointers bound to the host copies
dop::PlanDev bind(const dop::HostPlan& H)
{
    dop::PlanDev P = H.meta;
    for (int a = 0; a < 3; ++a) {
        dop::AxisDev& X = P.ax[a];
        const dop::AxisTables& T = H.axis[a];
        X.w64 = T.w64.empty() ? nullptr : T.w64.data();
        X.w32 = T.w32.empty() ? nullptr : T.w32.data();
        X.bJ0 = T.bJ0.empty() ? nullptr : T.bJ0.data();
        X.bn = T.bn.empty() ? nullptr : T.bn.data();
        X.bw = T.bw.empty() ? nullptr : T.bw.data();
        X.lbc = T.lbc.empty() ? nullptr : T.lbc.data();
    }
    return P;
}
}  // namespace

extern "C" {

const char* hc_last_error(void) { return g_err.c_str(); }

// Plan summary for tests: per axis K, dmin, rlo, rhi, uniform, KB, nb (7 ints per axis).
int hc_plan_info(const dop_desc* d, int32_t* info)
{
    dop::HostPlan H;
    const int rc = dop::build_host_plan(*d, &H, &g_err);
    if (rc != DOP_OK) return rc;
    for (int a = 0; a < 3; ++a) {
        const dop::AxisDev& X = H.axis[a].meta;
        int32_t* o = info + 7 * a;
        o[0] = X.K; o[1] = X.dmin; o[2] = X.rlo; o[3] = X.rhi;
        o[4] = X.uniform; o[5] = X.KB; o[6] = X.nb;
    }
    return H.meta.accum;
}

int hc_apply(const dop_desc* d, void* du, const void* u, int accumulate)
{
    dop::HostPlan H;
    const int rc = dop::build_host_plan(*d, &H, &g_err);
    if (rc != DOP_OK) return rc;
    dop::PlanDev P = bind(H);
    using namespace dop;
    switch (P.storage) {
    case kF64: run<double, kAccF64>(P, (double*)du, (const double*)u, accumulate); break;
    case kF32:
        if (P.accum == kAccF64) run<float, kAccF64>(P, (float*)du, (const float*)u, accumulate);
        else if (P.accum == kAccF32Kahan) run<float, kAccF32Kahan>(P, (float*)du, (const float*)u, accumulate);
        else if (P.accum == kAccF32Wide) run<float, kAccF32Wide>(P, (float*)du, (const float*)u, accumulate);
        else run<float, kAccF32>(P, (float*)du, (const float*)u, accumulate);
        break;
    case kBF16:
        if (P.accum == kAccF64) run<bf16, kAccF64>(P, (bf16*)du, (const bf16*)u, accumulate);
        else if (P.accum == kAccF32Kahan) run<bf16, kAccF32Kahan>(P, (bf16*)du, (const bf16*)u, accumulate);
        else if (P.accum == kAccF32Wide) run<bf16, kAccF32Wide>(P, (bf16*)du, (const bf16*)u, accumulate);
        else run<bf16, kAccF32>(P, (bf16*)du, (const bf16*)u, accumulate);
        break;
    }
    return DOP_OK;
}
}

extern "C" {
// sizes of the ABI structs as the C++ compiler lays them out (checked against ctypes)
int64_t hc_sizeof(int which)
{
    switch (which) {
    case 0: return (int64_t)sizeof(dop_operator);
    case 1: return (int64_t)sizeof(dop_affine_bc);
    case 2: return (int64_t)sizeof(dop_desc);
    }
    return -1;
}

// The CSR form of the plan's map (same enumeration and merge as the device export):
// row_ptr[n_rows+1]; col/val may be NULL (count only, *nnz set); b[n_rows] may be NULL.
int hc_csr(const dop_desc* d, int64_t* row_ptr, int64_t* col, double* val, double* b,
           int64_t* nnz)
{
    dop::HostPlan H;
    const int rc = dop::build_host_plan(*d, &H, &g_err);
    if (rc != DOP_OK) return rc;
    const dop::PlanDev P = bind(H);
    if (dop::row_entry_bound(P) > 128) return DOP_EUNSUPPORTED;
    int64_t o = 0;
    row_ptr[0] = 0;
    for (int64_t p = 0; p < P.total; ++p) {
        int64_t idx[3];
        dop::point_index(P, p, idx);
        dop::RowList<128> L;
        dop::row_entries(P, idx, L);
        for (int i = 0; i < L.n; ++i) {
            if (L.val[i] == 0.0) continue;
            if (col) {
                col[o] = L.col[i];
                val[o] = L.val[i];
            }
            ++o;
        }
        row_ptr[p + 1] = o;
        if (b) b[p] = L.b;
    }
    *nnz = o;
    return DOP_OK;
}
}
