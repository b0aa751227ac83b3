// dop_hostcheck.cpp -- TEST INFRASTRUCTURE.  Evaluates a plan on the CPU with the same
// plan builder (dop_build.cpp, dop_rows.cpp) and the same per-point arithmetic
// (dop_eval.hpp) that the gfx950 kernels run, so that the index math of the device path
// can be compared with the FP64 oracle without a GPU.  Never part of libdop.so.
#include <cstdint>
#include <string>

#include "../../diffeqoperators.jl_amd/csrc/dop_build.hpp"
#include "../../diffeqoperators.jl_amd/csrc/dop_concrete.hpp"
#include "../../diffeqoperators.jl_amd/csrc/dop_eval.hpp"
#include "../../diffeqoperators.jl_amd/csrc/dop_fast.hpp"
#include "../../diffeqoperators.jl_amd/csrc/dop_strip.hpp"

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

namespace {
template <typename T, int ACC, int KM>
void run_shell_k(const dop::PlanDev& P, const dop::StripArgs& S, T* du, const T* src, int accumulate)
{
    T stage[dop::strip::kStage + 1];
    dop::strip::Slot<ACC> slots[dop::strip::kSW];
    for (int s = 0; s < S.nseg; ++s) {
        const dop::ShellSeg& G = S.seg[s];
        if (G.kind == 4) {
            // the device marches these x-face points along z; the same points as face items
            for (int l = 0; l < S.fn[0][0] * S.fn[0][1]; ++l)
                dop::strip::eval_face<T, ACC, KM, 0, dop::strip::ScalarLoad<T>>(P, S, G.r0, G.xend, l, stage, slots,
                                                                               du, src, accumulate);
            continue;
        }
        for (int j = 0; j < G.nr; ++j) {
            for (int l = 0; l < G.items; ++l) {
                if (S.vec_ok)
                    dop::strip::eval_item<T, ACC, KM, true, dop::strip::ScalarLoad<T>>(P, S, s, j, l, stage, slots,
                                                                                     du, src, accumulate);
                else
                    dop::strip::eval_item<T, ACC, KM, false, dop::strip::ScalarLoad<T>>(P, S, s, j, l, stage,
                                                                                      slots, du, src, accumulate);
            }
        }
    }
}

// the tiles' window K (3, 5, 7 or 9) as the strips' compile-time width, as the launch does
template <typename T, int ACC>
void run_shell(const dop::PlanDev& P, const dop::StripArgs& S, int K, T* du, const T* src, int accumulate)
{
    switch (K) {
    case 3: run_shell_k<T, ACC, 3>(P, S, du, src, accumulate); break;
    case 5: run_shell_k<T, ACC, 5>(P, S, du, src, accumulate); break;
    case 7: run_shell_k<T, ACC, 7>(P, S, du, src, accumulate); break;
    default: run_shell_k<T, ACC, 9>(P, S, du, src, accumulate); break;
    }
}
}  // namespace

extern "C" {
// The 3-D shell of the tiled path (every point outside the tiles' box) evaluated strip by strip
// with the device kernel's per-strip code; other points of du are left as they are.  Returns the
// number of shell points, or -1 when the plan has no 3-D tiled path.
int64_t hc_shell(const dop_desc* d, void* du, const void* u, int accumulate)
{
    dop::HostPlan H;
    if (dop::build_host_plan(*d, &H, &g_err) != DOP_OK) return -1;
    dop::FastPlan F;
    dop::build_fast(H, &F);
    if (F.kind != 3) return -1;
    const dop::PlanDev P = bind(H);
    using namespace dop;
    switch (P.storage) {
    case kF64: run_shell<double, kAccF64>(P, F.strips, F.args.K, (double*)du, (const double*)u, accumulate); break;
    case kF32:
        if (P.accum == kAccF32Wide) run_shell<float, kAccF32Wide>(P, F.strips, F.args.K, (float*)du, (const float*)u, accumulate);
        else run_shell<float, kAccF32>(P, F.strips, F.args.K, (float*)du, (const float*)u, accumulate);
        break;
    case kBF16:
        if (P.accum == kAccF32Wide) run_shell<bf16, kAccF32Wide>(P, F.strips, F.args.K, (bf16*)du, (const bf16*)u, accumulate);
        else run_shell<bf16, kAccF32>(P, F.strips, F.args.K, (bf16*)du, (const bf16*)u, accumulate);
        break;
    }
    return F.shell_total;
}

// Shell points handled per segment kind of the 3-D tiled plan (kinds 0, 1, 2: faces of that
// axis; 3: general boxes) into pts[4]; -1 when the plan has no 3-D tiled path.
int hc_shell_kinds(const dop_desc* d, int64_t* pts)
{
    dop::HostPlan H;
    if (dop::build_host_plan(*d, &H, &g_err) != DOP_OK) return -1;
    dop::FastPlan F;
    dop::build_fast(H, &F);
    if (F.kind != 3) return -1;
    const dop::StripArgs& S = F.strips;
    for (int k = 0; k < 4; ++k) pts[k] = 0;
    for (int s = 0; s < S.nseg; ++s) {
        const dop::ShellSeg& G = S.seg[s];
        if (G.kind == 3) {
            const int b = G.r0;
            pts[3] += (int64_t)S.ext[b][0] * S.ext[b][1] * S.ext[b][2];
        } else if (G.kind == 0) {
            pts[0] += (int64_t)(G.xend - G.r0) * G.items;
        } else if (G.kind == 4) {
            const int w = G.xend - G.r0 < dop::kStripPts ? G.xend - G.r0 : dop::kStripPts;
            pts[0] += (int64_t)w * S.fn[0][0] * S.fn[0][1];
        } else {
            pts[G.kind] += (int64_t)G.nr * G.items * dop::kStripPts;
        }
    }
    return S.nseg;
}

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
