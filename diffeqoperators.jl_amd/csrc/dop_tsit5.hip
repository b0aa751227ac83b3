// dop_tsit5.hip -- adaptive explicit time stepping of du/dt = scale * (L*Q) u on the device.
//
// The reference's method-of-lines tests hand `(u, p, t) -> A*bc*u` to OrdinaryDiffEq's Tsit5
// (heat_eqn.jl:4-60).  Here the whole integration stays on the device: the Runge-Kutta pair of
// Tsitouras (2011, "Runge-Kutta pairs of order 5(4) satisfying only the first column
// simplifying assumption", Comput. Math. Appl. 62: 770-775; 7 stages, first-same-as-last),
// the scaled RMS error norm sqrt(mean((e_i / (atol + rtol max(|u_i|, |unew_i|)))^2)) as a
// deterministic two-level reduction, and the step-size controller of Hairer, Norsett & Wanner
// (Solving ODEs I, 2nd ed., II.4 and the DOPRI5 code: error exponent 1/5 - 0.75 beta,
// beta = 0.04, safety 0.9, step ratio in [0.2, 10], no growth right after a rejection) run as
// kernels that read and write a device-resident state.  One attempted step -- six stage
// combinations, six fused operator applications, the error reduction and the controller -- is
// one captured hipGraph, replayed until the end time is reached, with no host round trip
// between steps.  Accepted steps are committed (u <- unew, k1 <- k7) by the next attempt's
// first kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/dop.h"
#include "dop_eval.hpp"
#include "dop_objects.hpp"

namespace dop {
namespace {

// Tsitouras 5(4): c, A (lower triangle, row s = stage s+1), b = A[6], btilde = b - bhat
__constant__ double kA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
     -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
     2.324710524099774},
};
__constant__ double kBt[7] = {-0.00178001105222577714, -0.0008164344596567469,
                              0.007880878010261995, -0.1447110071732629,
                              0.5823571654525552, -0.45808210592918697,
                              0.015151515151515152};

constexpr int kTsBlock = 256;
constexpr int kTsParts = 1024;   // blocks of the error reduction (fixed: deterministic sums)

struct TsState {
    double t, dt, t_end, facold;
    double scale, atol, rtol;
    int64_t n;
    int64_t naccept, nreject, nattempt;
    int32_t commit;     // the last attempt was accepted: u <- unew, k1 <- k7 pending
    int32_t done;       // t reached t_end (or the step size collapsed: failed = 1)
    int32_t rejected;   // the last attempt was rejected (no step growth on the next accept)
    int32_t failed;
};

struct TsArrays {
    void* u;
    void* unew;
    void* us;   // stage input
    void* k[7];
};

template <typename T>
__device__ inline double ld(const void* p, int64_t i)
{
    return (double)Store<T>::ld((const T*)p, i);
}
template <typename T>
__device__ inline void st(void* p, int64_t i, double v)
{
    Store<T>::st((T*)p, i, v);
}

// stage S (2..7): input of stage S = u + h sum_{j<S} A[S-1][j] k_j, h = dt*scale; stage 7's
// input is the new solution.  Stage 2 first commits an accepted step.
template <typename T, int S>
__global__ __launch_bounds__(kTsBlock) void k_ts_stage(TsState* __restrict__ st_, TsArrays a)
{
    const TsState& S_ = *st_;
    const int64_t n = S_.n;
    const bool commit = S == 2 && S_.commit;
    if (!commit && S_.done) return;
    const double h = S_.dt * S_.scale;
    void* out = S == 7 ? a.unew : a.us;
    for (int64_t i = (int64_t)blockIdx.x * kTsBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kTsBlock) {
        double u = ld<T>(a.u, i);
        double k0 = ld<T>(a.k[0], i);
        if (commit) {
            u = ld<T>(a.unew, i);
            k0 = ld<T>(a.k[6], i);
            st<T>(a.u, i, u);
            st<T>(a.k[0], i, k0);
        }
        if (S_.done) continue;
        double acc = kA[S - 1][0] * k0;
#pragma unroll
        for (int j = 1; j < S - 1; ++j) acc = fma(kA[S - 1][j], ld<T>(a.k[j], i), acc);
        st<T>(out, i, fma(h, acc, u));
    }
}

// sum over the block's points of (e_i / sc_i)^2, e = h sum_j btilde_j k_j
template <typename T>
__global__ __launch_bounds__(kTsBlock) void k_ts_error(const TsState* __restrict__ st_, TsArrays a,
                                                       double* __restrict__ part)
{
    __shared__ double sh[kTsBlock];
    const TsState& S_ = *st_;
    double s = 0.0;
    if (!S_.done) {
        const double h = S_.dt * S_.scale;
        for (int64_t i = (int64_t)blockIdx.x * kTsBlock + threadIdx.x; i < S_.n;
             i += (int64_t)gridDim.x * kTsBlock) {
            double e = 0.0;
#pragma unroll
            for (int j = 0; j < 7; ++j) e = fma(kBt[j], ld<T>(a.k[j], i), e);
            e *= h;
            const double sc = S_.atol + S_.rtol * fmax(fabs(ld<T>(a.u, i)), fabs(ld<T>(a.unew, i)));
            const double q = e / sc;
            s = fma(q, q, s);
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kTsBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// the controller: one block reduces the partial sums in a fixed order and updates the state
__global__ __launch_bounds__(kTsBlock) void k_ts_control(TsState* __restrict__ st_,
                                                         const double* __restrict__ part)
{
    __shared__ double sh[kTsBlock];
    TsState& S_ = *st_;
    if (S_.done) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < kTsParts; b += kTsBlock) s += part[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kTsBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double err = sqrt(sh[0] / (double)S_.n);
    const double beta = 0.04, safe = 0.9;
    const double fac11 = pow(err, 0.2 - beta * 0.75);
    double dtn;
    S_.nattempt += 1;
    if (err <= 1.0) {
        // accepted
        S_.t += S_.dt;
        S_.naccept += 1;
        S_.commit = 1;
        double fac = fac11 / pow(S_.facold, beta);
        fac = fmax(0.1, fmin(5.0, fac / safe));
        dtn = S_.dt / fac;
        if (S_.rejected) dtn = fmin(dtn, S_.dt);
        S_.facold = fmax(err, 1.0e-4);
        S_.rejected = 0;
    } else {
        // rejected (a NaN error is a rejection: fmin drops the NaN, the step shrinks by 5)
        S_.nreject += 1;
        S_.commit = 0;
        dtn = S_.dt / fmin(5.0, fac11 / safe);
        S_.rejected = 1;
    }
    const double left = S_.t_end - S_.t;
    if (left <= 1.0e-14 * fmax(fabs(S_.t_end), 1.0)) {
        S_.done = 1;
    } else if (!(dtn > 1.0e-14 * fmax(fabs(S_.t), 1.0))) {
        S_.done = 1;
        S_.failed = 1;
    } else {
        S_.dt = fmin(dtn, left);
    }
}

// sum of ((x - c*y) / sc)^2, sc = atol + rtol |u0|  (the starting step size's norms)
template <typename T>
__global__ __launch_bounds__(kTsBlock) void k_ts_sq(const void* x, const void* y, double c,
                                                    const void* u0, double xs, int64_t n,
                                                    double atol, double rtol,
                                                    double* __restrict__ part)
{
    __shared__ double sh[kTsBlock];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kTsBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kTsBlock) {
        const double v = xs * ld<T>(x, i) - (y ? c * ld<T>(y, i) : 0.0);
        const double q = v / (atol + rtol * fabs(ld<T>(u0, i)));
        s = fma(q, q, s);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kTsBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// x = u0 + h * y
template <typename T>
__global__ __launch_bounds__(kTsBlock) void k_ts_axpy(void* x, const void* u0, const void* y,
                                                      double h, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kTsBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kTsBlock)
        st<T>(x, i, fma(h, ld<T>(y, i), ld<T>(u0, i)));
}

int grid_of(int64_t n)
{
    const int64_t g = (n + kTsBlock - 1) / kTsBlock;
    return (int)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

}  // namespace
}  // namespace dop

struct dop_tsit5 {
    const dop_plan* plan = nullptr;
    int storage = 0;
    dop::TsArrays a{};
    dop::TsState* st = nullptr;
    double* part = nullptr;
    std::vector<void*> bufs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~dop_tsit5()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : bufs) (void)hipFree(p);
        if (st) (void)hipFree(st);
        if (part) (void)hipFree(part);
    }
};

namespace {
int tfail(int code, const std::string& m) { return dop::set_error(code, m); }

template <typename T>
hipError_t ts_attempt(dop_tsit5* I, hipStream_t s)
{
    using namespace dop;
    const int64_t n = I->plan->d.host.total;
    const dim3 g(grid_of(n)), b(kTsBlock);
    hipError_t e;
#define TS_STAGE(SS)                                                                            \
    hipLaunchKernelGGL((k_ts_stage<T, SS>), g, b, 0, s, I->st, I->a);                           \
    if ((e = hipGetLastError()) != hipSuccess) return e;                                        \
    if ((e = launch_apply(I->plan->d, SS == 7 ? I->a.k[6] : I->a.k[SS - 1],                     \
                          SS == 7 ? I->a.unew : I->a.us, 0, -1, 0, s)) != hipSuccess)           \
        return e;
    TS_STAGE(2)
    TS_STAGE(3)
    TS_STAGE(4)
    TS_STAGE(5)
    TS_STAGE(6)
    TS_STAGE(7)
#undef TS_STAGE
    hipLaunchKernelGGL((k_ts_error<T>), dim3(kTsParts), b, 0, s, I->st, I->a, I->part);
    hipLaunchKernelGGL(k_ts_control, dim3(1), b, 0, s, I->st, I->part);
    return hipGetLastError();
}

template <typename T>
double ts_norm2(dop_tsit5* I, const void* x, const void* y, double c, double xs, double atol,
                double rtol, hipStream_t s, int* rc)
{
    using namespace dop;
    const int64_t n = I->plan->d.host.total;
    hipLaunchKernelGGL((k_ts_sq<T>), dim3(kTsParts), dim3(kTsBlock), 0, s, x, y, c, I->a.u, xs, n,
                       atol, rtol, I->part);
    std::vector<double> h(kTsParts);
    if (hipMemcpyAsync(h.data(), I->part, sizeof(double) * kTsParts, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        *rc = tfail(DOP_EHIP, "tsit5: norm reduction failed");
        return 0.0;
    }
    double t = 0.0;
    for (double v : h) t += v;
    return std::sqrt(t / (double)n);
}

template <typename T>
int ts_setup(dop_tsit5* I, double scale, double t0, double t1, double dt0, double atol, double rtol,
             hipStream_t s)
{
    using namespace dop;
    const int64_t n = I->plan->d.host.total;
    int rc = DOP_OK;
    // k1 = L Q u0 (unscaled; f = scale * k).  Without dt0: the starting step size of Hairer,
    // Norsett & Wanner (Solving ODEs I, II.4) from the norms of u0, f(u0) and f(u0 + h0 f(u0)).
    // Through dop_apply: an untuned plan of at least 2^20 points picks its tile variant here
    // (on k1's buffer and u0, before the attempt is captured), as on its first apply elsewhere.
    if ((rc = dop_apply(I->plan, I->a.k[0], I->a.u, 0, s)) != DOP_OK) return rc;
    double dt = dt0;
    if (!(dt > 0.0)) {
        const double d0 = ts_norm2<T>(I, I->a.u, nullptr, 0.0, 1.0, atol, rtol, s, &rc);
        const double d1 = std::fabs(scale) * ts_norm2<T>(I, I->a.k[0], nullptr, 0.0, 1.0, atol, rtol, s, &rc);
        if (rc != DOP_OK) return rc;
        double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
        h0 = std::fmin(h0, t1 - t0);
        // u1 = u0 + h0 f0 in the stage buffer, f1 = f(u1) in k2
        hipLaunchKernelGGL((k_ts_axpy<T>), dim3(grid_of(n)), dim3(kTsBlock), 0, s, I->a.us, I->a.u,
                           I->a.k[0], h0 * scale, n);
        if (launch_apply(I->plan->d, I->a.k[1], I->a.us, 0, -1, 0, s) != hipSuccess)
            return tfail(DOP_EHIP, "tsit5: initial apply failed");
        const double d2 =
            std::fabs(scale) * ts_norm2<T>(I, I->a.k[1], I->a.k[0], 1.0, 1.0, atol, rtol, s, &rc) / h0;
        if (rc != DOP_OK) return rc;
        const double m = std::fmax(d1, d2);
        const double h1 = m <= 1e-15 ? std::fmax(1e-6, h0 * 1e-3) : std::pow(0.01 / m, 1.0 / 5.0);
        dt = std::fmin(100.0 * h0, h1);
    }
    TsState H{};
    H.t = t0;
    H.dt = std::fmin(dt, t1 - t0);
    H.t_end = t1;
    H.facold = 1.0e-4;
    H.scale = scale;
    H.atol = atol;
    H.rtol = rtol;
    H.n = n;
    H.done = t1 - t0 <= 0.0;
    if (hipMemcpyAsync(I->st, &H, sizeof(H), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return tfail(DOP_EHIP, "tsit5: state upload failed");
    return DOP_OK;
}
}  // namespace

extern "C" {

int dop_tsit5_create(const dop_plan* plan, double scale, void* u, double t0, double t1, double dt0,
                     double abstol, double reltol, void* stream, dop_tsit5** out)
{
    if (!plan || !u || !out) return tfail(DOP_EINVAL, "null plan, u or out");
    *out = nullptr;
    const int storage = plan->d.host.storage;
    if (!plan->d.host.ghost || plan->in_elems != plan->out_elems)
        return tfail(DOP_EINVAL, "tsit5 needs a square L*Q plan (u and du of one size)");
    if (storage != DOP_F64 && storage != DOP_F32)
        return tfail(DOP_EUNSUPPORTED, "tsit5 needs f32 or f64 storage");
    if (!(abstol > 0.0) || !(reltol >= 0.0) || !(t1 >= t0))
        return tfail(DOP_EINVAL, "tsit5: need abstol > 0, reltol >= 0, t1 >= t0");
    const hipStream_t s = (hipStream_t)stream;
    dop_tsit5* I = new dop_tsit5();
    I->plan = plan;
    I->storage = storage;
    I->a.u = u;
    const size_t bytes = (size_t)plan->d.host.total * (storage == DOP_F64 ? 8 : 4);
    int rc = DOP_OK;
    void** slots[9] = {&I->a.unew, &I->a.us, &I->a.k[0], &I->a.k[1], &I->a.k[2],
                       &I->a.k[3],  &I->a.k[4], &I->a.k[5], &I->a.k[6]};
    for (void** p : slots) {
        if (hipMalloc(p, bytes) != hipSuccess) {
            rc = tfail(DOP_ENOMEM, "tsit5: buffer allocation failed");
            break;
        }
        I->bufs.push_back(*p);
    }
    if (rc == DOP_OK && (hipMalloc((void**)&I->st, sizeof(dop::TsState)) != hipSuccess ||
                         hipMalloc((void**)&I->part, sizeof(double) * dop::kTsParts) != hipSuccess))
        rc = tfail(DOP_ENOMEM, "tsit5: state allocation failed");
    if (rc == DOP_OK)
        rc = storage == DOP_F64 ? ts_setup<double>(I, scale, t0, t1, dt0, abstol, reltol, s)
                                : ts_setup<float>(I, scale, t0, t1, dt0, abstol, reltol, s);
    // capture one attempt
    hipStream_t cs = nullptr;
    if (rc == DOP_OK && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        rc = tfail(DOP_EHIP, "tsit5: stream creation failed");
    if (rc == DOP_OK) {
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = tfail(DOP_EHIP, "tsit5: begin capture failed");
        } else {
            const hipError_t e =
                storage == DOP_F64 ? ts_attempt<double>(I, cs) : ts_attempt<float>(I, cs);
            if (e != hipSuccess) rc = tfail(DOP_EHIP, std::string("tsit5: capture: ") + hipGetErrorString(e));
            if (hipStreamEndCapture(cs, &I->graph) != hipSuccess && rc == DOP_OK)
                rc = tfail(DOP_EHIP, "tsit5: end capture failed");
            if (rc == DOP_OK && hipGraphInstantiate(&I->exec, I->graph, nullptr, nullptr, 0) != hipSuccess)
                rc = tfail(DOP_EHIP, "tsit5: graph instantiation failed");
        }
    }
    if (cs) (void)hipStreamDestroy(cs);
    if (rc != DOP_OK) {
        delete I;
        return rc;
    }
    *out = I;
    return DOP_OK;
}

int dop_tsit5_set_end(dop_tsit5* I, double t1, void* stream)
{
    if (!I) return tfail(DOP_EINVAL, "null integrator");
    const hipStream_t s = (hipStream_t)stream;
    dop::TsState H;
    if (hipMemcpyAsync(&H, I->st, sizeof(H), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return tfail(DOP_EHIP, "tsit5: state read failed");
    if (!(t1 >= H.t)) return tfail(DOP_EINVAL, "tsit5: the new end time lies before t");
    if (H.failed) return tfail(DOP_EINVAL, "tsit5: the integration failed (step size underflow)");
    H.t_end = t1;
    H.done = t1 - H.t <= 1.0e-14 * std::fmax(std::fabs(t1), 1.0);
    if (!H.done) H.dt = std::fmin(H.dt, t1 - H.t);
    if (hipMemcpyAsync(I->st, &H, sizeof(H), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return tfail(DOP_EHIP, "tsit5: state write failed");
    return DOP_OK;
}

int dop_tsit5_run(dop_tsit5* I, int64_t max_attempts, int check_every, void* stream)
{
    if (!I || max_attempts < 0) return tfail(DOP_EINVAL, "null integrator or negative count");
    const hipStream_t s = (hipStream_t)stream;
    if (check_every < 1) check_every = 64;
    int64_t left = max_attempts;
    dop::TsState H;
    if (hipMemcpyAsync(&H, I->st, sizeof(H), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return tfail(DOP_EHIP, "tsit5: state read failed");
    // replays between two reads of the state: the attempts the current step size still needs
    // (plus one), at most `batch`, which doubles from 1 up to check_every -- the operator applies
    // of a replay run even after the end time is reached, so launching far past it would be
    // wasted work, and a step size that grows by up to 10x per step makes the first estimates
    // far too large
    int64_t batch = 1;
    while (left > 0 && !H.done) {
        const double need = H.dt > 0.0 ? std::ceil((H.t_end - H.t) / H.dt) + 1.0 : 1.0;
        int64_t k = need < (double)batch ? (int64_t)need : batch;
        if (k < 1) k = 1;
        batch = batch * 2 < check_every ? batch * 2 : check_every;
        if (k > left) k = left;
        for (int64_t i = 0; i < k; ++i)
            if (hipGraphLaunch(I->exec, s) != hipSuccess) return tfail(DOP_EHIP, "tsit5: graph launch failed");
        left -= k;
        if (hipMemcpyAsync(&H, I->st, sizeof(H), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return tfail(DOP_EHIP, "tsit5: state read failed");
    }
    // commit an accepted last step (the commit is idempotent)
    const int64_t n = I->plan->d.host.total;
    if (I->storage == DOP_F64)
        hipLaunchKernelGGL((dop::k_ts_stage<double, 2>), dim3(dop::grid_of(n)), dim3(dop::kTsBlock), 0,
                           s, I->st, I->a);
    else
        hipLaunchKernelGGL((dop::k_ts_stage<float, 2>), dim3(dop::grid_of(n)), dim3(dop::kTsBlock), 0,
                           s, I->st, I->a);
    if (hipGetLastError() != hipSuccess) return tfail(DOP_EHIP, "tsit5: commit launch failed");
    return DOP_OK;
}

int dop_tsit5_status(const dop_tsit5* I, double* t, double* dt, int64_t* naccept, int64_t* nreject,
                     int* done, void* stream)
{
    if (!I) return tfail(DOP_EINVAL, "null integrator");
    const hipStream_t s = (hipStream_t)stream;
    dop::TsState H;
    if (hipMemcpyAsync(&H, I->st, sizeof(H), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return tfail(DOP_EHIP, "tsit5: state read failed");
    if (t) *t = H.t;
    if (dt) *dt = H.dt;
    if (naccept) *naccept = H.naccept;
    if (nreject) *nreject = H.nreject;
    if (done) *done = H.failed ? -1 : H.done;
    return DOP_OK;
}

void dop_tsit5_destroy(dop_tsit5* I) { delete I; }

}  // extern "C"
