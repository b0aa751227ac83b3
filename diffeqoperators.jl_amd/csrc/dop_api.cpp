// dop_api.cpp -- the C ABI of libdop.so (include/dop.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dop.h"
#include "dop_build.hpp"
#include "dop_objects.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define DOP_HIP(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(DOP_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));         \
    } while (0)

int device_grid(int per_cu)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256 * per_cu;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return cus * per_cu;
}

}  // namespace

namespace {

template <typename E>
int upload(dop_plan* p, const std::vector<E>& v, const E** out)
{
    *out = nullptr;
    if (v.empty()) return DOP_OK;
    void* d = nullptr;
    const size_t bytes = v.size() * sizeof(E);
    if (hipMalloc(&d, bytes) != hipSuccess) return fail(DOP_ENOMEM, "hipMalloc of plan tables failed");
    p->bufs.push_back(d);
    DOP_HIP(hipMemcpy(d, v.data(), bytes, hipMemcpyHostToDevice));
    *out = (const E*)d;
    return DOP_OK;
}

size_t storage_bytes(int storage)
{
    return storage == DOP_F64 ? 8 : storage == DOP_F32 ? 4 : 2;
}

}  // namespace

namespace dop {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace dop

namespace {

// ---- tuning ----
// Process-wide choices of the tile autotuner, keyed by what decides the candidates and their
// speed: device, storage, accumulation, grid kind (uniform or not), window, extents.
std::mutex g_tune_mu;
std::map<std::string, int> g_tune_cache;

std::string tune_key(const dop_plan* p)
{
    const dop::PlanDev& P = p->d.host;
    const dop::FastPlan& F = p->d.fast;
    std::string k = std::to_string(p->device) + ":" + std::to_string(P.storage) + ":" +
                    std::to_string(P.accum) + ":" + std::to_string(F.kind) + ":" +
                    std::to_string(F.args.K) + ":" + std::to_string(F.VX) + ":" +
                    std::to_string(dop::apply_variants(p->d));
    for (int a = 0; a < dop::kMaxDim; ++a)
        k += ":" + std::to_string(P.n[a]) + "/" + std::to_string(F.args.lo[a]) + "/" +
             std::to_string(F.args.hi[a]) + "/" + std::to_string(F.args.uniform[a]);
    return k;
}

bool tune_cache_get(const dop_plan* p, int* v)
{
    if (dop::apply_variants(p->d) <= 1) return false;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_tune_cache.find(tune_key(p));
    if (it == g_tune_cache.end()) return false;
    *v = it->second;
    return true;
}

constexpr int64_t kAutoTuneMinPoints = (int64_t)1 << 20;

// Tune on the first apply: plans with a tiled path of at least 2^20 points, when the stream
// is not being captured and DOP_AUTOTUNE is not 0.
bool auto_tune_wanted(const dop_plan* p, hipStream_t s)
{
    if (dop::apply_variants(p->d) <= 1 || p->d.host.total < kAutoTuneMinPoints) return false;
    const char* e = std::getenv("DOP_AUTOTUNE");
    if (e && e[0] == '0') return false;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return false;
    return true;
}

// Times every candidate on (du, u) -- du is overwritten, u only read -- and keeps the fastest in
// the plan and the cache.  The caller holds p->tune_mu.
int tune_variants(dop_plan* p, void* du, const void* u, hipStream_t s, int* choice)
{
    const int nv = dop::apply_variants(p->d);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return fail(DOP_EHIP, "autotune: event creation failed");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(DOP_EHIP, "autotune: event creation failed");
    }
    // steady state, as an application sees it: back-to-back launches of one candidate between
    // two events; the candidates take turns over three rounds (drift in clocks or
    // temperature hits them alike) and each keeps its fastest round
    const char* vb = std::getenv("DOP_AUTOTUNE_VERBOSE");
    const bool verbose = vb && vb[0] == '1';
    constexpr int kRounds = 3;
    std::vector<float> per(nv, 1e30f);
    std::vector<char> okv(nv, 1);
    for (int v = 0; v < nv; ++v)   // warm-up (code objects, caches)
        okv[v] = dop::launch_apply(p->d, du, u, 0, v, 0, s) == hipSuccess;
    // launches per measurement: enough for about 1 ms of back-to-back work (8 to 64), so that
    // short kernels are timed in their steady state rather than by their first launches
    int kReps = 8;
    {
        (void)hipEventRecord(e0, s);
        for (int rep = 0; rep < 4; ++rep) (void)dop::launch_apply(p->d, du, u, 0, 0, 0, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const float t = ms / 4;
        if (t > 0.0f) kReps = std::max(8, std::min(64, (int)(1.0f / t) + 1));
    }
    for (int round = 0; round < kRounds; ++round) {
        for (int v = 0; v < nv; ++v) {
            if (!okv[v]) continue;
            (void)hipEventRecord(e0, s);
            for (int rep = 0; rep < kReps && okv[v]; ++rep)
                okv[v] = dop::launch_apply(p->d, du, u, 0, v, 0, s) == hipSuccess;
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            per[v] = std::min(per[v], ms / kReps);
        }
    }
    float best = 1e30f;
    int best_v = 0;
    for (int v = 0; v < nv; ++v) {
        if (verbose)
            fprintf(stderr, "dop_plan_autotune: variant %d %s %.4f ms\n", v, okv[v] ? "ok" : "failed",
                    per[v]);
        if (okv[v] && per[v] < best) {
            best = per[v];
            best_v = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGetLastError();   // a candidate that failed to launch is skipped, not reported
    p->d.variant = best_v;
    p->tuned.store(1, std::memory_order_release);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[tune_key(p)] = best_v;
    }
    if (choice) *choice = best_v;
    return DOP_OK;
}

}  // namespace

extern "C" {

int dop_version(void) { return DOP_ABI_VERSION; }

const char* dop_last_error(void) { return g_err.c_str(); }

int dop_plan_create(const dop_desc* desc, dop_plan** out)
{
    if (!desc || !out) return fail(DOP_EINVAL, "null argument");
    *out = nullptr;
    dop::HostPlan Hb;
    std::string err;
    const int rc = dop::build_host_plan(*desc, &Hb, &err);
    if (rc != DOP_OK) return fail(rc, err);
    dop_plan* p = new dop_plan();
    p->H = std::move(Hb);
    const dop::HostPlan& H = p->H;
    p->d.host = H.meta;
    p->in_elems = H.in_elems;
    p->out_elems = H.out_elems;
    (void)hipGetDevice(&p->device);
    int r = DOP_OK;
    // DOP_FORCE_GENERIC=1 (read when a plan is created) keeps every point on the generic
    // kernel: the reference for checking the tiled and single-axis kernels bit for bit.
    const char* fg = std::getenv("DOP_FORCE_GENERIC");
    p->d.generic_only = (fg && fg[0] == '1') ? 1 : 0;
    if (!p->d.generic_only) dop::build_fast(H, &p->d.fast);
    // DOP_TILE_VARIANT=v pins the tile-shape candidate; otherwise the tuning cache's choice
    // for this shape, or 0 until the plan is tuned
    if (const char* tv = std::getenv("DOP_TILE_VARIANT")) {
        const int v = std::atoi(tv);
        if (v >= 0 && v < dop::apply_variants(p->d)) p->d.variant = v;
        p->tuned = 1;
    } else {
        int v = -1;
        if (tune_cache_get(p, &v)) {
            p->d.variant = v;
            p->tuned = 1;
        }
    }
    for (int a = 0; a < dop::kMaxDim && r == DOP_OK && p->d.fast.kind != 0; ++a) {
        if ((r = upload(p, p->d.fast.w64[a], &p->d.fast.args.w64[a])) != DOP_OK) break;
        if ((r = upload(p, p->d.fast.w32[a], &p->d.fast.args.w32[a])) != DOP_OK) break;
    }
    for (int a = 0; a < dop::kMaxDim && r == DOP_OK; ++a) {
        dop::AxisDev& X = p->d.host.ax[a];
        const dop::AxisTables& T = H.axis[a];
        if ((r = upload(p, T.w64, &X.w64)) != DOP_OK) break;
        if ((r = upload(p, T.w32, &X.w32)) != DOP_OK) break;
        if ((r = upload(p, T.bJ0, &X.bJ0)) != DOP_OK) break;
        if ((r = upload(p, T.bn, &X.bn)) != DOP_OK) break;
        if ((r = upload(p, T.bw, &X.bw)) != DOP_OK) break;
        if ((r = upload(p, T.lbc, &X.lbc)) != DOP_OK) break;
    }
    if (r == DOP_OK) {
        if (hipMalloc((void**)&p->dev, sizeof(dop::PlanDev)) != hipSuccess) {
            r = fail(DOP_ENOMEM, "hipMalloc of plan failed");
        } else if (hipMemcpy(p->dev, &p->d.host, sizeof(dop::PlanDev), hipMemcpyHostToDevice) !=
                   hipSuccess) {
            r = fail(DOP_EHIP, "upload of plan failed");
        }
        p->d.dev = p->dev;
    }
    if (r != DOP_OK) {
        delete p;
        return r;
    }
    *out = p;
    return DOP_OK;
}

void dop_plan_destroy(dop_plan* plan) { delete plan; }

int64_t dop_plan_input_elems(const dop_plan* plan) { return plan ? plan->in_elems : -1; }
int64_t dop_plan_output_elems(const dop_plan* plan) { return plan ? plan->out_elems : -1; }

int dop_apply(const dop_plan* plan, void* du, const void* u, int accumulate, void* stream)
{
    if (!plan || !du || !u) return fail(DOP_EINVAL, "null plan or buffer");
    if (du == u) return fail(DOP_EINVAL, "du must not alias u");
    const hipStream_t s = (hipStream_t)stream;
    if (!plan->tuned.load(std::memory_order_acquire) && !accumulate && auto_tune_wanted(plan, s)) {
        // the plan's tuning state is not part of what it computes (every candidate gives the
        // same values), so a const plan may record its choice
        dop_plan* mp = const_cast<dop_plan*>(plan);
        std::lock_guard<std::mutex> lk(mp->tune_mu);
        if (!mp->tuned.load(std::memory_order_acquire)) {
            const int rc = tune_variants(mp, du, u, s, nullptr);
            if (rc != DOP_OK) return rc;
        }
    }
    DOP_HIP(dop::launch_apply(plan->d, du, u, accumulate ? 1 : 0, -1, 0, s));
    return DOP_OK;
}

int dop_plan_variant(const dop_plan* plan, int* tuned)
{
    if (!plan) return fail(DOP_EINVAL, "null plan");
    if (tuned) *tuned = plan->tuned.load();
    return plan->d.variant.load();
}

int dop_plan_update(dop_plan* plan, const dop_desc* desc, void* stream)
{
    if (!plan || !desc) return fail(DOP_EINVAL, "null plan or descriptor");
    dop::HostPlan H;
    std::string err;
    int rc = dop::build_host_plan(*desc, &H, &err);
    if (rc != DOP_OK) return fail(rc, err);
    dop::FastPlan F;
    if (!plan->d.generic_only) dop::build_fast(H, &F);
    const dop::FastPlan& F0 = plan->d.fast;
    bool same = dop::same_structure(plan->H.meta, H.meta) && F.kind == F0.kind &&
                F.args.K == F0.args.K && F.VX == F0.VX && F.shell_total == F0.shell_total;
    for (int a = 0; a < dop::kMaxDim && same; ++a) {
        const dop::AxisTables &A0 = plan->H.axis[a], &A1 = H.axis[a];
        same = A0.w64.size() == A1.w64.size() && A0.w32.size() == A1.w32.size() &&
               A0.bJ0.size() == A1.bJ0.size() && A0.bn.size() == A1.bn.size() &&
               A0.bw.size() == A1.bw.size() && A0.lbc.size() == A1.lbc.size() &&
               F.w64[a].size() == F0.w64[a].size() && F.w32[a].size() == F0.w32[a].size() &&
               F.args.lo[a] == F0.args.lo[a] && F.args.hi[a] == F0.args.hi[a];
    }
    if (!same)
        return fail(DOP_EINVAL, "dop_plan_update: the descriptor's structure differs from the "
                                "plan's (create a new plan)");
    const hipStream_t s = (hipStream_t)stream;
    // tables whose contents changed: stream-ordered copies, waited for before the host copies
    // they come from are kept
    bool copied = false;
    auto put = [&](const void* dst, const auto& oldv, const auto& newv) -> int {
        const size_t bytes = newv.size() * sizeof(newv[0]);
        if (bytes == 0 || !dst || std::memcmp(oldv.data(), newv.data(), bytes) == 0) return DOP_OK;
        DOP_HIP(hipMemcpyAsync((void*)dst, newv.data(), bytes, hipMemcpyHostToDevice, s));
        copied = true;
        return DOP_OK;
    };
    for (int a = 0; a < dop::kMaxDim; ++a) {
        const dop::AxisDev& X = plan->d.host.ax[a];
        const dop::AxisTables &A0 = plan->H.axis[a], &A1 = H.axis[a];
        if ((rc = put(X.w64, A0.w64, A1.w64)) != DOP_OK) return rc;
        if ((rc = put(X.w32, A0.w32, A1.w32)) != DOP_OK) return rc;
        if ((rc = put(X.bJ0, A0.bJ0, A1.bJ0)) != DOP_OK) return rc;
        if ((rc = put(X.bn, A0.bn, A1.bn)) != DOP_OK) return rc;
        if ((rc = put(X.bw, A0.bw, A1.bw)) != DOP_OK) return rc;
        if ((rc = put(X.lbc, A0.lbc, A1.lbc)) != DOP_OK) return rc;
        if (F0.kind != 0) {
            if ((rc = put(F0.args.w64[a], F0.w64[a], F.w64[a])) != DOP_OK) return rc;
            if ((rc = put(F0.args.w32[a], F0.w32[a], F.w32[a])) != DOP_OK) return rc;
        }
    }
    if (copied) DOP_HIP(hipStreamSynchronize(s));
    // the plan's own values (BC taps and constants among them): a kernel argument, so the
    // update is one stream-ordered launch
    dop::PlanDev v = H.meta;
    for (int a = 0; a < dop::kMaxDim; ++a) {
        const dop::AxisDev& X = plan->d.host.ax[a];
        v.ax[a].w64 = X.w64;
        v.ax[a].w32 = X.w32;
        v.ax[a].bJ0 = X.bJ0;
        v.ax[a].bn = X.bn;
        v.ax[a].bw = X.bw;
        v.ax[a].lbc = X.lbc;
    }
    DOP_HIP(dop::launch_set_plan(plan->dev, v, s));
    plan->d.host = v;
    for (int a = 0; a < dop::kMaxDim; ++a) {
        plan->d.fast.w64[a] = std::move(F.w64[a]);
        plan->d.fast.w32[a] = std::move(F.w32[a]);
    }
    plan->H = std::move(H);
    return DOP_OK;
}

int dop_plan_autotune(dop_plan* plan, void* stream, int* choice)
{
    if (!plan) return fail(DOP_EINVAL, "null plan");
    const int nv = dop::apply_variants(plan->d);
    std::lock_guard<std::mutex> lk(plan->tune_mu);
    if (nv <= 1) {
        plan->d.variant = 0;
        plan->tuned = 1;
        if (choice) *choice = 0;
        return DOP_OK;
    }
    const size_t sb = storage_bytes(plan->d.host.storage);
    void *du = nullptr, *u = nullptr;
    if (hipMalloc(&du, plan->out_elems * sb) != hipSuccess) return fail(DOP_ENOMEM, "autotune alloc");
    if (hipMalloc(&u, plan->in_elems * sb) != hipSuccess) {
        (void)hipFree(du);
        return fail(DOP_ENOMEM, "autotune alloc");
    }
    hipStream_t s = (hipStream_t)stream;
    (void)hipMemsetAsync(u, 0, plan->in_elems * sb, s);
    const int rc = tune_variants(plan, du, u, s, choice);
    (void)hipStreamSynchronize(s);
    (void)hipFree(du);
    (void)hipFree(u);
    return rc;
}

}  // extern "C"

namespace {

// (re)builds the work list for row chunk `chunk` and uploads probs + items
int batch_upload(dop_batch* b, int chunk)
{
    for (auto& P : b->probs) P.chunk = chunk;
    dop::batch_items(b->probs, &b->items, &b->groups);
    b->chunk = chunk;
    b->nitems = (int)b->items.size();
    if (!b->d_probs) {
        if (hipMalloc((void**)&b->d_probs, sizeof(dop::BProb) * b->probs.size()) != hipSuccess)
            return fail(DOP_ENOMEM, "hipMalloc of batch problem table failed");
    }
    if (b->d_items) (void)hipFree(b->d_items);
    b->d_items = nullptr;
    if (hipMalloc((void**)&b->d_items, sizeof(dop::BItem) * std::max<size_t>(b->items.size(), 1)) !=
        hipSuccess)
        return fail(DOP_ENOMEM, "hipMalloc of batch work list failed");
    DOP_HIP(hipMemcpy(b->d_probs, b->probs.data(), sizeof(dop::BProb) * b->probs.size(),
                      hipMemcpyHostToDevice));
    DOP_HIP(hipMemcpy(b->d_items, b->items.data(), sizeof(dop::BItem) * b->items.size(),
                      hipMemcpyHostToDevice));
    return DOP_OK;
}

int batch_init(dop_batch* b, int n, const dop_plan* const* plans)
{
    b->storage = plans[0]->d.host.storage;
    b->grid = device_grid(8);
    b->probs.resize(n);
    for (int i = 0; i < n; ++i) {
        if (plans[i]->d.host.storage != b->storage)
            return fail(DOP_EINVAL, "mixed storage types in batch");
        dop::BProb& P = b->probs[i];
        P = dop::BProb{};
        dop::batch_geometry(plans[i]->d, b->chunk, &P);
    }
    return DOP_OK;
}

hipError_t batch_launch(const dop_batch* b, int stage, int accumulate, hipStream_t s)
{
    return dop::launch_batch_items(b->storage, b->d_probs, b->d_items, b->groups, stage,
                                   accumulate, b->grid, s);
}

// times `stage` launches for each candidate chunk height and keeps the fastest
int batch_tune(dop_batch* b, int stage, hipStream_t s, int* choice)
{
    static const int cands[4] = {16, 32, 8, 64};
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const char* vb = std::getenv("DOP_AUTOTUNE_VERBOSE");
    const bool verbose = vb && vb[0] == '1';
    float best = 1e30f;
    int best_c = 0, r = DOP_OK;
    for (int c = 0; c < 4 && r == DOP_OK; ++c) {
        if ((r = batch_upload(b, cands[c])) != DOP_OK) break;
        float tsum = 0.0f;
        for (int rep = 0; rep < 4; ++rep) {
            (void)hipEventRecord(e0, s);
            if (batch_launch(b, stage, 0, s) != hipSuccess) {
                r = fail(DOP_EHIP, "batched launch failed during tuning");
                break;
            }
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) tsum += ms;
        }
        if (verbose) fprintf(stderr, "batch autotune: chunk %d %.4f ms\n", cands[c], tsum / 3.0f);
        if (tsum < best) {
            best = tsum;
            best_c = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r != DOP_OK) return r;
    if (choice) *choice = cands[best_c];
    return batch_upload(b, cands[best_c]);
}

}  // namespace

extern "C" {

int dop_batch_create(int n, const dop_plan* const* plans, void* const* du, const void* const* u,
                     dop_batch** out)
{
    if (!out || n < 1 || !plans || !du || !u) return fail(DOP_EINVAL, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!plans[i] || !du[i] || !u[i]) return fail(DOP_EINVAL, "null plan or buffer in batch");
        if (du[i] == u[i]) return fail(DOP_EINVAL, "du must not alias u");
    }
    dop_batch* b = new dop_batch();
    int r = batch_init(b, n, plans);
    for (int i = 0; i < n && r == DOP_OK; ++i) {
        b->probs[i].src = u[i];
        b->probs[i].out = du[i];
    }
    if (r == DOP_OK) r = batch_upload(b, b->chunk);
    if (r != DOP_OK) {
        delete b;
        return r;
    }
    *out = b;
    return DOP_OK;
}

int dop_batch_apply(const dop_batch* batch, int accumulate, void* stream)
{
    if (!batch) return fail(DOP_EINVAL, "null batch");
    DOP_HIP(batch_launch(batch, -1, accumulate ? 1 : 0, (hipStream_t)stream));
    return DOP_OK;
}

int dop_batch_autotune(dop_batch* batch, void* stream, int* choice)
{
    if (!batch) return fail(DOP_EINVAL, "null batch");
    return batch_tune(batch, -1, (hipStream_t)stream, choice);
}

void dop_batch_destroy(dop_batch* batch) { delete batch; }

namespace {
// One stage of a batch captured with its groups on ns forked streams (fresh events per stage:
// each fork and join becomes its own graph edge).
int capture_stage_forked(const dop_batch* b, int stage, hipStream_t cs, hipStream_t* ws, int ns,
                         std::vector<hipEvent_t>* evs)
{
    const size_t first = evs->size();
    for (int i = 0; i <= ns; ++i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail(DOP_EHIP, "rk4: event creation failed");
        evs->push_back(e);
    }
    if (dop::launch_batch_items_forked(b->storage, b->d_probs, b->d_items, b->groups, stage, 0,
                                       b->grid, cs, ws, evs->data() + first, ns) != hipSuccess)
        return fail(DOP_EHIP, "rk4: stage launch during capture failed");
    return DOP_OK;
}
}  // namespace

int dop_rk4_create(int n, const dop_plan* const* plans, const double* scale, const double* dt,
                   void* const* u, void* stream, dop_rk4** out)
{
    if (!out || n < 1 || !plans || !scale || !dt || !u) return fail(DOP_EINVAL, "bad rk4 arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i)
        if (!plans[i] || !u[i] || !plans[i]->d.host.ghost)
            return fail(DOP_EINVAL, "rk4: every problem needs a Q*u plan and a state buffer");
    const int storage = plans[0]->d.host.storage;
    if (storage != DOP_F64 && storage != DOP_F32)
        return fail(DOP_EUNSUPPORTED, "rk4 needs f32 or f64 storage");
    dop_rk4* rk = new dop_rk4();
    int r = batch_init(&rk->b, n, plans);
    const size_t sb = storage_bytes(storage);
    for (int i = 0; i < n && r == DOP_OK; ++i) {
        dop::BProb& P = rk->b.probs[i];
        P.u = u[i];
        P.scale = scale[i];
        P.dt = dt[i];
        void* bufs[3] = {nullptr, nullptr, nullptr};
        for (int k = 0; k < 3 && r == DOP_OK; ++k) {
            if (hipMalloc(&bufs[k], plans[i]->d.host.total * sb) != hipSuccess)
                r = fail(DOP_ENOMEM, "rk4: stage buffer allocation failed");
            else
                rk->bufs.push_back(bufs[k]);
        }
        P.s1 = bufs[0];
        P.s2 = bufs[1];
        P.acc = bufs[2];
    }
    hipStream_t s = (hipStream_t)stream;
    // tile shape: stage 0 only reads u and writes acc / s1, so it can be timed repeatedly
    const char* at = std::getenv("DOP_RK4_AUTOTUNE");
    if (r == DOP_OK) {
        if (at && at[0] == '0') r = batch_upload(&rk->b, rk->b.chunk);
        else r = batch_tune(&rk->b, 0, s, nullptr);
    }
    // capture one step (four stages; each stage's groups on forked streams) into a graph.
    // DOP_RK4_STREAMS sets the number of streams (1: every group in order on one stream).
    hipStream_t cs = nullptr;
    int ns = 4;
    if (const char* e = std::getenv("DOP_RK4_STREAMS")) ns = std::max(1, std::min(8, std::atoi(e)));
    std::vector<hipStream_t> ws(ns, nullptr);
    std::vector<hipEvent_t> evs;
    if (r == DOP_OK && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        r = fail(DOP_EHIP, "rk4: stream creation failed");
    for (int i = 0; i < ns && r == DOP_OK; ++i)
        if (hipStreamCreateWithFlags(&ws[i], hipStreamNonBlocking) != hipSuccess)
            r = fail(DOP_EHIP, "rk4: stream creation failed");
    if (r == DOP_OK) {
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            r = fail(DOP_EHIP, "rk4: begin capture failed");
        } else {
            for (int st = 0; st < 4 && r == DOP_OK; ++st) {
                if (ns == 1) {
                    if (batch_launch(&rk->b, st, 0, cs) != hipSuccess)
                        r = fail(DOP_EHIP, "rk4: stage launch during capture failed");
                } else {
                    r = capture_stage_forked(&rk->b, st, cs, ws.data(), ns, &evs);
                }
            }
            if (hipStreamEndCapture(cs, &rk->graph) != hipSuccess && r == DOP_OK)
                r = fail(DOP_EHIP, "rk4: end capture failed");
            if (r == DOP_OK &&
                hipGraphInstantiate(&rk->exec, rk->graph, nullptr, nullptr, 0) != hipSuccess)
                r = fail(DOP_EHIP, "rk4: graph instantiation failed");
        }
    }
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    for (hipStream_t w : ws)
        if (w) (void)hipStreamDestroy(w);
    if (cs) (void)hipStreamDestroy(cs);
    if (r != DOP_OK) {
        delete rk;
        return r;
    }
    *out = rk;
    return DOP_OK;
}

int dop_rk4_step(dop_rk4* rk, int nsteps, void* stream)
{
    if (!rk || nsteps < 0) return fail(DOP_EINVAL, "bad rk4 step arguments");
    for (int i = 0; i < nsteps; ++i) DOP_HIP(hipGraphLaunch(rk->exec, (hipStream_t)stream));
    return DOP_OK;
}

int dop_rk4_chunk(const dop_rk4* rk) { return rk ? rk->b.chunk : -1; }

void dop_rk4_destroy(dop_rk4* rk) { delete rk; }

}  // extern "C"
