// dop_server.cpp -- continuous-batching method-of-lines service (include/dop.h, serving).
//
// One host thread runs the service loop on one stream:
//   release   requests whose release time has passed join the FIFO pending queue
//   admit     the head of the queue is admitted while its state (u + 3 RK4 stage buffers)
//             fits in the device arena and fewer than max_active requests are resident;
//             its initial field is copied in asynchronously
//   step      every resident request advances one RK4 step: four persistent launches,
//             each over all resident requests (the work list is rebuilt only when the
//             resident set changes)
//   retire    a request whose last step is enqueued gets its state copied back
//             asynchronously; when that copy's event completes it is done and its arena
//             blocks return to the free list
// At most kMaxInflight steps are queued ahead of the GPU, so completions are observed
// within a few step times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "dop_objects.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxInflight = 2;
constexpr int kTableSlots = kMaxInflight + 2;
constexpr size_t kAlign = 256;

// First-fit allocator over one device allocation, with coalescing of freed blocks.
class Arena {
public:
    void init(char* base, size_t size)
    {
        base_ = base;
        free_.clear();
        free_[0] = size;
    }
    bool alloc(size_t n, size_t* off)
    {
        n = (n + kAlign - 1) / kAlign * kAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n;
                const size_t at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += n;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = (n + kAlign - 1) / kAlign * kAlign;
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* ptr(size_t off) const { return base_ + off; }
    size_t used() const { return used_; }

private:
    char* base_ = nullptr;
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

enum State { kScheduled, kPending, kActive, kDraining, kDone };

struct Request {
    const dop_plan* plan = nullptr;
    double scale = 0, dt = 0;
    int nsteps = 0, done_steps = 0;
    const void* u0 = nullptr;
    void* result = nullptr;
    double release = 0, admitted = -1, completed = -1;
    size_t off = 0, bytes = 0;
    hipEvent_t done_ev = nullptr;
    State state = kScheduled;
};

struct TableSlot {
    dop::BProb* h_probs = nullptr;   // pinned staging
    dop::BItem* h_items = nullptr;
    dop::BProb* d_probs = nullptr;
    dop::BItem* d_items = nullptr;
    size_t cap_probs = 0, cap_items = 0;
    std::vector<dop::BGroup> groups;  // launch ranges of the work list in this slot
    hipEvent_t used = nullptr;       // last step that read this slot
    bool armed = false;
};

}  // namespace

struct dop_server {
    int64_t budget = 0;
    int max_active = 0;
    int chunk = 16;
    int storage = -1;
    hipStream_t stream = nullptr;
    char* arena_base = nullptr;
    Arena arena;
    std::vector<Request> reqs;
    TableSlot slots[kTableSlots];
    int grid = 2048;
    int peak_active = 0;
    int64_t peak_bytes = 0;
    int64_t iterations = 0;
    ~dop_server()
    {
        for (auto& r : reqs)
            if (r.done_ev) (void)hipEventDestroy(r.done_ev);
        for (auto& s : slots) {
            if (s.h_probs) (void)hipHostFree(s.h_probs);
            if (s.h_items) (void)hipHostFree(s.h_items);
            if (s.d_probs) (void)hipFree(s.d_probs);
            if (s.d_items) (void)hipFree(s.d_items);
            if (s.used) (void)hipEventDestroy(s.used);
        }
        if (arena_base) (void)hipFree(arena_base);
    }
};

namespace {

size_t storage_bytes(int storage) { return storage == DOP_F64 ? 8 : storage == DOP_F32 ? 4 : 2; }

int ensure_slot(TableSlot& s, size_t np, size_t ni)
{
    if (np > s.cap_probs) {
        if (s.h_probs) (void)hipHostFree(s.h_probs);
        if (s.d_probs) (void)hipFree(s.d_probs);
        s.cap_probs = std::max(np, 2 * s.cap_probs);
        if (hipHostMalloc((void**)&s.h_probs, sizeof(dop::BProb) * s.cap_probs) != hipSuccess ||
            hipMalloc((void**)&s.d_probs, sizeof(dop::BProb) * s.cap_probs) != hipSuccess)
            return dop::set_error(DOP_ENOMEM, "server: work-list allocation failed");
    }
    if (ni > s.cap_items) {
        if (s.h_items) (void)hipHostFree(s.h_items);
        if (s.d_items) (void)hipFree(s.d_items);
        s.cap_items = std::max(ni, 2 * s.cap_items);
        if (hipHostMalloc((void**)&s.h_items, sizeof(dop::BItem) * s.cap_items) != hipSuccess ||
            hipMalloc((void**)&s.d_items, sizeof(dop::BItem) * s.cap_items) != hipSuccess)
            return dop::set_error(DOP_ENOMEM, "server: work-list allocation failed");
    }
    return DOP_OK;
}

}  // namespace

extern "C" {

int dop_server_create(int64_t hbm_budget_bytes, int max_active, int chunk, void* stream,
                      dop_server** out)
{
    if (!out || hbm_budget_bytes <= 0 || max_active < 1)
        return dop::set_error(DOP_EINVAL, "server: bad budget or slot count");
    *out = nullptr;
    dop_server* s = new dop_server();
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    s->chunk = chunk > 0 ? chunk : 16;
    s->stream = (hipStream_t)stream;
    if (hipMalloc((void**)&s->arena_base, (size_t)hbm_budget_bytes) != hipSuccess) {
        delete s;
        return dop::set_error(DOP_ENOMEM, "server: arena allocation failed");
    }
    s->arena.init(s->arena_base, (size_t)hbm_budget_bytes);
    for (auto& sl : s->slots)
        if (hipEventCreateWithFlags(&sl.used, hipEventDisableTiming) != hipSuccess) {
            delete s;
            return dop::set_error(DOP_EHIP, "server: event creation failed");
        }
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    s->grid = (cus > 0 ? cus : 256) * 8;
    *out = s;
    return DOP_OK;
}

int dop_server_submit(dop_server* srv, const dop_plan* plan, double scale, double dt, int nsteps,
                      const void* u0_host, void* result_host, double release_s, int64_t* id)
{
    if (!srv || !plan || !u0_host || !result_host || nsteps < 1 || !plan->d.host.ghost)
        return dop::set_error(DOP_EINVAL, "server: bad request");
    const int st = plan->d.host.storage;
    if (st != DOP_F32 && st != DOP_F64) return dop::set_error(DOP_EUNSUPPORTED, "server: f32/f64 only");
    if (srv->storage >= 0 && st != srv->storage)
        return dop::set_error(DOP_EINVAL, "server: one storage type per server");
    srv->storage = st;
    Request r;
    r.plan = plan;
    r.scale = scale;
    r.dt = dt;
    r.nsteps = nsteps;
    r.u0 = u0_host;
    r.result = result_host;
    r.release = release_s;
    r.bytes = 4 * ((size_t)plan->d.host.total * storage_bytes(st) + kAlign);
    if ((int64_t)r.bytes > srv->budget)
        return dop::set_error(DOP_EBUDGET, "server: request state exceeds the device budget");
    srv->reqs.push_back(r);
    if (id) *id = (int64_t)srv->reqs.size() - 1;
    return DOP_OK;
}

int dop_server_run(dop_server* srv)
{
    if (!srv) return dop::set_error(DOP_EINVAL, "null server");
    std::vector<size_t> order(srv->reqs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return srv->reqs[a].release < srv->reqs[b].release; });
    for (auto& r : srv->reqs) {
        r.state = kScheduled;
        r.done_steps = 0;
        r.admitted = r.completed = -1;
        if (!r.done_ev && hipEventCreateWithFlags(&r.done_ev, hipEventDisableTiming) != hipSuccess)
            return dop::set_error(DOP_EHIP, "server: event creation failed");
    }
    srv->peak_active = 0;
    srv->peak_bytes = 0;
    srv->iterations = 0;
    hipStream_t st = srv->stream;
    const size_t sb = storage_bytes(srv->storage < 0 ? DOP_F32 : srv->storage);
    std::deque<size_t> pending;
    std::vector<size_t> active, draining;
    std::deque<hipEvent_t> inflight;
    std::vector<hipEvent_t> step_events(kMaxInflight + 1);
    for (auto& e : step_events)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return dop::set_error(DOP_EHIP, "server: event creation failed");
    size_t next_rel = 0, ndone = 0, ev_i = 0;
    int slot = 0, cur_slot = -1;
    bool dirty = true;
    std::vector<dop::BProb> probs;
    std::vector<dop::BItem> items;
    std::vector<dop::BGroup> groups;
    const auto t0 = Clock::now();
    auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    int rc = DOP_OK;
    while (ndone < srv->reqs.size() && rc == DOP_OK) {
        const double t = now();
        // release
        while (next_rel < order.size() && srv->reqs[order[next_rel]].release <= t) {
            srv->reqs[order[next_rel]].state = kPending;
            pending.push_back(order[next_rel++]);
        }
        // admit (FIFO)
        while (!pending.empty() && (int)active.size() < srv->max_active) {
            Request& r = srv->reqs[pending.front()];
            size_t off;
            if (!srv->arena.alloc(r.bytes, &off)) break;
            r.off = off;
            r.admitted = now();
            r.state = kActive;
            const size_t n = (size_t)r.plan->d.host.total * sb;
            if (hipMemcpyAsync(srv->arena.ptr(off), r.u0, n, hipMemcpyHostToDevice, st) != hipSuccess) {
                rc = dop::set_error(DOP_EHIP, "server: initial-field copy failed");
                break;
            }
            active.push_back(pending.front());
            pending.pop_front();
            dirty = true;
        }
        srv->peak_active = std::max(srv->peak_active, (int)active.size());
        srv->peak_bytes = std::max(srv->peak_bytes, (int64_t)srv->arena.used());
        if (rc != DOP_OK) break;
        if (!active.empty()) {
            if (dirty) {
                // rebuild the work list in a table slot the GPU no longer reads
                slot = (cur_slot + 1) % kTableSlots;
                TableSlot& S = srv->slots[slot];
                if (S.armed) (void)hipEventSynchronize(S.used);
                probs.resize(active.size());
                for (size_t k = 0; k < active.size(); ++k) {
                    Request& r = srv->reqs[active[k]];
                    dop::BProb& B = probs[k];
                    B = dop::BProb{};
                    dop::batch_geometry(r.plan->d, srv->chunk, &B);
                    const size_t n = ((size_t)r.plan->d.host.total * sb + kAlign - 1) / kAlign * kAlign;
                    B.u = srv->arena.ptr(r.off);
                    B.s1 = srv->arena.ptr(r.off + n);
                    B.s2 = srv->arena.ptr(r.off + 2 * n);
                    B.acc = srv->arena.ptr(r.off + 3 * n);
                    B.scale = r.scale;
                    B.dt = r.dt;
                }
                dop::batch_items(probs, &items, &groups);
                if ((rc = ensure_slot(S, probs.size(), items.size())) != DOP_OK) break;
                std::copy(probs.begin(), probs.end(), S.h_probs);
                std::copy(items.begin(), items.end(), S.h_items);
                if (hipMemcpyAsync(S.d_probs, S.h_probs, sizeof(dop::BProb) * probs.size(),
                                   hipMemcpyHostToDevice, st) != hipSuccess ||
                    hipMemcpyAsync(S.d_items, S.h_items, sizeof(dop::BItem) * items.size(),
                                   hipMemcpyHostToDevice, st) != hipSuccess) {
                    rc = dop::set_error(DOP_EHIP, "server: work-list upload failed");
                    break;
                }
                cur_slot = slot;
                S.groups = groups;
                dirty = false;
            }
            TableSlot& S = srv->slots[cur_slot];
            for (int stage = 0; stage < 4 && rc == DOP_OK; ++stage)
                if (dop::launch_batch_items(srv->storage, S.d_probs, S.d_items, S.groups, stage, 0,
                                            srv->grid, st) != hipSuccess)
                    rc = dop::set_error(DOP_EHIP, "server: stage launch failed");
            if (rc != DOP_OK) break;
            (void)hipEventRecord(S.used, st);
            S.armed = true;
            ++srv->iterations;
            // retire requests whose last step is now enqueued
            for (size_t k = 0; k < active.size();) {
                Request& r = srv->reqs[active[k]];
                if (++r.done_steps >= r.nsteps) {
                    const size_t n = (size_t)r.plan->d.host.total * sb;
                    if (hipMemcpyAsync(r.result, srv->arena.ptr(r.off), n, hipMemcpyDeviceToHost,
                                       st) != hipSuccess) {
                        rc = dop::set_error(DOP_EHIP, "server: result copy failed");
                        break;
                    }
                    (void)hipEventRecord(r.done_ev, st);
                    r.state = kDraining;
                    draining.push_back(active[k]);
                    active.erase(active.begin() + k);
                    dirty = true;
                } else {
                    ++k;
                }
            }
            // bound the queue ahead of the GPU
            hipEvent_t e = step_events[ev_i];
            ev_i = (ev_i + 1) % step_events.size();
            (void)hipEventRecord(e, st);
            inflight.push_back(e);
            if ((int)inflight.size() > kMaxInflight) {
                (void)hipEventSynchronize(inflight.front());
                inflight.pop_front();
            }
        } else if (draining.empty()) {
            // idle until the next release
            if (next_rel < order.size()) {
                const double wait = srv->reqs[order[next_rel]].release - now();
                if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
            }
        } else {
            std::this_thread::yield();
        }
        // completions
        for (size_t k = 0; k < draining.size();) {
            Request& r = srv->reqs[draining[k]];
            if (hipEventQuery(r.done_ev) == hipSuccess) {
                r.completed = now();
                r.state = kDone;
                srv->arena.release(r.off, r.bytes);
                draining.erase(draining.begin() + k);
                ++ndone;
            } else {
                ++k;
            }
        }
    }
    (void)hipStreamSynchronize(st);
    for (auto& e : step_events) (void)hipEventDestroy(e);
    return rc;
}

int dop_server_times(const dop_server* srv, int64_t id, double* release, double* admitted,
                     double* completed)
{
    if (!srv || id < 0 || id >= (int64_t)srv->reqs.size())
        return dop::set_error(DOP_EINVAL, "server: unknown request");
    const Request& r = srv->reqs[id];
    if (release) *release = r.release;
    if (admitted) *admitted = r.admitted;
    if (completed) *completed = r.completed;
    return DOP_OK;
}

int dop_server_stats(const dop_server* srv, int* peak_active, int64_t* peak_bytes,
                     int64_t* iterations)
{
    if (!srv) return dop::set_error(DOP_EINVAL, "null server");
    if (peak_active) *peak_active = srv->peak_active;
    if (peak_bytes) *peak_bytes = srv->peak_bytes;
    if (iterations) *iterations = srv->iterations;
    return DOP_OK;
}

void dop_server_destroy(dop_server* srv) { delete srv; }

}  // extern "C"
