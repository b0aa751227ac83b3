// dop_server.cpp -- continuous-batching method-of-lines service (include/dop.h, serving).
//
// One host thread runs the service loop; any thread may submit, also while it runs:
//   release   requests whose release time has passed join the FIFO pending queue
//   admit     the head of the queue is admitted while its state (u + 3 RK4 stage buffers)
//             fits in the device arena and fewer than max_active requests are resident;
//             its initial field is copied in asynchronously
//   step      resident requests are split into size classes by grid points; each class has
//             its own stream (the library's side stream of that index, so the classes and the
//             caller's stream fill HIP's four hardware queues without sharing) and advances its requests one RK4 step per iteration (four
//             grouped launches over the class, the work list rebuilt only when the class
//             changes).  The classes run beside each other on the GPU, so a small request
//             steps at the pace of its own class instead of the largest resident grid's.
//   retire    a request whose last step is enqueued gets its state copied back
//             asynchronously; when that copy's event completes it is done and its arena
//             blocks return to the free list
// At most kMaxInflight steps per class are queued ahead of the GPU, so completions are
// observed within a few step times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "dop_objects.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxInflight = 2;
constexpr int kTableSlots = kMaxInflight + 2;
constexpr size_t kAlign = 256;
// size classes by grid points: up to 256^2, up to 1024^2, larger
constexpr int kClasses = 3;
constexpr int64_t kClassPoints[kClasses - 1] = {(int64_t)256 * 256, (int64_t)1024 * 1024};

size_t round_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation, with coalescing of freed blocks.
class Arena {
public:
    void init(char* base, size_t size)
    {
        base_ = base;
        free_.clear();
        free_[0] = size;
        used_ = 0;
    }
    bool alloc(size_t n, size_t* off)
    {
        n = round_up(n);
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
        n = round_up(n);
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

enum State { kScheduled, kPending, kActive, kDraining, kDone, kFailed };

struct Request {
    const dop_plan* plan = nullptr;
    double scale = 0, dt = 0;
    int nsteps = 0, done_steps = 0;
    const void* u0 = nullptr;
    void* result = nullptr;
    double release = 0, admitted = -1, completed = -1;
    size_t off = 0, bytes = 0;
    int cls = 0;
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

// One size class: its stream, resident requests, work-list slots and queued steps.
struct Class {
    hipStream_t st = nullptr;
    std::vector<Request*> active;
    TableSlot slots[kTableSlots];
    int cur_slot = -1;
    bool dirty = true;
    std::deque<hipEvent_t> inflight;
    std::vector<hipEvent_t> step_events;
    size_t ev_i = 0;
};

int class_of(int64_t points)
{
    int c = 0;
    while (c < kClasses - 1 && points > kClassPoints[c]) ++c;
    return c;
}

}  // namespace

struct dop_server {
    int64_t budget = 0;
    int max_active = 0;
    int chunk = 16;
    int storage = -1;
    hipStream_t stream = nullptr;
    char* arena_base = nullptr;
    Arena arena;
    std::mutex mu;                 // guards reqs, storage, open, running, t0
    std::deque<Request> reqs;      // stable references: requests are only appended
    bool open = false;             // run() waits for submissions while open
    bool running = false;
    Clock::time_point t0;
    Class cls[kClasses];
    int grid = 2048;
    int peak_active = 0;
    int64_t peak_bytes = 0;
    int64_t iterations = 0;
    ~dop_server()
    {
        for (auto& r : reqs)
            if (r.done_ev) (void)hipEventDestroy(r.done_ev);
        for (auto& c : cls) {
            for (auto& s : c.slots) {
                if (s.h_probs) (void)hipHostFree(s.h_probs);
                if (s.h_items) (void)hipHostFree(s.h_items);
                if (s.d_probs) (void)hipFree(s.d_probs);
                if (s.d_items) (void)hipFree(s.d_items);
                if (s.used) (void)hipEventDestroy(s.used);
            }
            for (auto& e : c.step_events) (void)hipEventDestroy(e);
        }
        if (arena_base) (void)hipFree(arena_base);
    }
};

namespace {

size_t storage_bytes(int storage) { return storage == DOP_F64 ? 8 : storage == DOP_F32 ? 4 : 2; }

// device bytes of a request's state: u and three stage buffers, each rounded to kAlign
size_t state_bytes(const dop_plan* plan, int storage)
{
    return 4 * round_up((size_t)plan->d.host.total * storage_bytes(storage));
}

int ensure_slot(TableSlot& s, size_t np, size_t ni)
{
    if (np > s.cap_probs) {
        if (s.h_probs) (void)hipHostFree(s.h_probs);
        if (s.d_probs) (void)hipFree(s.d_probs);
        s.h_probs = nullptr;
        s.d_probs = nullptr;
        s.cap_probs = std::max(np, 2 * s.cap_probs);
        if (hipHostMalloc((void**)&s.h_probs, sizeof(dop::BProb) * s.cap_probs) != hipSuccess ||
            hipMalloc((void**)&s.d_probs, sizeof(dop::BProb) * s.cap_probs) != hipSuccess)
            return dop::set_error(DOP_ENOMEM, "server: work-list allocation failed");
    }
    if (ni > s.cap_items) {
        if (s.h_items) (void)hipHostFree(s.h_items);
        if (s.d_items) (void)hipFree(s.d_items);
        s.h_items = nullptr;
        s.d_items = nullptr;
        s.cap_items = std::max(ni, 2 * s.cap_items);
        if (hipHostMalloc((void**)&s.h_items, sizeof(dop::BItem) * s.cap_items) != hipSuccess ||
            hipMalloc((void**)&s.d_items, sizeof(dop::BItem) * s.cap_items) != hipSuccess)
            return dop::set_error(DOP_ENOMEM, "server: work-list allocation failed");
    }
    return DOP_OK;
}

// One RK4 step of every resident request of class C on its stream (the work list rebuilt
// first if the class changed); requests whose last step this is start their copy back.
int class_step(dop_server* srv, Class& C, size_t sb, std::vector<Request*>* draining)
{
    if (C.dirty) {
        // rebuild the work list in a table slot the GPU no longer reads
        const int slot = (C.cur_slot + 1) % kTableSlots;
        TableSlot& S = C.slots[slot];
        if (S.armed) (void)hipEventSynchronize(S.used);
        std::vector<dop::BProb> probs(C.active.size());
        std::vector<dop::BItem> items;
        std::vector<dop::BGroup> groups;
        for (size_t k = 0; k < C.active.size(); ++k) {
            Request& r = *C.active[k];
            dop::BProb& B = probs[k];
            B = dop::BProb{};
            dop::batch_geometry(r.plan->d, srv->chunk, &B);
            const size_t n = round_up((size_t)r.plan->d.host.total * sb);
            B.u = srv->arena.ptr(r.off);
            B.s1 = srv->arena.ptr(r.off + n);
            B.s2 = srv->arena.ptr(r.off + 2 * n);
            B.acc = srv->arena.ptr(r.off + 3 * n);
            B.scale = r.scale;
            B.dt = r.dt;
        }
        dop::batch_items(probs, &items, &groups);
        int rc = ensure_slot(S, probs.size(), items.size());
        if (rc != DOP_OK) return rc;
        std::copy(probs.begin(), probs.end(), S.h_probs);
        std::copy(items.begin(), items.end(), S.h_items);
        if (hipMemcpyAsync(S.d_probs, S.h_probs, sizeof(dop::BProb) * probs.size(),
                           hipMemcpyHostToDevice, C.st) != hipSuccess ||
            hipMemcpyAsync(S.d_items, S.h_items, sizeof(dop::BItem) * items.size(),
                           hipMemcpyHostToDevice, C.st) != hipSuccess)
            return dop::set_error(DOP_EHIP, "server: work-list upload failed");
        C.cur_slot = slot;
        S.groups = groups;
        C.dirty = false;
    }
    TableSlot& S = C.slots[C.cur_slot];
    // one block per work item (no persistent blocks): blocks retire all the time, so the
    // classes' streams share the CUs instead of the largest class holding them for a whole step
    for (int stage = 0; stage < 4; ++stage)
        if (dop::launch_batch_items(srv->storage, S.d_probs, S.d_items, S.groups, stage, 0,
                                    INT32_MAX, C.st) != hipSuccess)
            return dop::set_error(DOP_EHIP, "server: stage launch failed");
    (void)hipEventRecord(S.used, C.st);
    S.armed = true;
    for (size_t k = 0; k < C.active.size();) {
        Request& r = *C.active[k];
        if (++r.done_steps >= r.nsteps) {
            const size_t n = (size_t)r.plan->d.host.total * sb;
            if (hipMemcpyAsync(r.result, srv->arena.ptr(r.off), n, hipMemcpyDeviceToHost, C.st) !=
                hipSuccess)
                return dop::set_error(DOP_EHIP, "server: result copy failed");
            (void)hipEventRecord(r.done_ev, C.st);
            r.state = kDraining;
            draining->push_back(C.active[k]);
            C.active.erase(C.active.begin() + k);
            C.dirty = true;
        } else {
            ++k;
        }
    }
    // bound the queue ahead of the GPU
    hipEvent_t e = C.step_events[C.ev_i];
    C.ev_i = (C.ev_i + 1) % C.step_events.size();
    (void)hipEventRecord(e, C.st);
    C.inflight.push_back(e);
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
    // the arena holds whole kAlign blocks
    s->budget = hbm_budget_bytes / (int64_t)kAlign * (int64_t)kAlign;
    s->max_active = max_active;
    s->chunk = chunk > 0 ? chunk : 16;
    s->stream = (hipStream_t)stream;
    if (s->budget <= 0 || hipMalloc((void**)&s->arena_base, (size_t)s->budget) != hipSuccess) {
        delete s;
        return dop::set_error(DOP_ENOMEM, "server: arena allocation failed");
    }
    s->arena.init(s->arena_base, (size_t)s->budget);
    static_assert(kClasses <= dop::kSideStreams, "a side stream per class");
    for (int k = 0; k < kClasses; ++k) {
        Class& c = s->cls[k];
        bool ok = dop::side_stream(k, &c.st) == hipSuccess;
        for (auto& sl : c.slots)
            ok = ok && hipEventCreateWithFlags(&sl.used, hipEventDisableTiming) == hipSuccess;
        c.step_events.resize(kMaxInflight + 1, nullptr);
        for (auto& e : c.step_events)
            ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            delete s;
            return dop::set_error(DOP_EHIP, "server: stream or event creation failed");
        }
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
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->storage >= 0 && st != srv->storage)
        return dop::set_error(DOP_EINVAL, "server: one storage type per server");
    Request r;
    r.plan = plan;
    r.scale = scale;
    r.dt = dt;
    r.nsteps = nsteps;
    r.u0 = u0_host;
    r.result = result_host;
    r.release = release_s;
    r.bytes = state_bytes(plan, st);
    r.cls = class_of(plan->d.host.total);
    if ((int64_t)r.bytes > srv->budget)
        return dop::set_error(DOP_EBUDGET, "server: request state exceeds the device budget");
    if (hipEventCreateWithFlags(&r.done_ev, hipEventDisableTiming) != hipSuccess)
        return dop::set_error(DOP_EHIP, "server: event creation failed");
    srv->storage = st;
    srv->reqs.push_back(r);
    if (id) *id = (int64_t)srv->reqs.size() - 1;
    return DOP_OK;
}

int dop_server_set_open(dop_server* srv, int open)
{
    if (!srv) return dop::set_error(DOP_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->open = open != 0;
    return DOP_OK;
}

int dop_server_run(dop_server* srv)
{
    if (!srv) return dop::set_error(DOP_EINVAL, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->running) return dop::set_error(DOP_EINVAL, "server: already running");
        srv->running = true;
        for (auto& r : srv->reqs) {
            if (r.state == kDone) continue;
            r.state = kScheduled;
            r.done_steps = 0;
            r.admitted = r.completed = -1;
        }
        srv->t0 = Clock::now();
    }
    srv->peak_active = 0;
    srv->peak_bytes = 0;
    srv->iterations = 0;
    const auto t0 = srv->t0;
    auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    // Requests are appended by submitters (under the mutex); the deque keeps references to its
    // elements valid, so the loop takes a pointer to each new request under the mutex and then
    // works through the pointers alone.
    std::deque<Request*> pending;
    std::vector<Request*> draining;
    std::vector<Request*> scheduled;   // known requests not yet released, by release time
    size_t known = 0;                  // requests taken from srv->reqs so far
    size_t ndone = 0;
    int rc = DOP_OK;
    int nactive = 0;
    for (;;) {
        const double t = now();
        size_t nreq;
        bool open;
        int storage;
        std::vector<Request*> fresh;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            nreq = srv->reqs.size();
            open = srv->open;
            storage = srv->storage;
            for (; known < nreq; ++known) fresh.push_back(&srv->reqs[known]);
        }
        for (Request* r : fresh) {
            if (r->state == kDone) ++ndone;
            else scheduled.push_back(r);
        }
        if (!fresh.empty())
            std::stable_sort(scheduled.begin(), scheduled.end(),
                             [](const Request* a, const Request* b) { return a->release < b->release; });
        // release
        size_t nrel = 0;
        while (nrel < scheduled.size() && scheduled[nrel]->release <= t) {
            scheduled[nrel]->state = kPending;
            pending.push_back(scheduled[nrel++]);
        }
        scheduled.erase(scheduled.begin(), scheduled.begin() + nrel);
        if (ndone == nreq && !open) break;
        const size_t sb = storage_bytes(storage < 0 ? DOP_F32 : storage);
        // admit (FIFO)
        while (!pending.empty() && nactive < srv->max_active) {
            Request& r = *pending.front();
            size_t off;
            if (!srv->arena.alloc(r.bytes, &off)) {
                if (nactive == 0 && draining.empty())
                    rc = dop::set_error(DOP_EBUDGET, "server: a request does not fit the empty arena");
                break;
            }
            r.off = off;
            r.admitted = now();
            r.state = kActive;
            Class& C = srv->cls[r.cls];
            const size_t n = (size_t)r.plan->d.host.total * sb;
            C.active.push_back(&r);
            C.dirty = true;
            ++nactive;
            pending.pop_front();
            if (hipMemcpyAsync(srv->arena.ptr(off), r.u0, n, hipMemcpyHostToDevice, C.st) != hipSuccess) {
                rc = dop::set_error(DOP_EHIP, "server: initial-field copy failed");
                break;
            }
        }
        srv->peak_active = std::max(srv->peak_active, nactive);
        srv->peak_bytes = std::max(srv->peak_bytes, (int64_t)srv->arena.used());
        if (rc != DOP_OK) break;
        // step every class whose queue ahead of the GPU has room
        bool stepped = false;
        for (auto& C : srv->cls) {
            while (!C.inflight.empty() && hipEventQuery(C.inflight.front()) == hipSuccess)
                C.inflight.pop_front();
            if (C.active.empty() || (int)C.inflight.size() >= kMaxInflight) continue;
            const size_t before = draining.size();
            if ((rc = class_step(srv, C, sb, &draining)) != DOP_OK) break;
            nactive -= (int)(draining.size() - before);
            ++srv->iterations;
            stepped = true;
        }
        if (rc != DOP_OK) break;
        // completions
        for (size_t k = 0; k < draining.size();) {
            Request& r = *draining[k];
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
        if (!stepped) {
            if (nactive == 0 && draining.empty() && pending.empty()) {
                // idle until the next release or submission
                double wait = 2e-4;
                if (!scheduled.empty()) wait = std::min(wait, scheduled.front()->release - now());
                if (wait > 2e-5) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            } else {
                std::this_thread::yield();
            }
        }
    }
    for (auto& C : srv->cls) (void)hipStreamSynchronize(C.st);
    if (rc != DOP_OK) {
        // give back the arena blocks of everything resident; those requests failed
        for (auto& C : srv->cls) {
            for (Request* r : C.active) {
                srv->arena.release(r->off, r->bytes);
                r->state = kFailed;
            }
            C.active.clear();
            C.dirty = true;
        }
        for (Request* r : draining) {
            srv->arena.release(r->off, r->bytes);
            r->state = kFailed;
        }
    }
    for (auto& C : srv->cls) C.inflight.clear();
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->running = false;
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

double dop_server_now(const dop_server* srv)
{
    if (!srv) return -1.0;
    dop_server* s = const_cast<dop_server*>(srv);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->running) return -1.0;
    return std::chrono::duration<double>(Clock::now() - s->t0).count();
}

void dop_server_destroy(dop_server* srv) { delete srv; }

}  // extern "C"
