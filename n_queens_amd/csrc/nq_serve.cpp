// Serving with continuous batching over a persistent kernel (no reference counterpart).
//
// Callers submit counting requests (an nq_instance each) at any time and get a ticket.  Submit
// validates and expands the request into tasks on the caller's thread and admits it while the
// pending work fits the server's task budget (else NQ_ERR_BUSY).  A scheduler thread feeds one
// device-resident ring of tasks that a persistent kernel (serve_kernel, nq_dfs.hip) consumes
// while the host keeps appending to it:
//
//   * it publishes requests in slices (at most batch_work estimated DFS nodes each), least
//     estimated remaining work first, aged by waiting time, and only while the published but
//     unclaimed work stays under a lookahead of two slices, so a request that arrives behind a
//     large one waits for at most that much work (the claim cursor is read from the device when
//     the kernel's progress hint would hold publication back);
//   * a slice is packed into pinned staging and copied into the ring on a copy stream outside
//     the server lock; once the copy has completed the scheduler raises the ring's tail in
//     pinned host memory, which the running kernel reads;
//   * every published slice occupies one of slot_requests slots; the kernel adds every finished
//     task's count to its slot and, when the slice's last task is counted, writes {count,
//     sequence number} into pinned host memory, where the scheduler finds it; the slice's ring
//     space is then free, and a request is done when its last slice is counted;
//   * kernel waves exit after a short idle time without work (every wait in the kernel is
//     bounded), and an instance one of whose waves has exited retires: it claims nothing new and
//     ends when its tasks are counted.  The scheduler launches a new instance, on the next of
//     `slots` streams, whenever published tasks are unclaimed and no instance is running or the
//     latest one retires.  Instances share the ring's claim cursor, so they may overlap.
//
// A request's count (the sum over its tasks) is exactly the one nq_instance_count gives.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nq.h"
#include "nq_internal.h"
#include "nq_kernels.h"

namespace nq {
int append_instance_tasks(int N, int k, uint32_t c, uint32_t l, uint32_t r, int depth_target, uint32_t idx,
                          std::vector<uint4> &out);
bool encode_instance(const nq_instance &in, uint32_t *c, uint32_t *l, uint32_t *r);
}  // namespace nq

using namespace nq;

namespace {

struct Request {
    uint64_t ticket = 0;
    int N = 0;
    std::vector<uint4> tasks;  // tag = twice << 5 | N; the slice's slot is added when published
    size_t next = 0;           // first task not yet published
    int outstanding = 0;       // published slices not yet counted
    double task_work = 1;      // estimated DFS nodes per task
    double t_submit = 0, t_start = -1, t_done = 0;
    uint64_t count = 0;        // sum of the counted slices
    bool done = false;
};

// Estimated DFS nodes below a prefix with m free rows (measured: ~6.3e5 at 14 free rows, a
// factor ~5.7 per additional row).  Only orders and sizes slices; never affects counts.
double est_task_nodes(int m) {
    double w = 6.3e5;
    for (int i = 14; i < m; ++i) w *= 5.7;
    for (int i = m; i < 14; ++i) w /= 5.7;
    return std::max(w, 1.0);
}

// A published slice: ring positions [begin, end) of one request, its estimated work, and the
// slot the kernel counts it in (every slice has its own slot and sequence number, so that its
// ring space is freed as soon as it is counted, whatever the rest of its request does).
struct Slice {
    std::shared_ptr<Request> req;
    uint64_t begin = 0, end = 0;
    double work = 0;
    int slot = -1;
    uint64_t seq = 0;
    bool counted = false;
};

constexpr uint64_t STAGING_TASKS = 1u << 20;  // tasks per ring copy
constexpr double IDLE_US = 200;               // a kernel wave's idle time before it exits
constexpr double CLOSE_DRAIN_MS = 120000;     // close() waits this long for requests in flight

}  // namespace

struct nq_server {
    nq_server_config cfg{};
    int device = 0, cus = 0, wall_khz = 0;
    // Device memory.
    uint4 *ring = nullptr;
    uint64_t ring_cap = 0;  // a power of two
    unsigned long long *head = nullptr, *remaining = nullptr, *count = nullptr, *seq = nullptr, *error = nullptr;
    // Pinned host memory.
    ServeCtl *ctl = nullptr;
    ServeDone *done = nullptr;
    uint4 *staging = nullptr;
    unsigned long long *slot_init = nullptr;  // [remaining, count, seq] for a slot being opened
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_ev = nullptr;
    std::vector<hipStream_t> kstreams;
    std::vector<hipEvent_t> kevents;
    std::vector<bool> kused;
    int block = 0, grid = 0, stack_depth = 0;
    bool apush = false;
    unsigned epoch = 0;   // number of the latest kernel instance
    int next_stream = 0;  // stream of the next instance
    // Scheduling state, under mu.
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::shared_ptr<Request>> pending;  // requests with tasks not yet published
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    std::vector<Slice *> slot_slice;  // slice counted in each slot (null: free)
    std::vector<int> free_slots;
    std::deque<Slice> published;  // ring FIFO, freed from the front as slices are counted
    uint64_t tail = 0;            // tasks published (host copy of ctl->tail)
    uint64_t next_ticket = 1, next_seq = 1, pending_tasks = 0;  // next_seq numbers slices
    int inflight = 0;  // slices published and not yet counted
    double head_checked = -1e30;  // last read of the device claim cursor for a launch (ms)
    double head_read = -1e30;     // last read of it for the lookahead (ms)
    uint64_t claimed_seen = 0;    // largest claim cursor value seen (a lower bound of it)
    bool stop = false;
    double t_stop = 0;  // when close() was called
    int error_code = NQ_OK;
    std::string error_msg;
    std::thread thr;
    // Statistics.
    uint64_t completed = 0, slices = 0, solutions = 0, launches = 0;
    double busy_ms = 0, busy_since = -1, t_open = 0;
};

static void free_server(nq_server *S) {
    if (S->ring) (void)hipFree(S->ring);
    for (unsigned long long *p : {S->head, S->remaining, S->count, S->seq, S->error})
        if (p) (void)hipFree(p);
    if (S->ctl) (void)hipHostFree(S->ctl);
    if (S->done) (void)hipHostFree(S->done);
    if (S->staging) (void)hipHostFree(S->staging);
    if (S->slot_init) (void)hipHostFree(S->slot_init);
    if (S->copy_ev) (void)hipEventDestroy(S->copy_ev);
    if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
    for (hipEvent_t e : S->kevents)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : S->kstreams)
        if (s) (void)hipStreamDestroy(s);
    delete S;
}

static void set_error(nq_server *S, int code, const std::string &msg) {
    if (S->error_code == NQ_OK) {
        S->error_code = code;
        S->error_msg = msg;
    }
}

// Work of the published slices not yet claimed by the kernel, for a claim cursor `claimed`.
static double unclaimed_work(const nq_server *S, uint64_t claimed) {
    double w = 0;
    for (const Slice &sl : S->published) {
        if (sl.end <= claimed) continue;
        const uint64_t from = std::max(sl.begin, claimed);
        w += static_cast<double>(sl.end - from) * sl.req->task_work;
    }
    return w;
}

// The next slice to publish (lock held): the pending request with the least estimated remaining
// work, aged by its waiting time; its tasks up to the slice budget and the ring and staging
// space.  Returns false when nothing can be published now (no request, no free slot, no space).
static bool pick_slice(nq_server *S, double now, Slice *out) {
    if (S->pending.empty() || S->free_slots.empty()) return false;
    const uint64_t front = S->published.empty() ? S->tail : S->published.front().begin;
    const uint64_t space = S->ring_cap - (S->tail - front);
    if (space == 0) return false;
    auto key = [&](const std::shared_ptr<Request> &r) {
        const double rem = static_cast<double>(r->tasks.size() - r->next) * r->task_work;
        return rem / 1e9 - 2.0 * (now - r->t_submit);  // ~1 ms per 1e9 nodes; 2x aging
    };
    std::shared_ptr<Request> best;
    for (auto &r : S->pending)
        if (!best || key(r) < key(best)) best = r;
    const uint64_t left = best->tasks.size() - best->next;
    uint64_t take = static_cast<uint64_t>(std::max(1.0, S->cfg.batch_work / best->task_work));
    take = std::min<uint64_t>({take, left, space, STAGING_TASKS});
    out->req = best;
    out->begin = S->tail;
    out->end = S->tail + take;
    out->work = static_cast<double>(take) * best->task_work;
    return true;
}

// Copy a slice into the ring with its slot's initial state (tasks left = the slice's tasks,
// count 0, the slice's sequence number) and wait for the copies; called without the lock.
static hipError_t copy_slice(nq_server *S, const Slice &sl, size_t first) {
    const Request &r = *sl.req;
    const uint64_t n = sl.end - sl.begin;
    for (uint64_t j = 0; j < n; ++j) {
        uint4 u = r.tasks[first + j];
        u.w = (static_cast<uint32_t>(sl.slot) << 6) | (u.w & 63u);
        S->staging[j] = u;
    }
    S->slot_init[0] = n;
    S->slot_init[1] = 0;
    S->slot_init[2] = sl.seq;
    hipError_t e = hipMemcpyAsync(S->remaining + sl.slot, &S->slot_init[0], 8, hipMemcpyHostToDevice, S->copy_stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(S->count + sl.slot, &S->slot_init[1], 8, hipMemcpyHostToDevice, S->copy_stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(S->seq + sl.slot, &S->slot_init[2], 8, hipMemcpyHostToDevice, S->copy_stream);
    // Ring positions wrap: up to two contiguous pieces.
    const uint64_t pos = sl.begin & (S->ring_cap - 1);
    const uint64_t n1 = std::min(n, S->ring_cap - pos);
    if (e == hipSuccess && n1)
        e = hipMemcpyAsync(S->ring + pos, S->staging, n1 * sizeof(uint4), hipMemcpyHostToDevice, S->copy_stream);
    if (e == hipSuccess && n > n1)
        e = hipMemcpyAsync(S->ring, S->staging + n1, (n - n1) * sizeof(uint4), hipMemcpyHostToDevice,
                           S->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(S->copy_ev, S->copy_stream);
    if (e == hipSuccess) e = hipEventSynchronize(S->copy_ev);
    return e;
}

// Launch a kernel instance when published tasks are unclaimed and no instance is running, or
// the latest one has begun to exit (its idle waves leave; busy ones finish their tasks first).
// With an instance running the device's claim cursor decides, read at most every 100 us (lock
// held; the read waits for a small copy).  Returns the launch error, if any.
static hipError_t maybe_launch(nq_server *S, double now) {
    if (S->inflight == 0) return hipSuccess;
    int running = 0;
    for (size_t i = 0; i < S->kstreams.size(); ++i) {
        if (!S->kused[i]) continue;
        const hipError_t q = hipEventQuery(S->kevents[i]);
        if (q == hipErrorNotReady) {
            ++running;
        } else {
            S->kused[i] = false;
            if (q != hipSuccess) return q;
        }
    }
    if (running > 0) {
        const unsigned exited = __atomic_load_n(&S->ctl->exited, __ATOMIC_ACQUIRE);
        if (exited != S->epoch || now - S->head_checked < 0.1) return hipSuccess;
        S->head_checked = now;
        unsigned long long h = 0;
        hipError_t e = hipMemcpyAsync(&h, S->head, sizeof h, hipMemcpyDeviceToHost, S->copy_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(S->copy_stream);
        if (e != hipSuccess) return e;
        S->claimed_seen = std::max<uint64_t>(S->claimed_seen, h);
        if (h >= S->tail) return hipSuccess;  // every published task is claimed: busy lanes finish them
    }
    // (No instance running while requests are in flight: their tasks are unclaimed, since waves
    // exit only with every claimed task counted and collect() ran first.)
    const int i = S->next_stream;
    if (S->kused[i]) return hipSuccess;  // every stream still runs an instance
    S->next_stream = (i + 1) % static_cast<int>(S->kstreams.size());
    ServeArgs a{};
    a.ring = S->ring;
    a.ring_mask = S->ring_cap - 1;
    a.head = S->head;
    a.ctl = S->ctl;
    a.done = S->done;
    a.remaining = S->remaining;
    a.count = S->count;
    a.seq = S->seq;
    a.error = S->error;
    a.idle_ticks = static_cast<unsigned long long>(IDLE_US * S->wall_khz / 1000.0);
    a.epoch = ++S->epoch;
    a.block = S->block;
    a.grid = S->grid;
    a.stack_depth = S->stack_depth;
    a.chunk = S->cfg.chunk;
    a.apush = S->apush;
    hipError_t e = launch_serve(a, S->kstreams[i]);
    if (e == hipSuccess) e = hipEventRecord(S->kevents[i], S->kstreams[i]);
    if (e == hipSuccess) {
        S->kused[i] = true;
        S->launches++;
    }
    return e;
}

// Collect counted slices from the kernel's completion records; a request is done when its last
// slice is counted (lock held).
static bool collect(nq_server *S, double now) {
    bool any = false;
    for (size_t s = 0; s < S->slot_slice.size(); ++s) {
        Slice *sl = S->slot_slice[s];
        if (!sl) continue;
        if (__atomic_load_n(&S->done[s].seq, __ATOMIC_ACQUIRE) != sl->seq) continue;
        Request &r = *sl->req;
        r.count += __atomic_load_n(&S->done[s].count, __ATOMIC_RELAXED);
        r.outstanding--;
        sl->counted = true;
        S->slot_slice[s] = nullptr;
        S->free_slots.push_back(static_cast<int>(s));
        S->inflight--;
        if (r.outstanding == 0 && r.next >= r.tasks.size() && !r.done) {
            r.done = true;
            r.t_done = now;
            S->completed++;
            S->solutions += r.count;
        }
        any = true;
    }
    while (!S->published.empty() && S->published.front().counted) S->published.pop_front();
    return any;
}

static void scheduler(nq_server *S) {
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        const double now = now_ms();
        bool progressed = collect(S, now);
        if (progressed) S->cv_done.notify_all();
        // Publish slices while the unclaimed published work is within the lookahead.  The claim
        // cursor is known from below: the kernel's hint in ctl->claimed (stores of two waves may
        // land out of order) and, when the hint would hold publication back, a read of the device
        // cursor itself (at most every 100 us), so a stale hint never stops the queue.
        S->claimed_seen = std::max<uint64_t>(S->claimed_seen, __atomic_load_n(&S->ctl->claimed, __ATOMIC_ACQUIRE));
        const double lookahead = 2.0 * S->cfg.batch_work;
        if (!S->pending.empty() && unclaimed_work(S, S->claimed_seen) >= lookahead && now - S->head_read > 0.1) {
            S->head_read = now;
            unsigned long long h = 0;
            hipError_t e = hipMemcpyAsync(&h, S->head, sizeof h, hipMemcpyDeviceToHost, S->copy_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(S->copy_stream);
            if (e == hipSuccess)
                S->claimed_seen = std::max<uint64_t>(S->claimed_seen, h);
            else
                set_error(S, NQ_ERR_HIP, std::string("serving cursor read failed: ") + hipGetErrorString(e));
        }
        Slice sl;
        if (S->error_code == NQ_OK && unclaimed_work(S, S->claimed_seen) < lookahead && pick_slice(S, now, &sl)) {
            Request &r = *sl.req;
            sl.slot = S->free_slots.back();
            S->free_slots.pop_back();
            sl.seq = S->next_seq++;
            const size_t first = r.next;
            r.next += sl.end - sl.begin;
            r.outstanding++;
            if (r.t_start < 0) r.t_start = now;
            S->pending_tasks -= sl.end - sl.begin;
            if (r.next >= r.tasks.size())
                S->pending.erase(std::remove(S->pending.begin(), S->pending.end(), sl.req), S->pending.end());
            lk.unlock();
            const hipError_t e = copy_slice(S, sl, first);
            lk.lock();
            if (e != hipSuccess) {
                set_error(S, NQ_ERR_HIP, std::string("serving copy failed: ") + hipGetErrorString(e));
            } else {
                S->published.push_back(sl);
                S->slot_slice[sl.slot] = &S->published.back();
                S->inflight++;
                S->tail = sl.end;
                __atomic_store_n(&S->ctl->tail, S->tail, __ATOMIC_RELEASE);
                S->slices++;
                progressed = true;
            }
        }
        if (S->error_code == NQ_OK) {
            const hipError_t e = maybe_launch(S, now_ms());
            if (e != hipSuccess) set_error(S, NQ_ERR_HIP, std::string("serving kernel failed: ") + hipGetErrorString(e));
        }
        if (S->inflight > 0 && S->busy_since < 0) S->busy_since = now;
        if (S->inflight == 0 && S->busy_since >= 0) {
            S->busy_ms += now - S->busy_since;
            S->busy_since = -1;
        }
        if (S->error_code != NQ_OK) {
            // Fail everything still queued or in flight so that no caller waits for it.
            for (auto &kv : S->all)
                if (!kv.second->done) {
                    kv.second->done = true;
                    kv.second->t_done = now;
                }
            S->pending.clear();
            S->cv_done.notify_all();
        }
        if (S->stop && S->pending.empty() && S->inflight == 0) break;
        if (S->stop && now - S->t_stop > CLOSE_DRAIN_MS) {
            // Requests still in flight long after close: give up on them rather than wait forever.
            set_error(S, NQ_ERR_STATE, "server closed with requests in flight");
            for (auto &kv : S->all) kv.second->done = true;
            S->cv_done.notify_all();
            break;
        }
        if (!progressed) S->cv_work.wait_for(lk, std::chrono::microseconds(S->inflight ? 20 : 2000));
    }
    // Let running instances see the stop flag and drain.
    __atomic_store_n(&S->ctl->stop, 1u, __ATOMIC_RELEASE);
}

extern "C" {

int nq_server_config_default(nq_server_config *cfg) {
    if (!cfg) return fail(NQ_ERR_ARG, "cfg is NULL");
    cfg->slots = 4;                    // streams for kernel instances
    cfg->slot_tasks = 1u << 25;        // ring capacity (tasks; 512 MB)
    cfg->slot_requests = 4096;         // slices with tasks in the ring at once
    cfg->max_pending_tasks = 1ull << 26;
    cfg->depth_target = 12;
    cfg->block = 128;
    cfg->chunk = 16;
    cfg->batch_work = 1e10;
    return NQ_OK;
}

int nq_server_open(const nq_server_config *cfg_in, nq_server **out) {
    if (!out) return fail(NQ_ERR_ARG, "out is NULL");
    *out = nullptr;
    nq_server_config cfg;
    nq_server_config_default(&cfg);
    if (cfg_in) {
        if (cfg_in->slots) cfg.slots = cfg_in->slots;
        if (cfg_in->slot_tasks) cfg.slot_tasks = cfg_in->slot_tasks;
        if (cfg_in->slot_requests) cfg.slot_requests = cfg_in->slot_requests;
        if (cfg_in->max_pending_tasks) cfg.max_pending_tasks = cfg_in->max_pending_tasks;
        if (cfg_in->depth_target) cfg.depth_target = cfg_in->depth_target;
        if (cfg_in->block) cfg.block = cfg_in->block;
        if (cfg_in->chunk) cfg.chunk = cfg_in->chunk;
        if (cfg_in->batch_work > 0) cfg.batch_work = cfg_in->batch_work;
    }
    if (cfg.slots < 1 || cfg.slots > 16 || cfg.slot_requests < 1 || cfg.slot_requests > (1 << 20) ||
        cfg.depth_target < 2 || cfg.depth_target > 26 || cfg.block < 64 || cfg.block > MAX_BLOCK ||
        cfg.block % 64 || cfg.chunk < 1 || cfg.chunk > 4096 || cfg.slot_tasks < 1 || cfg.slot_tasks > (1ull << 30))
        return fail(NQ_ERR_ARG, "invalid server configuration");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");
    auto *S = new nq_server();
    S->cfg = cfg;
    S->ring_cap = 1;
    while (S->ring_cap < cfg.slot_tasks) S->ring_cap <<= 1;
    const size_t nslots = static_cast<size_t>(cfg.slot_requests);
    hipError_t e = hipGetDevice(&S->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, S->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&S->wall_khz, hipDeviceAttributeWallClockRate, S->device);
    if (e == hipSuccess) e = hipMalloc(&S->ring, S->ring_cap * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&S->head, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->error, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->remaining, nslots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->count, nslots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->seq, nslots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc(&S->ctl, sizeof(ServeCtl), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc(&S->done, nslots * sizeof(ServeDone), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc(&S->staging, STAGING_TASKS * sizeof(uint4), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&S->slot_init, 4 * sizeof(unsigned long long), hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S->copy_ev, hipEventDisableTiming);
    S->kstreams.assign(cfg.slots, nullptr);
    S->kevents.assign(cfg.slots, nullptr);
    S->kused.assign(cfg.slots, false);
    for (int i = 0; i < cfg.slots && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&S->kstreams[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&S->kevents[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) {
        // Zero the claim cursor, error word and slot counters with kernels (no memset).
        e = launch_zero(S->head, 1, S->copy_stream);
        if (e == hipSuccess) e = launch_zero(S->error, 1, S->copy_stream);
        if (e == hipSuccess) e = launch_zero(S->remaining, nslots, S->copy_stream);
        if (e == hipSuccess) e = launch_zero(S->count, nslots, S->copy_stream);
        if (e == hipSuccess) e = launch_zero(S->seq, nslots, S->copy_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(S->copy_stream);
    }
    if (e == hipSuccess) {
        const int frames = std::max(1, cfg.depth_target - 1);  // tasks keep at most depth_target free rows
        int bpc = 0;
        e = serve_fit(cfg.block, frames, 0, &S->apush, &bpc);
        S->block = cfg.block;
        S->stack_depth = frames;
        S->grid = std::max(1, S->cus * std::max(1, bpc));
    }
    if (e != hipSuccess) {
        free_server(S);
        return fail_hip(e, "server allocation");
    }
    memset(S->ctl, 0, sizeof(ServeCtl));
    memset(S->done, 0, nslots * sizeof(ServeDone));
    S->slot_slice.assign(nslots, nullptr);
    for (int s = static_cast<int>(nslots) - 1; s >= 0; --s) S->free_slots.push_back(s);
    S->t_open = now_ms();
    S->thr = std::thread(scheduler, S);
    *out = S;
    return NQ_OK;
}

int nq_server_submit(nq_server *S, const nq_instance *inst, uint64_t *ticket) {
    if (!S || !inst || !ticket) return fail(NQ_ERR_ARG, "NULL argument");
    uint32_t c, l, r;
    if (!encode_instance(*inst, &c, &l, &r)) return fail(NQ_ERR_ARG, "invalid N/k or attacking prefix");
    auto req = std::make_shared<Request>();
    req->N = inst->N;
    const int d0 = append_instance_tasks(inst->N, inst->k, c, l, r, S->cfg.depth_target, 0, req->tasks);
    req->task_work = est_task_nodes(inst->N - d0);
    // A request that can never fit the ring is refused for good (NQ_ERR_NOMEM), unlike a full
    // queue (NQ_ERR_BUSY), which a retry may pass.
    if (req->tasks.size() > S->ring_cap)
        return fail(NQ_ERR_NOMEM, "request larger than the server's ring (raise slot_tasks or depth_target)");
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) return fail(NQ_ERR_STATE, "server is closing");
    if (S->pending_tasks + req->tasks.size() > S->cfg.max_pending_tasks)
        return fail(NQ_ERR_BUSY, "pending task budget exhausted");
    req->ticket = S->next_ticket++;
    req->t_submit = now_ms();
    S->all[req->ticket] = req;
    if (req->tasks.empty()) {
        // No legal continuation: count 0, done at once.
        req->done = true;
        req->t_done = req->t_submit;
        S->completed++;
    } else {
        S->pending_tasks += req->tasks.size();
        S->pending.push_back(req);
    }
    *ticket = req->ticket;
    S->cv_work.notify_one();
    return NQ_OK;
}

static int take_result(nq_server *S, std::unordered_map<uint64_t, std::shared_ptr<Request>>::iterator it,
                       uint64_t *count, nq_request_stats *st) {
    Request &r = *it->second;
    if (count) *count = r.count;
    if (st) {
        st->queue_ms = (r.t_start < 0 ? r.t_done : r.t_start) - r.t_submit;
        st->latency_ms = r.t_done - r.t_submit;
        st->ntasks = r.tasks.size();
    }
    S->all.erase(it);
    if (S->error_code) return fail(S->error_code, S->error_msg);
    return NQ_OK;
}

int nq_server_poll(nq_server *S, uint64_t ticket, uint64_t *count, nq_request_stats *st) {
    if (!S) return fail(NQ_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return fail(NQ_ERR_ARG, "unknown ticket");
    if (!it->second->done) return fail(NQ_ERR_STATE, "pending");
    return take_result(S, it, count, st);
}

int nq_server_wait(nq_server *S, uint64_t ticket, int timeout_ms, uint64_t *count, nq_request_stats *st) {
    if (!S) return fail(NQ_ERR_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return fail(NQ_ERR_ARG, "unknown ticket");
    auto req = it->second;
    auto pred = [&] { return req->done; };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, pred);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        return fail(NQ_ERR_STATE, "pending");
    return take_result(S, S->all.find(ticket), count, st);
}

int nq_server_stats(nq_server *S, uint64_t *completed, uint64_t *batches, uint64_t *solutions) {
    if (!S) return fail(NQ_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> g(S->mu);
    if (completed) *completed = S->completed;
    if (batches) *batches = S->slices;
    if (solutions) *solutions = S->solutions;
    return NQ_OK;
}

int nq_server_timing(nq_server *S, double *device_ms, double *open_ms) {
    if (!S) return fail(NQ_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> g(S->mu);
    const double now = now_ms();
    if (device_ms) *device_ms = S->busy_ms + (S->busy_since >= 0 ? now - S->busy_since : 0);
    if (open_ms) *open_ms = now - S->t_open;
    return NQ_OK;
}

int nq_server_close(nq_server *S) {
    if (!S) return NQ_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->t_stop = now_ms();
        S->cv_work.notify_all();
    }
    if (S->thr.joinable()) S->thr.join();
    // Every instance sees the stop flag within its idle wait; wait for them to drain.
    for (size_t i = 0; i < S->kstreams.size(); ++i) (void)hipStreamSynchronize(S->kstreams[i]);
    unsigned long long err = 0;
    (void)hipMemcpy(&err, S->error, sizeof err, hipMemcpyDeviceToHost);
    int rc = S->error_code;
    if (!rc && err) rc = (err & DEV_ERR_LDS) ? NQ_ERR_STATE : NQ_ERR_DEPTH;
    free_server(S);
    return rc ? fail(rc, "server had a device error") : NQ_OK;
}

}  // extern "C"
