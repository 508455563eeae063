// Serving with continuous batching over a persistent kernel (no reference counterpart).
//
// Callers submit counting requests (an nq_instance each) at any time and get a ticket.  Submit
// validates and expands the request into tasks on the caller's thread and admits it while the
// pending work fits the server's task budget (else NQ_ERR_BUSY).  A scheduler thread feeds one
// device-resident ring of tasks that a persistent kernel (serve_kernel, nq_dfs.hip) consumes
// while the host keeps appending to it:
//
//   * it publishes requests in slices (at most batch_work estimated DFS nodes each), least
//     estimated remaining work first (aged after a second of waiting), and only while the published but
//     unclaimed work stays under a lookahead of two slices, so a request that arrives behind a
//     large one waits for at most that much work (every kernel wave records the end of its last
//     claim in host memory; the largest record is the claim cursor);
//   * the ring, the tail and the completion records live in pinned, coherent host memory that
//     the kernel reads and writes directly: publishing a slice is plain stores (its tasks and a
//     terminator entry carrying its size, outside the server lock) followed by raising the tail,
//     and no copy is ever queued behind the running kernel;
//   * every published slice occupies one of slot_requests slots; the kernel adds every finished
//     task's count to its slot and, when the slice's last entry is counted, writes {count, the
//     slot's completion number} into host memory, where the scheduler finds it; the slice's ring
//     space is then free, and a request is done when its last slice is counted;
//   * kernel waves claim tasks as tickets with a fetch-and-add on the device cursor (a claim
//     never fails, however many waves claim at once) and read their entries with one coalesced
//     load per claim; a ticket past the published tail waits for the host to publish it;
//   * one kernel instance runs while requests are in flight; the scheduler stops it once every
//     published task has been counted and nothing has arrived for IDLE_STOP_MS (so that an idle
//     server leaves the device to other work), and launches the next one on the same stream when
//     work comes again.
//
// A request's count (the sum over its tasks) is exactly the one nq_instance_count gives.
//
// With NQ_SERVE_TRACE=<file> in the environment the server records its events (requests
// admitted, slices published, kernel instances launched and stopped, slices counted) and writes
// them to <file> at close, one per line: "<ms since open> <event> <fields>".
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nq.h"
#include "nq_internal.h"
#include "nq_kernels.h"
#include "nq_tasks.h"

namespace nq {
bool encode_instance(const nq_instance &in, uint32_t *c, uint32_t *l, uint32_t *r);
}  // namespace nq

using namespace nq;

namespace {

struct Request {
    uint64_t ticket = 0;
    int N = 0;
    TaskVec tasks;             // tag = twice << 5 | N; the slice's slot is added when published
    size_t next = 0;           // first task not yet published
    int outstanding = 0;       // published slices not yet counted
    double task_work = 1;      // estimated DFS nodes per task
    int free_rows = 0;         // rows every task leaves to the kernel
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
// slot the kernel counts it in (every slice has its own slot and completion number, so that its
// ring space is freed as soon as it is counted, whatever the rest of its request does).
struct Slice {
    std::shared_ptr<Request> req;
    uint64_t begin = 0, end = 0;  // ring positions: the tasks, then the terminator at end - 1
    uint64_t ntasks = 0;
    double work = 0;
    int slot = -1;
    uint64_t seq = 0;  // the slice's number (trace only)
    uint64_t gen = 0;  // the slot's completion number this slice's record will carry
    bool counted = false;
};

// Tasks one slice writes into the ring at most (SERVE_SLICE_MAX = 2^16): the scheduler collects
// completions between slices, so a slice's write time (about 0.5 ms here) bounds how late a
// finished request is seen.

// Most tasks of `free_rows` free rows and weight `weight` whose completions certainly fit the
// kernel's count field: a task has at most free_rows! completions.
uint64_t slice_cap(int free_rows, int weight) {
    long double most = weight;
    for (int i = 2; i <= free_rows; ++i) most *= i;
    const long double cap = std::ldexp(1.0L, SERVE_COUNT_BITS) / most;
    return cap >= static_cast<long double>(SERVE_SLICE_MAX) ? SERVE_SLICE_MAX
                                                            : std::max<uint64_t>(1, static_cast<uint64_t>(cap));
}
constexpr size_t SMALL_REQUEST_TASKS = 4096;    // a request is split into at least this many tasks
constexpr int MIN_FREE_ROWS = 5;                // ... unless they would keep fewer free rows
constexpr double IDLE_STOP_MS = 0.2;          // an idle server stops its kernel after this long
constexpr double ORPHAN_S = 30;               // a kernel wave without work for this long exits
constexpr double AGING_AFTER_MS = 1000;       // waiting time after which a request gains priority
constexpr double CLOSE_DRAIN_MS = 120000;     // close() waits this long for requests in flight

struct TraceEvent {
    double t;
    char kind;  // 'A' request admitted, 'P' slice published, 'L' instance launched, 'S' instance
                // stopped, 'C' slice counted
    uint64_t a, b, c;
    double d;
};

}  // namespace

struct nq_server {
    nq_server_config cfg{};
    int device = 0, cus = 0, wall_khz = 0;
    // Device memory.
    unsigned long long *head = nullptr, *acc = nullptr, *gen = nullptr, *error = nullptr;
    // Pinned, coherent host memory.
    uint4 *ring = nullptr;
    uint64_t ring_cap = 0;  // a power of two
    ServeCtl *ctl = nullptr;
    ServeDone *done = nullptr;
    unsigned long long *claimed = nullptr;  // one entry per wave of an instance
    uint64_t waves = 0;                     // waves of one kernel instance
    hipStream_t stream = nullptr;
    hipEvent_t kev = nullptr;  // recorded after the latest instance
    int block = 0, grid = 0, stack_depth = 0;
    bool apush = false;
    unsigned epoch = 0;      // number of the latest kernel instance
    bool running = false;    // an instance was launched and has not been seen to end
    bool stopping = false;   // ... and it was asked to stop
    uint64_t base = 0;       // ring position of the latest instance's ticket 0
    int parity = 0;          // the tail the next (or running) instance reads
    uint64_t gen_base = 0;   // ring position of the next instance's ticket 0 (the tail at the last stop)
    double idle_since = -1;  // since when nothing is pending or in flight (ms)
    // Scheduling state, under mu.
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::shared_ptr<Request>> pending;  // requests with tasks not yet published
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    std::vector<int> free_slots;
    std::deque<Slice> published;  // ring FIFO, freed from the front as slices are counted
    uint64_t tail = 0;            // ring entries published
    uint64_t next_ticket = 1, next_seq = 1, pending_tasks = 0;  // next_seq numbers slices
    std::vector<uint64_t> slot_gen;  // completions each slot will have had once its slice is counted
    int inflight = 0;  // slices published and not yet counted
    double head_read = -1e30;   // last scan of the waves' claim records (ms)
    uint64_t claimed_seen = 0;  // largest claim record seen (a lower bound of the claim cursor)
    bool stop = false;
    double t_stop = 0;  // when close() was called
    int error_code = NQ_OK;
    std::string error_msg;
    std::thread thr;
    // Statistics and the event trace (NQ_SERVE_TRACE).
    uint64_t completed = 0, slices = 0, solutions = 0, launches = 0;
    std::string trace_path;
    std::vector<TraceEvent> trace;
    double busy_ms = 0, busy_since = -1, t_open = 0;
};

static void free_server(nq_server *S) {
    for (unsigned long long *p : {S->head, S->acc, S->gen, S->error})
        if (p) (void)hipFree(p);
    if (S->ring) (void)hipHostFree(S->ring);
    if (S->ctl) (void)hipHostFree(S->ctl);
    if (S->done) (void)hipHostFree(S->done);
    if (S->claimed) (void)hipHostFree(S->claimed);
    if (S->kev) (void)hipEventDestroy(S->kev);
    if (S->stream) (void)hipStreamDestroy(S->stream);
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
// work (aged once it has waited AGING_AFTER_MS); its tasks up to the slice budget, the ring space and
// slice_cap.  Returns false when nothing can be published now (no request, no free slot, no space).
static bool pick_slice(nq_server *S, double now, Slice *out) {
    if (S->pending.empty() || S->free_slots.empty()) return false;
    const uint64_t front = S->published.empty() ? S->tail : S->published.front().begin;
    const uint64_t space = S->ring_cap - (S->tail - front);
    if (space < 2) return false;  // a task and the terminator
    auto key = [&](const std::shared_ptr<Request> &r) {
        const double rem = static_cast<double>(r->tasks.size() - r->next) * r->task_work;
        // ~1 ms per 1e9 nodes; a request that has waited longer than AGING_AFTER_MS gains 2 ms
        // of priority per ms, so that a stream of small requests cannot starve a large one.
        return rem / 1e9 - 2.0 * std::max(0.0, now - r->t_submit - AGING_AFTER_MS);
    };
    std::shared_ptr<Request> best;
    for (auto &r : S->pending)
        if (!best || key(r) < key(best)) best = r;
    const uint64_t left = best->tasks.size() - best->next;
    uint64_t take = static_cast<uint64_t>(std::max(1.0, S->cfg.batch_work / best->task_work));
    const int weight = (!best->tasks.empty() && (best->tasks[0].w & (1u << 5))) ? 2 : 1;
    take = std::min<uint64_t>({take, left, space - 1, slice_cap(best->free_rows, weight)});
    out->req = best;
    out->begin = S->tail;
    out->end = S->tail + take + 1;
    out->ntasks = take;
    out->work = static_cast<double>(take) * best->task_work;
    return true;
}

// Write a slice's tasks (tagged with its slot) and its terminator into its ring positions; the
// scheduler publishes them by raising the tail afterwards.  Called without the lock: the
// positions and the slot are the slice's alone until then.
static void write_slice(nq_server *S, const Slice &sl, size_t first) {
    const Request &r = *sl.req;
    const uint64_t mask = S->ring_cap - 1;
    const uint32_t tag = static_cast<uint32_t>(sl.slot) << SERVE_TAG_SLOT_SHIFT;
    for (uint64_t j = 0; j < sl.ntasks; ++j) {
        uint4 u = r.tasks[first + j];
        u.w = tag | (u.w & 63u);
        S->ring[(sl.begin + j) & mask] = u;
    }
    S->ring[(sl.end - 1) & mask] = make_uint4(static_cast<uint32_t>(sl.ntasks), 0, 0, tag | SERVE_TAG_TERM);
}

// The largest claim end the latest instance's waves have recorded: a lower bound of its claim
// cursor (tickets past the tail included), equal to it once the waves' stores have landed.
static uint64_t claimed_max(const nq_server *S) {
    const uint64_t tag = static_cast<uint64_t>(S->epoch & 0xFFFFu);
    const uint64_t low = (1ull << SERVE_CLAIM_EPOCH_SHIFT) - 1;
    uint64_t m = 0;
    for (uint64_t i = 0; i < S->waves; ++i) {
        const uint64_t rec = __atomic_load_n(&S->claimed[i], __ATOMIC_RELAXED);
        if ((rec >> SERVE_CLAIM_EPOCH_SHIFT) == tag) m = std::max(m, rec & low);
    }
    return S->base + m;
}

// Ask the running instance to stop and switch publication to the other tail, which the next
// instance reads from the current tail on (lock held).  Called with every published task counted
// (or, at close after an error, with the instance left to count what it has): the stopped
// instance sees no task published from now on, and its waves exit once idle.
static void stop_instance(nq_server *S) {
    __atomic_store_n(&S->ctl->stop_epoch, S->epoch, __ATOMIC_RELEASE);
    S->stopping = true;
    S->parity ^= 1;
    S->gen_base = S->tail;
    __atomic_store_n(&S->ctl->tail[S->parity], S->tail, __ATOMIC_RELEASE);
    if (!S->trace_path.empty()) S->trace.push_back({now_ms() - S->t_open, 'S', S->epoch, S->tail, 0, 0});
}

// Keep one kernel instance running while slices are in flight (lock held): note an instance
// that has ended, stop the running one once the server has been idle for IDLE_STOP_MS (every
// published task counted, nothing pending), and launch a new one when slices are in flight and
// none runs (or the running one is stopping: the new one waits for it on the stream).
static hipError_t manage_instance(nq_server *S, double now) {
    if (S->running) {
        const hipError_t q = hipEventQuery(S->kev);
        if (q == hipSuccess) {
            S->running = false;
            if (!S->stopping && S->inflight > 0) {
                set_error(S, NQ_ERR_STATE, "serving kernel exited with requests in flight");
                return hipSuccess;
            }
        } else if (q != hipErrorNotReady) {
            return q;
        }
    }
    if (S->inflight == 0 && S->pending.empty()) {
        if (S->idle_since < 0) S->idle_since = now;
        if (S->running && !S->stopping && (now - S->idle_since > IDLE_STOP_MS || S->stop)) {
            stop_instance(S);
        }
        return hipSuccess;
    }
    S->idle_since = -1;
    if (S->inflight == 0 || (S->running && !S->stopping)) return hipSuccess;
    // The next instance: ticket 0 is the first task published since the last one stopped (or
    // since open), and its cursor is zeroed behind the stopping instance on the same stream.
    ServeArgs a{};
    a.ring = S->ring;
    a.ring_mask = S->ring_cap - 1;
    a.base = S->gen_base;
    a.parity = S->parity;
    a.head = S->head;
    a.ctl = S->ctl;
    a.done = S->done;
    a.claimed = S->claimed;
    a.acc = S->acc;
    a.gen = S->gen;
    a.error = S->error;
    a.orphan_ticks = static_cast<unsigned long long>(ORPHAN_S * 1000.0 * S->wall_khz);
    a.epoch = ++S->epoch;
    a.block = S->block;
    a.grid = S->grid;
    a.stack_depth = S->stack_depth;
    a.chunk = S->cfg.chunk;
    a.apush = S->apush;
    hipError_t e = launch_zero(S->head, 1, S->stream);
    if (e == hipSuccess) e = launch_serve(a, S->stream);
    if (e == hipSuccess) e = hipEventRecord(S->kev, S->stream);
    if (e != hipSuccess) return e;
    S->base = a.base;
    S->claimed_seen = a.base;
    S->running = true;
    S->stopping = false;
    S->launches++;
    if (!S->trace_path.empty())
        S->trace.push_back({now - S->t_open, 'L', a.epoch, a.base, S->tail, 0});
    return hipSuccess;
}

// Collect counted slices from the kernel's completion records; a request is done when its last
// slice is counted (lock held).
static bool collect(nq_server *S, double now) {
    bool any = false;
    for (Slice &sl : S->published) {
        if (sl.counted) continue;
        const int s = sl.slot;
        if (__atomic_load_n(&S->done[s].gen, __ATOMIC_ACQUIRE) != sl.gen) continue;
        Request &r = *sl.req;
        r.count += __atomic_load_n(&S->done[s].count, __ATOMIC_RELAXED);
        r.outstanding--;
        sl.counted = true;
        if (!S->trace_path.empty()) S->trace.push_back({now - S->t_open, 'C', static_cast<uint64_t>(s), sl.seq, sl.ntasks, 0});
        S->free_slots.push_back(s);
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
        // cursor is known from below through the waves' claim records, scanned when the last
        // known value would hold publication back (at most every 20 us).
        const double lookahead = 2.0 * S->cfg.batch_work;
        if (!S->pending.empty() && unclaimed_work(S, S->claimed_seen) >= lookahead && now - S->head_read > 0.02) {
            S->head_read = now;
            S->claimed_seen = std::max<uint64_t>(S->claimed_seen, claimed_max(S));
        }
        Slice sl;
        if (S->error_code == NQ_OK && unclaimed_work(S, S->claimed_seen) < lookahead && pick_slice(S, now, &sl)) {
            Request &r = *sl.req;
            sl.slot = S->free_slots.back();
            S->free_slots.pop_back();
            sl.seq = S->next_seq++;
            sl.gen = ++S->slot_gen[sl.slot];
            const size_t first = r.next;
            r.next += sl.ntasks;
            r.outstanding++;
            if (r.t_start < 0) r.t_start = now;
            S->pending_tasks -= sl.ntasks;
            if (r.next >= r.tasks.size())
                S->pending.erase(std::remove(S->pending.begin(), S->pending.end(), sl.req), S->pending.end());
            lk.unlock();
            const double t_w = now_ms();
            write_slice(S, sl, first);
            lk.lock();
            if (!S->trace_path.empty())
                S->trace.push_back({t_w - S->t_open, 'P', sl.begin, sl.end, sl.req->ticket, now_ms() - t_w});
            S->published.push_back(sl);
            S->inflight++;
            S->tail = sl.end;
            __atomic_store_n(&S->ctl->tail[S->parity], S->tail, __ATOMIC_RELEASE);
            S->slices++;
            progressed = true;
        }
        if (S->error_code == NQ_OK) {
            const hipError_t e = manage_instance(S, now_ms());
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
        if (S->stop && (S->error_code != NQ_OK || (S->pending.empty() && S->inflight == 0))) break;
        if (S->stop && now - S->t_stop > CLOSE_DRAIN_MS) {
            // Requests still in flight long after close: give up on them rather than wait forever.
            set_error(S, NQ_ERR_STATE, "server closed with requests in flight");
            for (auto &kv : S->all) kv.second->done = true;
            S->cv_done.notify_all();
            break;
        }
        if (!progressed)
            S->cv_work.wait_for(lk, std::chrono::microseconds(S->inflight ? 20 : (S->running && !S->stopping) ? 50 : 2000));
    }
    // Stop the running instance (after an error, with tasks in flight: it counts every published
    // task it has taken or can still take, then exits).
    if (S->running && !S->stopping) stop_instance(S);
}

extern "C" {

int nq_server_config_default(nq_server_config *cfg) {
    if (!cfg) return fail(NQ_ERR_ARG, "cfg is NULL");
    cfg->waves_per_cu = 0;             // as many as fit
    cfg->slot_tasks = 1u << 23;        // ring capacity (tasks; 128 MB of pinned host memory)
    cfg->slot_requests = 4096;         // slices with tasks in the ring at once
    cfg->max_pending_tasks = 1ull << 26;
    cfg->depth_target = 12;
    cfg->block = 128;
    // Claims of 32 tickets and slices of 5e8 estimated nodes (the 400-request Poisson set at 70 %
    // load, two runs each: p50 2.38-2.52 ms and prefixed p50 1.77-1.80 ms against 3.07-3.09 and
    // 2.61-2.70 ms with 64 tickets and 1e9 nodes; capacity 248-250 against 257-261 requests/s).
    cfg->chunk = 32;
    cfg->batch_work = 5e8;
    return NQ_OK;
}

int nq_server_open(const nq_server_config *cfg_in, nq_server **out) {
    if (!out) return fail(NQ_ERR_ARG, "out is NULL");
    *out = nullptr;
    nq_server_config cfg;
    nq_server_config_default(&cfg);
    if (cfg_in) {
        if (cfg_in->waves_per_cu) cfg.waves_per_cu = cfg_in->waves_per_cu;
        if (cfg_in->slot_tasks) cfg.slot_tasks = cfg_in->slot_tasks;
        if (cfg_in->slot_requests) cfg.slot_requests = cfg_in->slot_requests;
        if (cfg_in->max_pending_tasks) cfg.max_pending_tasks = cfg_in->max_pending_tasks;
        if (cfg_in->depth_target) cfg.depth_target = cfg_in->depth_target;
        if (cfg_in->block) cfg.block = cfg_in->block;
        if (cfg_in->chunk) cfg.chunk = cfg_in->chunk;
        if (cfg_in->batch_work > 0) cfg.batch_work = cfg_in->batch_work;
    }
    if (cfg.waves_per_cu < 0 || cfg.slot_requests < 1 || cfg.slot_requests > (1 << 20) ||
        cfg.depth_target < 2 || cfg.depth_target > 26 || cfg.block < 64 || cfg.block > MAX_BLOCK ||
        cfg.block % 64 || cfg.chunk < 1 || cfg.chunk > SERVE_BUF || cfg.slot_tasks < 1 || cfg.slot_tasks > (1ull << 30))
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
    if (e == hipSuccess) e = hipMalloc(&S->head, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->error, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->acc, nslots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->gen, nslots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc(&S->ring, S->ring_cap * sizeof(uint4), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc(&S->ctl, sizeof(ServeCtl), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc(&S->done, nslots * sizeof(ServeDone), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S->kev, hipEventDisableTiming);
    if (e == hipSuccess) {
        // Zero the error word and slot counters with kernels (no memset).
        e = launch_zero(S->error, 1, S->stream);
        if (e == hipSuccess) e = launch_zero(S->acc, nslots, S->stream);
        if (e == hipSuccess) e = launch_zero(S->gen, nslots, S->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
    }
    if (e == hipSuccess) {
        const int frames = std::max(1, cfg.depth_target - 1);  // tasks keep at most depth_target free rows
        int bpc = 0;
        e = serve_fit(cfg.block, frames, 0, &S->apush, &bpc);
        S->block = cfg.block;
        S->stack_depth = frames;
        if (cfg.waves_per_cu > 0) bpc = std::min(bpc, std::max(1, cfg.waves_per_cu / (cfg.block / 64)));
        S->grid = std::max(1, S->cus * std::max(1, bpc));
        S->waves = static_cast<uint64_t>(S->grid) * (cfg.block / 64);
    }
    if (e == hipSuccess)
        e = hipHostMalloc(&S->claimed, S->waves * sizeof(unsigned long long), hipHostMallocCoherent);
    if (e != hipSuccess) {
        free_server(S);
        return fail_hip(e, "server allocation");
    }
    memset(S->ctl, 0, sizeof(ServeCtl));
    memset(S->done, 0, nslots * sizeof(ServeDone));
    memset(S->claimed, 0, S->waves * sizeof(unsigned long long));
    for (int s = static_cast<int>(nslots) - 1; s >= 0; --s) S->free_slots.push_back(s);
    S->slot_gen.assign(nslots, 0);
    S->t_open = now_ms();
    if (const char *tp = getenv("NQ_SERVE_TRACE")) S->trace_path = tp;
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
    // Tasks keep at most depth_target free rows; a small request is split further, until it has
    // SMALL_REQUEST_TASKS tasks or its tasks keep MIN_FREE_ROWS, so that many lanes share it
    // (one lane alone would count a 12-row subtree about a thousand times slower than the device).
    // (Never past the admission budget.)
    int depth = S->cfg.depth_target;
    int d0 = append_instance_tasks(inst->N, inst->k, c, l, r, depth, 0, req->tasks);
    while (req->tasks.size() < SMALL_REQUEST_TASKS && inst->N - d0 > MIN_FREE_ROWS && depth > MIN_FREE_ROWS) {
        TaskVec finer;
        const int d1 = append_instance_tasks(inst->N, inst->k, c, l, r, depth - 1, 0, finer);
        if (finer.size() > S->cfg.max_pending_tasks) break;
        req->tasks.swap(finer);
        d0 = d1;
        --depth;
    }
    req->task_work = est_task_nodes(inst->N - d0);
    req->free_rows = std::max(0, inst->N - d0);
    // A request that can never fit the admission budget is refused for good (NQ_ERR_NOMEM), unlike
    // a full queue (NQ_ERR_BUSY), which a retry may pass.  (The ring takes it slice by slice.)
    if (req->tasks.size() > S->cfg.max_pending_tasks)
        return fail(NQ_ERR_NOMEM, "request larger than the admission budget (raise max_pending_tasks or depth_target)");
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) return fail(NQ_ERR_STATE, "server is closing");
    if (S->pending_tasks + req->tasks.size() > S->cfg.max_pending_tasks)
        return fail(NQ_ERR_BUSY, "pending task budget exhausted");
    req->ticket = S->next_ticket++;
    req->t_submit = now_ms();
    if (!S->trace_path.empty())
        S->trace.push_back({req->t_submit - S->t_open, 'A', req->ticket, req->tasks.size(),
                            static_cast<uint64_t>(req->N), req->task_work});
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
    // The scheduler has stopped the instance; it exits once its waves see that.
    (void)hipStreamSynchronize(S->stream);
    unsigned long long err = 0;
    (void)hipMemcpy(&err, S->error, sizeof err, hipMemcpyDeviceToHost);
    int rc = S->error_code;
    if (!rc && err) rc = (err & DEV_ERR_LDS) ? NQ_ERR_STATE : NQ_ERR_DEPTH;
    if (!S->trace_path.empty()) {
        if (FILE *f = fopen(S->trace_path.c_str(), "w")) {
            for (const TraceEvent &ev : S->trace)
                fprintf(f, "%.3f %c %llu %llu %llu %.3f\n", ev.t, ev.kind, static_cast<unsigned long long>(ev.a),
                        static_cast<unsigned long long>(ev.b), static_cast<unsigned long long>(ev.c), ev.d);
            fclose(f);
        }
    }
    free_server(S);
    return rc ? fail(rc, "server had a device error") : NQ_OK;
}

}  // extern "C"
