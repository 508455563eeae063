// Serving loop with continuous batching (no reference counterpart).
//
// Callers submit counting requests (an nq_instance each) at any time and get a ticket.  Submit
// validates the instance and expands it into tasks on the caller's thread, and admits it only
// when the pending work stays within the server's task budget (else NQ_ERR_BUSY).  A scheduler
// thread keeps `slots` device-resident task queues (task buffer, per-request counters, cursor
// state, pinned result buffer, stream).  Whenever a slot is free it cuts the next batch: pending
// requests in order of least estimated remaining work (aged by their waiting time), each
// contributing tasks until the batch's estimated work budget is spent, so a large request is
// served as slices over consecutive batches and a small one joins the next batch instead of
// waiting behind it.  Batches launch on their slot's stream and overlap on the device.  A
// request's count (the sum over its slices) is exactly the one nq_instance_count gives.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
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
    std::vector<uint4> tasks;  // the tag's index bits are rewritten when placed in a batch
    size_t next = 0;           // first task not yet placed in a batch
    int outstanding = 0;       // batches holding slices of this request
    int frames = 1;
    double task_work = 1;      // estimated DFS nodes per task
    double t_submit = 0, t_start = -1, t_done = 0;
    uint64_t count = 0;
    bool done = false;
};

// Estimated DFS nodes below a prefix with m free rows (measured: ~6.3e5 at 14 free rows, a
// factor ~5.7 per additional row).  Only orders and sizes batches; never affects counts.
double est_task_nodes(int m) {
    double w = 6.3e5;
    for (int i = 14; i < m; ++i) w *= 5.7;
    for (int i = m; i < 14; ++i) w /= 5.7;
    return std::max(w, 1.0);
}

struct Slice {
    std::shared_ptr<Request> req;
    size_t begin = 0, end = 0;  // task range of the request placed in this batch
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done_ev = nullptr;
    uint4 *d_tasks = nullptr;
    uint4 *h_tasks = nullptr;  // pinned staging
    uint64_t *d_counts = nullptr;
    uint64_t *h_counts = nullptr;  // pinned
    DevState *d_state = nullptr;
    DevState *h_state = nullptr;  // pinned copy of the batch's flags
    std::vector<Slice> slices;
    double t_launch = 0;
    bool busy = false;
};

}  // namespace

struct nq_server {
    nq_server_config cfg{};
    int cus = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::shared_ptr<Request>> pending;  // requests with tasks not yet placed
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_ticket = 1;
    uint64_t pending_tasks = 0;
    bool stop = false;
    int error = NQ_OK;
    std::string error_msg;
    std::thread thr;
    // statistics
    uint64_t completed = 0, batches = 0, solutions = 0;
    double device_ms = 0, t_open = 0;
};

static void free_slot(Slot &s) {
    if (s.d_tasks) (void)hipFree(s.d_tasks);
    if (s.d_counts) (void)hipFree(s.d_counts);
    if (s.d_state) (void)hipFree(s.d_state);
    if (s.h_tasks) (void)hipHostFree(s.h_tasks);
    if (s.h_counts) (void)hipHostFree(s.h_counts);
    if (s.h_state) (void)hipHostFree(s.h_state);
    if (s.done_ev) (void)hipEventDestroy(s.done_ev);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s = Slot();
}

// Cut the next batch from the pending requests into slot s (least estimated remaining work
// first, aged by waiting time; up to the batch work budget).  Called with the server lock held.
static void plan_slot(nq_server *S, Slot &s, double now) {
    std::vector<std::shared_ptr<Request>> order = S->pending;
    auto key = [&](const std::shared_ptr<Request> &r) {
        const double rem = static_cast<double>(r->tasks.size() - r->next) * r->task_work;
        return rem / 1e9 - 2.0 * (now - r->t_submit);  // ~1 ms per 1e9 nodes; 2x aging
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](const std::shared_ptr<Request> &a, const std::shared_ptr<Request> &b) { return key(a) < key(b); });
    uint64_t n = 0;
    double budget = S->cfg.batch_work;
    s.slices.clear();
    for (auto &rp : order) {
        if (budget <= 0 || n >= S->cfg.slot_tasks || s.slices.size() >= static_cast<size_t>(S->cfg.slot_requests)) break;
        Request &r = *rp;
        const size_t left = r.tasks.size() - r.next;
        size_t take = static_cast<size_t>(std::max(1.0, budget / r.task_work));
        take = std::min<size_t>({take, left, static_cast<size_t>(S->cfg.slot_tasks - n)});
        s.slices.push_back({rp, r.next, r.next + take});
        r.next += take;
        r.outstanding++;
        if (r.t_start < 0) r.t_start = now;
        budget -= static_cast<double>(take) * r.task_work;
        n += take;
        S->pending_tasks -= take;
    }
    S->pending.erase(std::remove_if(S->pending.begin(), S->pending.end(),
                                    [](const std::shared_ptr<Request> &r) { return r->next >= r->tasks.size(); }),
                     S->pending.end());
}

// Pack the planned slices into the slot's pinned staging buffer and enqueue upload, batched
// kernel and result download on its stream (server lock held).
static hipError_t launch_slot(nq_server *S, Slot &s) {
    uint64_t n = 0;
    int frames = 1;
    // Pack the slices with the largest tasks first: the batch ends when its last task does, so
    // the longest tasks should be claimed first and the short ones fill the tail.  Counters are
    // per slice (the tag's index), so the order changes no count.
    std::vector<size_t> order(s.slices.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return s.slices[a].req->task_work > s.slices[b].req->task_work;
    });
    for (size_t i : order) {
        const Slice &sl = s.slices[i];
        const uint32_t idx = static_cast<uint32_t>(i);
        for (size_t j = sl.begin; j < sl.end; ++j) {
            uint4 u = sl.req->tasks[j];
            u.w = (idx << 6) | (u.w & 63u);
            s.h_tasks[n++] = u;
        }
        frames = std::max(frames, sl.req->frames);
    }
    hipError_t e = n ? hipMemcpyAsync(s.d_tasks, s.h_tasks, n * sizeof(uint4), hipMemcpyHostToDevice, s.stream)
                     : hipSuccess;
    int grid = 0;
    bool apush = false;
    if (e == hipSuccess) {
        int bpc = 0;
        e = batch_fit(S->cfg.block, frames, 0, 0, &apush, &bpc);
        grid = std::max(1, S->cus * std::max(1, bpc));
        const uint64_t need = (n + S->cfg.block - 1) / S->cfg.block;
        if (static_cast<uint64_t>(grid) > need) grid = static_cast<int>(std::max<uint64_t>(1, need));
    }
    if (e == hipSuccess) {
        BatchArgs a{s.d_tasks, n, s.d_counts, std::max<uint64_t>(1, s.slices.size()), S->cfg.block, grid, frames,
                    S->cfg.chunk, s.d_state, apush};
        e = launch_batch(a, s.stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.h_counts, s.d_counts, s.slices.size() * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.h_state, s.d_state, sizeof(DevState), hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done_ev, s.stream);
    return e;
}

// A request is complete when every task is placed and every batch holding a slice is back.
static void maybe_finish(nq_server *S, Request &r, double now) {
    if (r.done || r.outstanding > 0 || r.next < r.tasks.size()) return;
    r.t_done = now;
    r.done = true;
    S->completed++;
    S->solutions += r.count;
}

static void scheduler(nq_server *S) {
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        bool progressed = false;
        // Retire finished slots.
        for (Slot &s : S->slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done_ev);
            if (q == hipErrorNotReady) continue;
            const double now = now_ms();
            if (q != hipSuccess) {
                S->error = NQ_ERR_HIP;
                S->error_msg = "serving batch failed";
            } else if (s.h_state->error) {
                // Admission sizes every request's frames, so this is a kernel fault, not a user error.
                S->error = NQ_ERR_STATE;
                S->error_msg = (s.h_state->error & DEV_ERR_LDS) ? "batch kernel LDS does not start at address 0"
                                                                : "a served task needed more stack frames than sized";
            }
            for (size_t i = 0; i < s.slices.size(); ++i) {
                Request &r = *s.slices[i].req;
                r.count += s.h_counts[i];
                r.outstanding--;
                maybe_finish(S, r, now);
            }
            S->device_ms += now - s.t_launch;
            s.slices.clear();
            s.busy = false;
            progressed = true;
            S->cv_done.notify_all();
        }
        // Cut and launch batches into free slots.
        for (Slot &s : S->slots) {
            if (s.busy || S->pending.empty()) continue;
            const double now = now_ms();
            plan_slot(S, s, now);
            s.busy = true;
            s.t_launch = now;
            hipError_t e = launch_slot(S, s);
            if (e != hipSuccess) {
                S->error = NQ_ERR_HIP;
                S->error_msg = hipGetErrorString(e);
                for (Slice &sl : s.slices) {
                    sl.req->outstanding--;
                    sl.req->done = true;
                    sl.req->t_done = now;
                }
                s.slices.clear();
                s.busy = false;
                S->cv_done.notify_all();
                continue;
            }
            S->batches++;
            progressed = true;
        }
        bool any_busy = false;
        for (Slot &s : S->slots) any_busy |= s.busy;
        if (S->stop && S->pending.empty() && !any_busy) break;
        if (!progressed) {
            if (any_busy)
                S->cv_work.wait_for(lk, std::chrono::microseconds(50));
            else
                S->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

extern "C" {

int nq_server_config_default(nq_server_config *cfg) {
    if (!cfg) return fail(NQ_ERR_ARG, "cfg is NULL");
    // Measured on the seeded Poisson mix at 68 requests/s: 12 free rows per task keeps every
    // batch's longest task short (p50 27 ms, p99 225 ms; 13 rows: 93 / 519 ms; 14 rows
    // overloads), at up to ~13 M tasks for an N=19 whole board, hence 16 M-task slots.
    cfg->slots = 3;
    cfg->slot_tasks = 1u << 24;
    cfg->slot_requests = 4096;
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
        cfg.block % 64 || cfg.chunk < 1)
        return fail(NQ_ERR_ARG, "invalid server configuration");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");
    auto *S = new nq_server();
    S->cfg = cfg;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, dev);
    S->slots.resize(cfg.slots);
    for (Slot &s : S->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&s.d_tasks, cfg.slot_tasks * sizeof(uint4));
        if (e == hipSuccess) e = hipHostMalloc(&s.h_tasks, cfg.slot_tasks * sizeof(uint4), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&s.d_counts, cfg.slot_requests * sizeof(uint64_t));
        if (e == hipSuccess) e = hipHostMalloc(&s.h_counts, cfg.slot_requests * sizeof(uint64_t), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&s.d_state, sizeof(DevState));
        if (e == hipSuccess) e = hipHostMalloc(&s.h_state, sizeof(DevState), hipHostMallocDefault);
    }
    if (e != hipSuccess) {
        for (Slot &s : S->slots) free_slot(s);
        delete S;
        return fail_hip(e, "server slot allocation");
    }
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
    req->frames = d0 + 1 < inst->N ? inst->N - 1 - d0 : 1;
    req->task_work = est_task_nodes(inst->N - d0);
    // A request that can never fit one slot is refused for good (NQ_ERR_NOMEM), unlike a full
    // queue (NQ_ERR_BUSY), which a retry may pass.
    if (req->tasks.size() > S->cfg.slot_tasks)
        return fail(NQ_ERR_NOMEM, "request larger than a slot's task budget (raise slot_tasks or depth_target)");
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) return fail(NQ_ERR_STATE, "server is closing");
    if (S->pending_tasks + req->tasks.size() > S->cfg.max_pending_tasks)
        return fail(NQ_ERR_BUSY, "pending task budget exhausted");
    req->ticket = S->next_ticket++;
    req->t_submit = now_ms();
    S->pending_tasks += req->tasks.size();
    S->all[req->ticket] = req;
    if (req->tasks.empty())
        maybe_finish(S, *req, req->t_submit);  // no legal continuation: count 0, done at once
    else
        S->pending.push_back(req);
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
    if (S->error) return fail(S->error, S->error_msg);
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
    if (batches) *batches = S->batches;
    if (solutions) *solutions = S->solutions;
    return NQ_OK;
}

int nq_server_timing(nq_server *S, double *device_ms, double *open_ms) {
    if (!S) return fail(NQ_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> g(S->mu);
    if (device_ms) *device_ms = S->device_ms;
    if (open_ms) *open_ms = now_ms() - S->t_open;
    return NQ_OK;
}

int nq_server_close(nq_server *S) {
    if (!S) return NQ_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->thr.joinable()) S->thr.join();
    for (Slot &s : S->slots) free_slot(s);
    const int err = S->error;
    delete S;
    return err ? fail(err, "server had a device error") : NQ_OK;
}

}  // extern "C"
