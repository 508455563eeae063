// CPU model of the serving path: the library's own nq_serve.cpp (scheduler, slicing, ring,
// slots, stop and restart), nq_batch.cpp and nq_prefix.cpp (request expansion), built with g++
// against a mock of the few HIP runtime calls they make.  The "kernel" is an emulation of
// serve_kernel's protocol (n_queens_amd/csrc/nq_dfs.hip) with one host thread per wave of 64
// lanes: tickets from a fetch-and-add on the cursor, a wave's claim loaded in one go below the
// tail of its parity, ring entries handed to waiting lanes in rank order, runs of one slot
// flushed with one packed fetch-and-add (count << SERVE_FIN_BITS | weight, the terminator
// weighing SERVE_FIN_DONE - tasks), the completing lane clearing the word and writing {count,
// generation}, idle waves leaving on the stop epoch.  Lanes finish their tasks after a random
// number of steps, so claims, completions and publication interleave differently in each run.
//
// Test infrastructure only (tests/test_serve_model.py builds and runs it; the expected counts
// come from the oracle there).  Usage:
//
//   serve_model SEED NREQ     prints one line per request: "N k c0 c1 ... = count"
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "nq.h"
#include "nq_kernels.h"

// ---- mock runtime: streams run their operations in order on a worker thread ----

struct ihipStream_t {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> ops;
    bool busy = false, quit = false;
    std::thread worker;
    ihipStream_t() {
        worker = std::thread([this] {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                cv.wait(lk, [&] { return quit || !ops.empty(); });
                if (ops.empty()) return;
                auto op = std::move(ops.front());
                ops.pop_front();
                busy = true;
                lk.unlock();
                op();
                lk.lock();
                busy = false;
                cv.notify_all();
            }
        });
    }
    void push(std::function<void()> op) {
        std::lock_guard<std::mutex> g(mu);
        ops.push_back(std::move(op));
        cv.notify_all();
    }
    void drain() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return ops.empty() && !busy; });
    }
    ~ihipStream_t() {
        {
            std::lock_guard<std::mutex> g(mu);
            quit = true;
            cv.notify_all();
        }
        worker.join();
    }
};

struct ihipEvent_t {
    std::atomic<uint64_t> recorded{0}, completed{0};
};

static const int MODEL_CUS = 2;         // compute units of the modelled device
static const int MODEL_WALL_KHZ = 1000;  // wall clock ticks: microseconds

static uint64_t wall_us() {
    return static_cast<uint64_t>(
        std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now().time_since_epoch())
            .count());
}

extern "C" {
hipError_t hipMalloc(void **p, size_t n) {
    *p = aligned_alloc(256, (n + 255) / 256 * 256);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipFree(void *p) {
    free(p);
    return hipSuccess;
}
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t attr, int) {
    *v = attr == hipDeviceAttributeMultiprocessorCount ? MODEL_CUS
         : attr == hipDeviceAttributeWallClockRate     ? MODEL_WALL_KHZ
                                                        : 0;
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int) {
    *s = new ihipStream_t();
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    s->drain();
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    s->drain();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) {
    *e = new ihipEvent_t();
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    const uint64_t k = ++e->recorded;
    s->push([e, k] { e->completed = k; });
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) { return e->completed == e->recorded ? hipSuccess : hipErrorNotReady; }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    memcpy(d, s, n);
    return hipSuccess;
}
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "modelled HIP error"; }
// Used only by the batching path, which the model does not run.
hipError_t hipEventElapsedTime(float *, hipEvent_t, hipEvent_t) { return hipErrorNotSupported; }
hipError_t hipMemcpyAsync(void *, const void *, size_t, hipMemcpyKind, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t) {
    return hipErrorNotSupported;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipErrorNotSupported; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipErrorNotSupported; }
int nq_config_default(int, int, nq_config *) { return NQ_ERR_STATE; }
int nq_count(int, const uint32_t *, uint64_t, const nq_config *, uint64_t *, uint64_t *, uint64_t *, nq_stats *) {
    return NQ_ERR_STATE;
}
}  // extern "C"

// ---- the library's helpers that live in sources the model does not build ----

namespace nq {

static std::string g_last;

int fail(int status, const std::string &msg) {
    g_last = msg;
    return status;
}
int fail_hip(int e, const char *what) { return fail(NQ_ERR_HIP, std::string(what) + ": " + std::to_string(e)); }
double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
hipError_t batch_expand(uint4 **, uint64_t *, uint32_t, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t batch_fit(int, int, int, int, bool *, int *) { return hipErrorNotSupported; }
hipError_t launch_batch(const BatchArgs &, hipStream_t) { return hipErrorNotSupported; }

}  // namespace nq

extern "C" const char *nq_last_error(void) { return nq::g_last.c_str(); }

namespace nq {

hipError_t launch_zero(unsigned long long *p, uint64_t n, hipStream_t s) {
    s->push([p, n] {
        for (uint64_t i = 0; i < n; ++i) __atomic_store_n(&p[i], 0ull, __ATOMIC_RELAXED);
    });
    return hipSuccess;
}

hipError_t serve_fit(int, int, int, bool *apush, int *blocks_per_cu) {
    *apush = false;
    *blocks_per_cu = 2;
    return hipSuccess;
}

// The completions below a task, as serve_kernel counts them: an edge prefix directly, else the
// DFS (a completed board counts 1).
static uint64_t dfs(uint32_t last, uint32_t c, uint32_t l, uint32_t r) {
    if (c == last) return 1;
    uint64_t s = 0;
    for (uint32_t v = last & ~(c | l | r); v; v &= v - 1) {
        const uint32_t p = v & (0u - v);
        s += dfs(last, c | p, (l | p) << 1, (r | p) >> 1);
    }
    return s;
}

namespace {

struct LaneState {
    bool own = false;
    int steps = 0;  // steps left before the lane's task is done
    uint4 task{};
    uint32_t weight = 1;
    uint32_t run_slot = 0;
    uint64_t run_sum = 0, run_n = 0;
};

void wave_main(const ServeArgs a, uint32_t gwave) {
    std::mt19937 rng(gwave * 7919u + a.epoch);
    LaneState L[64];
    uint4 buf[SERVE_BUF];
    uint64_t bfirst = 0, bend = 0, bload = 0, bnext = 0, lim = 0;
    uint64_t idle_since = 0;
    int idle_polls = 0;
    const unsigned long long *tailp = &a.ctl->tail[a.parity];
    auto flush = [&](LaneState &x) {
        if (!x.run_n) return;
        const uint64_t add = (x.run_sum << SERVE_FIN_BITS) | x.run_n;
        // Acquire-release here stands in for what orders the kernel's reads of a slice's ring
        // entries before the flush that may let the host reuse them: the flushed slot and counts
        // are computed from those reads, so they have completed.
        const uint64_t now = __atomic_fetch_add(&a.acc[x.run_slot], add, __ATOMIC_ACQ_REL) + add;
        if ((now & ((1ull << SERVE_FIN_BITS) - 1)) == SERVE_FIN_DONE) {
            __atomic_exchange_n(&a.acc[x.run_slot], 0ull, __ATOMIC_RELAXED);
            const uint64_t g = __atomic_fetch_add(&a.gen[x.run_slot], 1ull, __ATOMIC_RELAXED) + 1;
            __atomic_store_n(&a.done[x.run_slot].count, now >> SERVE_FIN_BITS, __ATOMIC_RELAXED);
            __atomic_store_n(&a.done[x.run_slot].gen, g, __ATOMIC_RELEASE);
        }
        x.run_sum = 0;
        x.run_n = 0;
    };
    for (;;) {
        // Busy lanes advance; a finished task joins the lane's run.
        for (LaneState &x : L) {
            if (!x.own || --x.steps > 0) continue;
            const uint32_t N = x.task.w & 31u, last = (1u << N) - 1u;
            x.run_sum += dfs(last, x.task.x, x.task.y, x.task.z) * x.weight;
            x.run_n += 1;
            x.own = false;
        }
        auto waiting = [&] {
            std::vector<int> w;
            for (int i = 0; i < 64; ++i)
                if (!L[i].own) w.push_back(i);
            return w;
        };
        std::vector<int> need = waiting();
        while (!need.empty()) {
            if (bnext == bload) {
                if (bload == bend) {
                    const uint64_t k = need.size();
                    const uint64_t h = __atomic_load_n(a.head, __ATOMIC_RELAXED);
                    const uint64_t take =
                        (lim > h && lim - h >= 64ull * SERVE_BUF) ? std::max<uint64_t>(k, a.chunk) : k;
                    const uint64_t first = __atomic_fetch_add(a.head, take, __ATOMIC_RELAXED);
                    const uint64_t rec = (static_cast<uint64_t>(a.epoch & 0xFFFFu) << SERVE_CLAIM_EPOCH_SHIFT) |
                                         ((first + take) & ((1ull << SERVE_CLAIM_EPOCH_SHIFT) - 1));
                    __atomic_store_n(&a.claimed[gwave], rec, __ATOMIC_RELAXED);
                    bfirst = bload = bnext = first;
                    bend = first + take;
                }
                if (bload >= lim) {
                    const uint64_t t = __atomic_load_n(tailp, __ATOMIC_ACQUIRE);
                    lim = t > a.base ? t - a.base : 0;
                }
                const uint64_t n = (lim > bload ? std::min(bend, lim) : bload) - bload;
                if (n == 0) break;
                for (uint64_t i = 0; i < n; ++i) buf[bload - bfirst + i] = a.ring[(a.base + bload + i) & a.ring_mask];
                bload += n;
            }
            const uint64_t avail = bload - bnext;
            for (uint64_t rank = 0; rank < need.size() && rank < avail; ++rank) {
                LaneState &x = L[need[rank]];
                const uint4 w = buf[bnext - bfirst + rank];
                const uint32_t s = w.w >> SERVE_TAG_SLOT_SHIFT;
                if (x.run_n && s != x.run_slot) flush(x);
                x.run_slot = s;
                const uint32_t N = w.w & 31u, last = (1u << N) - 1u;
                x.weight = 1u + ((w.w >> 5) & 1u);
                const uint32_t v0 = last & ~(w.x | w.y | w.z);
                const uint32_t d0 = __builtin_popcount(w.x);
                if (w.w & SERVE_TAG_TERM) {
                    x.run_n += SERVE_FIN_DONE - w.x;
                } else if (d0 + 1u >= N || v0 == 0) {
                    x.run_sum += (d0 >= N ? (w.x == last ? 1u : 0u) : (d0 + 1u == N ? __builtin_popcount(v0) : 0u)) *
                                 x.weight;
                    x.run_n += 1;
                } else if (N - 1u - d0 > static_cast<uint32_t>(a.stack_depth)) {
                    __atomic_fetch_or(a.error, static_cast<unsigned long long>(DEV_ERR_DEPTH), __ATOMIC_RELAXED);
                    x.run_n += 1;
                } else {
                    x.task = w;
                    x.own = true;
                    x.steps = 1 + static_cast<int>(rng() % 4);
                }
            }
            bnext += std::min<uint64_t>(avail, need.size());
            need = waiting();
        }
        for (LaneState &x : L)
            if (!x.own) flush(x);
        bool live = false;
        for (const LaneState &x : L) live |= x.own;
        if (!live) {
            if (__atomic_load_n(&a.ctl->stop_epoch, __ATOMIC_RELAXED) == a.epoch) break;
            const uint64_t now = wall_us();
            if (idle_polls == 0) idle_since = now;
            if (now - idle_since > a.orphan_ticks) break;
            std::this_thread::sleep_for(std::chrono::microseconds(5 << std::min(idle_polls, 4)));
            ++idle_polls;
            continue;
        }
        idle_polls = 0;
        if (rng() % 8 == 0) std::this_thread::yield();
    }
}

}  // namespace

hipError_t launch_serve(const ServeArgs &a, hipStream_t s) {
    if (a.chunk < 1 || a.chunk > SERVE_BUF) return hipErrorInvalidValue;
    s->push([a] {
        const uint32_t waves = static_cast<uint32_t>(a.grid) * static_cast<uint32_t>(a.block / 64);
        std::vector<std::thread> th;
        for (uint32_t w = 0; w < waves; ++w) th.emplace_back(wave_main, a, w);
        for (auto &t : th) t.join();
    });
    return hipSuccess;
}

}  // namespace nq

// ---- driver ----

int main(int argc, char **argv) {
    const unsigned seed = argc > 1 ? static_cast<unsigned>(atoi(argv[1])) : 1;
    const int nreq = argc > 2 ? atoi(argv[2]) : 200;
    std::mt19937 rng(seed);
    // Random requests: N in 4..11, a legal prefix of 0..4 rows (every 10th an empty board).
    struct Req {
        int N;
        std::vector<int> cols;
    };
    std::vector<Req> reqs;
    while (static_cast<int>(reqs.size()) < nreq) {
        Req q;
        q.N = 4 + static_cast<int>(rng() % 8);
        const int k = reqs.size() % 10 == 0 ? 0 : static_cast<int>(rng() % 5);
        uint32_t c = 0, l = 0, r = 0, last = (1u << q.N) - 1u;
        bool ok = true;
        for (int i = 0; i < k && ok; ++i) {
            const uint32_t v = last & ~(c | l | r);
            if (!v) {
                ok = false;
                break;
            }
            std::vector<int> opts;
            for (int j = 0; j < q.N; ++j)
                if (v >> j & 1u) opts.push_back(j);
            const int col = opts[rng() % opts.size()];
            q.cols.push_back(col);
            const uint32_t p = 1u << col;
            c |= p;
            l = (l | p) << 1;
            r = (r | p) >> 1;
        }
        if (ok) reqs.push_back(q);
    }
    // A small ring (wraps many times), few slots (reused), a small task budget (BUSY) and
    // 4..6 free rows per task (many tasks per request), so that every path of the scheduler runs.
    nq_server_config cfg;
    nq_server_config_default(&cfg);
    cfg.slot_tasks = 512;
    cfg.slot_requests = 6;
    cfg.max_pending_tasks = 20000;
    cfg.depth_target = 6;
    cfg.chunk = 16;
    cfg.batch_work = 50;
    nq_server *S = nullptr;
    if (nq_server_open(&cfg, &S) != NQ_OK) {
        fprintf(stderr, "open: %s\n", nq_last_error());
        return 2;
    }
    // Three phases, each fed by three threads and waited for, with an idle gap longer than the
    // server's idle time between them: the kernel instance stops and a new one (the other tail
    // parity, a new base) starts for the next phase.
    std::vector<uint64_t> tickets(reqs.size(), 0);
    std::atomic<int> busy{0};
    const size_t phases = 3;
    for (size_t ph = 0; ph < phases; ++ph) {
        const size_t lo = reqs.size() * ph / phases, hi = reqs.size() * (ph + 1) / phases;
        auto feeder = [&](int part) {
            std::mt19937 r2(seed * 31u + static_cast<unsigned>(part + 3 * ph));
            for (size_t i = lo + part; i < hi; i += 3) {
                nq_instance in{};
                in.N = reqs[i].N;
                in.k = static_cast<int>(reqs[i].cols.size());
                for (int j = 0; j < in.k; ++j) in.cols[j] = static_cast<int8_t>(reqs[i].cols[j]);
                for (;;) {
                    const int rc = nq_server_submit(S, &in, &tickets[i]);
                    if (rc == NQ_OK) break;
                    if (rc != NQ_ERR_BUSY) {
                        fprintf(stderr, "submit %zu: %d %s\n", i, rc, nq_last_error());
                        exit(3);
                    }
                    ++busy;
                    std::this_thread::sleep_for(std::chrono::microseconds(200));
                }
                if (r2() % 16 == 0) std::this_thread::sleep_for(std::chrono::microseconds(300));
            }
        };
        std::vector<std::thread> feeders;
        for (int p = 0; p < 3; ++p) feeders.emplace_back(feeder, p);
        for (auto &t : feeders) t.join();
        for (size_t i = lo; i < hi; ++i) {
            uint64_t count = 0;
            nq_request_stats st{};
            const int rc = nq_server_wait(S, tickets[i], 60000, &count, &st);
            if (rc != NQ_OK) {
                fprintf(stderr, "wait %zu: %d %s\n", i, rc, nq_last_error());
                return 4;
            }
            printf("%d %zu", reqs[i].N, reqs[i].cols.size());
            for (int c : reqs[i].cols) printf(" %d", c);
            printf(" = %llu\n", static_cast<unsigned long long>(count));
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
    }
    uint64_t completed = 0, slices = 0, solutions = 0;
    nq_server_stats(S, &completed, &slices, &solutions);
    const int rc = nq_server_close(S);
    fprintf(stderr, "completed %llu slices %llu busy %d close %d\n", static_cast<unsigned long long>(completed),
            static_cast<unsigned long long>(slices), busy.load(), rc);
    return rc == NQ_OK ? 0 : 5;
}
