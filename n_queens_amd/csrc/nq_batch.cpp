// Batched instances: many independent counting instances (different N, different fixed prefixes)
// counted in one persistent launch over a device task queue, with the launch sequence captured
// into a hipGraph and replayed.  No reference counterpart (the reference counts one whole board
// per process, src/main.cpp:51); each instance's count is exactly what the reference's n_queens
// (src/n_queens.cpp:8-21) returns for the instance's prefix masks.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

#include "nq.h"
#include "nq_internal.h"
#include "nq_kernels.h"
#include "nq_tasks.h"

namespace nq {

namespace {

struct Prefix {
    uint32_t c, l, r;
};

// Compose the fixed rows exactly as the recursion does (cur|p, (left|p)<<1, (right|p)>>1).
bool encode(const nq_instance &in, Prefix *out) {
    if (in.N < 1 || in.N > MAX_N || in.k < 0 || in.k > in.N || in.k > 32) return false;
    const uint32_t last = (1u << in.N) - 1u;
    uint32_t c = 0, l = 0, r = 0;
    for (int i = 0; i < in.k; ++i) {
        const int col = in.cols[i];
        if (col < 0 || col >= in.N) return false;
        const uint32_t p = 1u << col;
        if (!(p & last & ~(c | l | r))) return false;
        c |= p;
        l = (l | p) << 1;
        r = (r | p) >> 1;
    }
    *out = {c, l, r};
    return true;
}

// Rows to expand on the host so that no task needs more than `depth_target` levels.
int split_rows(int N, int k, int depth_target) { return std::max(0, (N - k) - depth_target); }

}  // namespace

int append_instance_tasks(int N, int k, uint32_t c, uint32_t l, uint32_t r, int depth_target, uint32_t idx,
                          TaskVec &out) {
    const int s = split_rows(N, k, depth_target);
    if (k == 0 && N >= 4) {
        // The reference's mirror-symmetric task list (generate_tasks' order), weight 2.
        const int rows = std::max(2, s);
        const uint32_t tag = (idx << 6) | (1u << 5) | static_cast<uint32_t>(N);
        const ExpandPart parts[2] = {{EXPAND_EVEN, 0, 0, 0}, {EXPAND_ODD, 0, 0, 0}};
        expand_tagged(N, parts, (N & 1) ? 2 : 1, rows, tag, out);
        return rows;
    }
    const ExpandPart part{EXPAND_PLAIN, c, l, r};
    expand_tagged(N, &part, 1, s, (idx << 6) | static_cast<uint32_t>(N), out);
    return k + s;
}

double est_nodes_per_task(int N, int k, uint32_t c, uint32_t l, uint32_t r, int split) {
    constexpr int PROBES = 64;
    const int m = N - k;  // free rows
    if (m <= 0) return 0;
    split = std::min(std::max(split, 0), m);
    const uint32_t last = (1u << N) - 1u;
    // xorshift64 (Marsaglia 2003) seeded from the prefix words (never zero)
    uint64_t s = ((static_cast<uint64_t>(c) << 32 | l) ^ (static_cast<uint64_t>(r) << 7) ^ static_cast<uint64_t>(N)) |
                 (1ull << 63);
    auto rnd = [&s]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return static_cast<uint32_t>(s >> 32);
    };
    // est[j]: estimated nodes j rows below the prefix, summed over the probes.
    double est[33] = {0};
    for (int p = 0; p < PROBES; ++p) {
        uint32_t pc = c, pl = l, pr = r;
        double w = 1;
        for (int j = 0; j <= m; ++j) {
            est[j] += w;
            if (j == m) break;
            uint32_t v = last & ~(pc | pl | pr);
            const int b = __builtin_popcount(v);
            if (b == 0) break;
            w *= b;
            // a uniform candidate: the i-th set bit, i = floor(rnd * b / 2^32)
            for (uint32_t i = static_cast<uint32_t>((static_cast<uint64_t>(rnd()) * b) >> 32); i; --i) v &= v - 1;
            const uint32_t q = v & (0u - v);
            pc |= q;
            pl = (pl | q) << 1;
            pr = (pr | q) >> 1;
        }
    }
    if (est[split] <= 0) return 0;
    // DFS nodes of one task: placements below its root, up to the last row (a full board is
    // not a node).
    double below = 0;
    for (int j = split + 1; j < m; ++j) below += est[j];
    return below / est[split];
}

bool encode_instance(const nq_instance &in, uint32_t *c, uint32_t *l, uint32_t *r) {
    Prefix p;
    if (!encode(in, &p)) return false;
    *c = p.c;
    *l = p.l;
    *r = p.r;
    return true;
}

// Validate instances and build the seeds of the device task list, largest expected subtrees
// first: each instance's prefix, or for an empty board (N >= 4) the reference's
// mirror-symmetric tasks at depth 2 (weight 2).  batch_expand then expands every seed to
// `depth_target` free rows on the device, in the order append_instance_tasks would produce on
// the host.  *max_frames is the deepest DFS stack any final task needs, *levels the most rows
// any seed is expanded by.
int build_batch_seeds(const nq_instance *inst, int n, int depth_target, std::vector<uint4> &seeds, int *max_frames,
                      int *levels) {
    std::vector<Prefix> pre(n);
    for (int i = 0; i < n; ++i) {
        if (!encode(inst[i], &pre[i])) {
            char buf[128];
            snprintf(buf, sizeof buf, "instance %d: invalid N/k or attacking prefix", i);
            return fail(NQ_ERR_ARG, buf);
        }
    }
    // Order instances by estimated DFS nodes per task, descending: the longest tasks start
    // first and the shortest fill the tail.  (Free rows alone are a poor proxy: a 12-row task
    // of N=14 holds about 13 times the nodes of a 12-row task of N=20, and a lane that claims
    // one of those last runs alone for milliseconds.)
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    // About 15 us per instance of 20 free rows: large batches spread them over threads.
    std::vector<double> est(n);
    std::atomic<int> next_est{0};
    auto estimate = [&] {
        for (int i = next_est++; i < n; i = next_est++) {
            const int N = inst[i].N, k = inst[i].k;
            const int s = split_rows(N, k, depth_target);
            est[i] = est_nodes_per_task(N, k, pre[i].c, pre[i].l, pre[i].r, k == 0 && N >= 4 ? std::max(2, s) : s);
        }
    };
    const int threads = n >= 64 ? static_cast<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency()))) : 1;
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(estimate);
    estimate();
    for (auto &t : pool) t.join();
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return est[a] > est[b]; });
    seeds.clear();
    int frames = 1, lev = 0;
    for (int i : order) {
        const int N = inst[i].N, k = inst[i].k;
        const int s = split_rows(N, k, depth_target);
        int seed_depth = k, d0 = k + s;
        if (k == 0 && N >= 4) {
            std::vector<uint32_t> t;
            generate_tasks(N, 2, t);
            const uint32_t tag = (static_cast<uint32_t>(i) << 6) | (1u << 5) | static_cast<uint32_t>(N);
            for (size_t j = 0; j + 2 < t.size(); j += 3) seeds.push_back({t[j], t[j + 1], t[j + 2], tag});
            seed_depth = 2;
            d0 = std::max(2, s);
        } else {
            seeds.push_back({pre[i].c, pre[i].l, pre[i].r, (static_cast<uint32_t>(i) << 6) | static_cast<uint32_t>(N)});
        }
        lev = std::max(lev, d0 - seed_depth);
        if (d0 + 1 < N) frames = std::max(frames, N - 1 - d0);
    }
    *max_frames = frames;
    *levels = lev;
    return NQ_OK;
}

int instance_tasks(const nq_instance &in, int depth_target, std::vector<uint32_t> &tasks, int *weight) {
    Prefix p;
    if (!encode(in, &p)) return fail(NQ_ERR_ARG, "invalid N/k or attacking prefix");
    TaskVec t;
    append_instance_tasks(in.N, in.k, p.c, p.l, p.r, depth_target, 0, t);
    tasks.resize(3 * t.size());
    for (size_t i = 0; i < t.size(); ++i) {
        tasks[3 * i] = t[i].x;
        tasks[3 * i + 1] = t[i].y;
        tasks[3 * i + 2] = t[i].z;
    }
    *weight = (!t.empty() && (t[0].w & (1u << 5))) ? 2 : 1;
    return NQ_OK;
}

}  // namespace nq

using namespace nq;

#define NQ_HIP(call)                                             \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return fail_hip(e_, #call);        \
    } while (0)

struct nq_batch_plan {
    int device = 0;
    uint64_t ninst = 0, ntasks = 0;
    uint4 *d_tasks = nullptr;
    uint64_t *d_counts = nullptr;
    DevState *d_state = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    nq_config cfg{};
    double expand_ms = 0;
};

static void plan_free(nq_batch_plan *p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_tasks) (void)hipFree(p->d_tasks);
    if (p->d_counts) (void)hipFree(p->d_counts);
    if (p->d_state) (void)hipFree(p->d_state);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

extern "C" {

int nq_batch_plan_create(const nq_instance *inst, int n, const nq_config *cfg, nq_batch_plan **out) {
    if (!out) return fail(NQ_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n < 0 || (n > 0 && !inst)) return fail(NQ_ERR_ARG, "bad instance array");
    // A batched task carries its instance index in tag bits 6..31.
    if (n > (1 << 26)) return fail(NQ_ERR_ARG, "at most 2^26 instances per batch");
    const double t0 = now_ms();
    nq_config c;
    nq_config_default(MAX_N, 0, &c);
    c.block = 128;  // batches keep 128-lane workgroups
    c.stack_depth = 0;
    if (cfg) {
        if (cfg->block) c.block = cfg->block;
        if (cfg->waves_per_cu) c.waves_per_cu = cfg->waves_per_cu;
        if (cfg->chunk) c.chunk = cfg->chunk;
        if (cfg->grid) c.grid = cfg->grid;
        if (cfg->stack_depth) c.stack_depth = cfg->stack_depth;
        c.push = cfg->push;
    }
    if (c.push < 0 || c.push > NQ_PUSH_ALWAYS) return fail(NQ_ERR_ARG, "push must be 0, 1 or 2");
    if (c.block < 64 || c.block > MAX_BLOCK || c.block % 64) return fail(NQ_ERR_ARG, "block must be 64..256, multiple of 64");
    if (c.chunk < 1 || c.chunk > 4096) return fail(NQ_ERR_ARG, "chunk must be 1..4096");
    // Per-task depth target: the stack depth + 1 when given, else 12 free rows (11 frames).
    const int depth_target = c.stack_depth ? c.stack_depth + 1 : 12;
    std::vector<uint4> tasks;  // seeds; expanded on the device below
    int frames = 1, levels = 0;
    int rc = build_batch_seeds(inst, n, depth_target, tasks, &frames, &levels);
    if (rc) return rc;
    if (!c.stack_depth) c.stack_depth = frames;
    if (c.stack_depth < frames || c.stack_depth > 32) return fail(NQ_ERR_DEPTH, "stack depth out of range for the batch");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");

    auto *p = new nq_batch_plan();
    p->ninst = static_cast<uint64_t>(n);
    p->ntasks = tasks.size();
    hipError_t e = hipGetDevice(&p->device);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&p->ev0);
    if (e == hipSuccess) e = hipEventCreate(&p->ev1);
    if (e == hipSuccess) e = hipMalloc(&p->d_tasks, std::max<size_t>(1, tasks.size()) * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&p->d_counts, std::max<uint64_t>(1, p->ninst) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&p->d_state, sizeof(DevState));
    if (e == hipSuccess && !tasks.empty())
        e = hipMemcpyAsync(p->d_tasks, tasks.data(), tasks.size() * sizeof(uint4), hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e == hipSuccess) e = batch_expand(&p->d_tasks, &p->ntasks, static_cast<uint32_t>(depth_target), levels, p->stream);
    if (e != hipSuccess) {
        plan_free(p);
        return fail_hip(e, "batch plan allocation/upload/expansion");
    }
    p->expand_ms = now_ms() - t0;
    bool apush = false;
    int bpc = 0;
    e = batch_fit(c.block, c.stack_depth, c.waves_per_cu, c.push, &apush, &bpc);
    if (e != hipSuccess || bpc <= 0) {
        plan_free(p);
        return e != hipSuccess ? fail_hip(e, "batch_occupancy") : fail(NQ_ERR_ARG, "batch does not fit on a CU");
    }
    c.push = apush ? NQ_PUSH_ALWAYS : NQ_PUSH_MASKED;
    if (c.grid == 0) {
        c.grid = cus * bpc;
        const uint64_t need = (p->ntasks + c.block - 1) / c.block;
        if (static_cast<uint64_t>(c.grid) > need) c.grid = static_cast<int>(std::max<uint64_t>(1, need));
    }
    p->cfg = c;
    // Capture zero-counts + reset + batched DFS into a graph (kernel nodes only).
    BatchArgs a{p->d_tasks, p->ntasks, p->d_counts, std::max<uint64_t>(1, p->ninst), c.block, c.grid, c.stack_depth,
                c.chunk, p->d_state, apush};
    e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = launch_batch(a, p->stream);
        e = hipStreamEndCapture(p->stream, &p->graph);
        if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        plan_free(p);
        return fail_hip(e, "batch graph capture");
    }
    *out = p;
    return NQ_OK;
}

int nq_batch_plan_launch(nq_batch_plan *p) {
    if (!p) return fail(NQ_ERR_ARG, "plan is NULL");
    NQ_HIP(hipEventRecord(p->ev0, p->stream));
    NQ_HIP(hipGraphLaunch(p->exec, p->stream));
    NQ_HIP(hipEventRecord(p->ev1, p->stream));
    return NQ_OK;
}

int nq_batch_plan_wait(nq_batch_plan *p, uint64_t *counts, nq_batch_stats *stats) {
    if (!p) return fail(NQ_ERR_ARG, "plan is NULL");
    DevState st;
    if (counts && p->ninst)
        NQ_HIP(hipMemcpyAsync(counts, p->d_counts, p->ninst * sizeof(uint64_t), hipMemcpyDeviceToHost, p->stream));
    NQ_HIP(hipMemcpyAsync(&st, p->d_state, sizeof st, hipMemcpyDeviceToHost, p->stream));
    NQ_HIP(hipStreamSynchronize(p->stream));
    if (st.error & DEV_ERR_LDS) return fail(NQ_ERR_STATE, "batch kernel LDS does not start at address 0");
    if (st.error) return fail(NQ_ERR_DEPTH, "a batched task needed more stack frames than configured");
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
        stats->ntasks = p->ntasks;
        stats->expand_ms = p->expand_ms;
        stats->kernel_ms = ms;
        stats->total = st.total;
    }
    return NQ_OK;
}

int nq_batch_plan_run(nq_batch_plan *p, uint64_t *counts, nq_batch_stats *stats) {
    int rc = nq_batch_plan_launch(p);
    if (rc) return rc;
    return nq_batch_plan_wait(p, counts, stats);
}

int nq_batch_plan_destroy(nq_batch_plan *p) {
    plan_free(p);
    return NQ_OK;
}

int nq_batch_count(const nq_instance *inst, int n, const nq_config *cfg, uint64_t *counts, nq_batch_stats *stats) {
    const double t0 = now_ms();
    nq_batch_plan *p = nullptr;
    int rc = nq_batch_plan_create(inst, n, cfg, &p);
    if (rc) return rc;
    rc = nq_batch_plan_run(p, counts, stats);
    plan_free(p);
    if (!rc && stats) stats->total_ms = now_ms() - t0;
    return rc;
}

int nq_instance_count(const nq_instance *inst, const nq_config *cfg, uint64_t *count) {
    if (!inst || !count) return fail(NQ_ERR_ARG, "NULL argument");
    std::vector<uint32_t> tasks;
    int weight = 1;
    const int depth_target = (cfg && cfg->stack_depth) ? cfg->stack_depth + 1 : 12;
    int rc = instance_tasks(*inst, depth_target, tasks, &weight);
    if (rc) return rc;
    uint64_t sum = 0;
    rc = nq_count(inst->N, tasks.data(), tasks.size() / 3, cfg, &sum, nullptr, nullptr, nullptr);
    if (rc) return rc;
    *count = sum * static_cast<uint64_t>(weight);
    return NQ_OK;
}

}  // extern "C"
