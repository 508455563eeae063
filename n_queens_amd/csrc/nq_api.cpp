// C ABI of libnq_hip.so (include/nq.h): device management, the counting entry points and the
// whole reference pipeline.  Replaces the host driver cuda_n_queens
// (src/n_queens_cuda.cu:84-186) for one device.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "nq.h"
#include "nq_internal.h"
#include "nq_kernels.h"

namespace nq {

namespace {
thread_local std::string g_err;

struct Device {
    int id = -1;
    int cus = 0;
    int clock_khz = 0;
    std::string name;
    DevState *state = nullptr;
    uint64_t *result = nullptr;  // [2]
    uint32_t *tasks = nullptr;
    size_t tasks_cap = 0;  // in tasks
    uint64_t *per_task = nullptr;
    size_t per_task_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int wall_khz = 0;  // rate of the kernel's wall_clock64 counter
    uint64_t *trace = nullptr;
    size_t trace_cap = 0;  // in uint64
    uint32_t *xbuf[2] = {nullptr, nullptr};  // device-side expansion levels (ping-pong)
    size_t xcap[2] = {0, 0};                 // in tasks
    unsigned long long *xcount = nullptr;    // [0], [1] task counts of the two buffers, [2] flags,
                                             // [3] child count of the next level (exact sizing)
    std::mutex mu;
};

std::mutex g_devs_mu;
std::vector<std::unique_ptr<Device>> g_devs;

}  // namespace

int fail(int status, const std::string &msg) {
    g_err = msg;
    return status;
}

int fail_hip(int e, const char *what) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: hipError %d (%s)", what, e, hipGetErrorString(static_cast<hipError_t>(e)));
    g_err = buf;
    return NQ_ERR_HIP;
}

double now_ms() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

#define NQ_HIP(call)                                             \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return fail_hip(e_, #call);        \
    } while (0)

// The calling thread's current device, initialised on first use.
static int current_device(Device **out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");
    int id = 0;
    NQ_HIP(hipGetDevice(&id));
    std::lock_guard<std::mutex> g(g_devs_mu);
    if (g_devs.size() < static_cast<size_t>(n)) g_devs.resize(n);
    if (!g_devs[id]) {
        auto d = std::make_unique<Device>();
        d->id = id;
        hipDeviceProp_t prop;
        NQ_HIP(hipGetDeviceProperties(&prop, id));
        d->cus = prop.multiProcessorCount;
        d->clock_khz = prop.clockRate;
        d->name = prop.gcnArchName;
        NQ_HIP(hipMalloc(&d->state, sizeof(DevState)));
        NQ_HIP(hipMalloc(&d->result, 3 * sizeof(uint64_t)));
        NQ_HIP(hipMalloc(&d->xcount, 4 * sizeof(unsigned long long)));
        NQ_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        NQ_HIP(hipEventCreate(&d->ev0));
        NQ_HIP(hipEventCreate(&d->ev1));
        NQ_HIP(hipDeviceGetAttribute(&d->wall_khz, hipDeviceAttributeWallClockRate, id));
        g_devs[id] = std::move(d);
    }
    *out = g_devs[id].get();
    return NQ_OK;
}

// gcnArchName of the current device (the tuner cache's key); false without a device.
static bool current_arch(char *out, size_t len) {
    int n = 0, id = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || hipGetDevice(&id) != hipSuccess) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, id) != hipSuccess) return false;
    snprintf(out, len, "%s", prop.gcnArchName);
    return out[0] != 0;
}

static int check_cfg_shape(const nq_config &c) {
    if (c.block != 0 && (c.block < 64 || c.block > MAX_BLOCK || c.block % 64))
        return fail(NQ_ERR_ARG, "block must be 0 (automatic) or 64..256, multiple of 64");
    if (c.chunk < 1 || c.chunk > 4096) return fail(NQ_ERR_ARG, "chunk must be 1..4096");
    if (c.stack_depth < 1 || c.stack_depth > 32) return fail(NQ_ERR_ARG, "stack_depth must be 1..32");
    if (c.waves_per_cu < 0 || c.grid < 0) return fail(NQ_ERR_ARG, "negative waves_per_cu/grid");
    if (c.push < 0 || c.push > NQ_PUSH_ALWAYS) return fail(NQ_ERR_ARG, "push must be 0, 1 or 2");
    return NQ_OK;
}

static int auto_split(int N, int rows);

// Frames a prefix at depth d0 needs: one per depth d0..N-2 that can hold remaining candidates.
static int frames_needed(int N, int d0) { return d0 + 1 >= N ? 0 : N - 1 - d0; }

// Resolve defaults and the persistent grid for (N, shallowest DFS root depth d0, ntasks).  With
// split > 0 (plain totals) the kernel is split_kernel and d0 is the depth of the roots it walks
// to, split rows below the tasks.
static int resolve_config(int cus, int N, int d0, uint64_t ntasks, const nq_config *in, bool nodes,
                          bool per_task, int split, nq_config *out) {
    nq_config c;
    nq_config_default(N, d0, &c);
    if (in) {
        if (in->block) c.block = in->block;
        if (in->stack_depth) c.stack_depth = in->stack_depth;
        if (in->waves_per_cu) c.waves_per_cu = in->waves_per_cu;
        if (in->chunk) c.chunk = in->chunk;
        if (in->grid) c.grid = in->grid;
        c.split = in->split;
        c.push = in->push;
    }
    // With the split a wave claims whole tasks: one at a time unless told otherwise, so that no
    // wave holds claimed tasks back while others run dry at the end.
    if (split > 0 && !(nodes && per_task) && !(in && in->chunk)) c.chunk = 1;
    int rc = check_cfg_shape(c);
    if (rc) return rc;
    if (c.stack_depth < frames_needed(N, d0)) {
        char buf[160];
        snprintf(buf, sizeof buf, "stack depth %d too small: N=%d with prefix depth %d needs %d frames", c.stack_depth,
                 N, d0, frames_needed(N, d0));
        return fail(NQ_ERR_DEPTH, buf);
    }
    // Resident workgroups per CU of a workgroup size for both push forms (capped by the
    // waves_per_cu target); the unconditional push is taken when its spare stack level costs
    // none of them.
    auto fit = [&](int block, int *push, int *bpc) -> int {
        int occ[2] = {0, 0};
        NQ_HIP(dfs_occupancy(block, c.stack_depth, nodes, per_task, false, split, &occ[0]));
        NQ_HIP(dfs_occupancy(block, c.stack_depth, nodes, per_task, true, split, &occ[1]));
        if (c.waves_per_cu > 0)
            for (int &o : occ) o = std::min(o, std::max(1, c.waves_per_cu / (block / 64)));
        *push = c.push ? c.push : (prefer_always_push(occ[0], occ[1]) ? NQ_PUSH_ALWAYS : NQ_PUSH_MASKED);
        *bpc = occ[*push == NQ_PUSH_ALWAYS ? 1 : 0];
        return NQ_OK;
    };
    // Automatic workgroup size: 256 lanes (one wave of each workgroup per SIMD) when that keeps
    // the resident waves 128 lanes get, else 128.  With equal waves 256 is about 0.5% faster
    // (N=20 and N=21, 6 rows, 12 waves/CU, unconditional push); where 128 fits an odd number
    // of workgroups, 256 would lose waves (N=21 after a 3-row split, masked push: 14 against 12).
    const bool auto_block = c.block == 0;
    if (auto_block) c.block = 128;
    int push = 0, bpc = 0;
    rc = fit(c.block, &push, &bpc);
    if (rc) return rc;
    if (auto_block) {
        int push256 = 0, bpc256 = 0;
        rc = fit(256, &push256, &bpc256);
        if (rc) return rc;
        if (bpc256 > 0 && bpc256 * 4 >= bpc * 2) {
            c.block = 256;
            push = push256;
            bpc = bpc256;
        }
    }
    c.push = push;
    if (bpc <= 0) return fail(NQ_ERR_ARG, "configuration does not fit on a CU (LDS)");
    if (c.grid == 0) c.grid = cus * bpc;
    // Workgroups the tasks can keep busy: one lane per task, or with the split one wave per task
    // (a wave walks a claimed task's roots on its 64 lanes).
    const uint64_t per_wg = split > 0 && !(nodes && per_task) ? static_cast<uint64_t>(c.block / 64) : c.block;
    const uint64_t wgs_needed = (ntasks + per_wg - 1) / per_wg;
    if (static_cast<uint64_t>(c.grid) > wgs_needed) c.grid = static_cast<int>(std::max<uint64_t>(1, wgs_needed));
    c.split = split;
    *out = c;
    return NQ_OK;
}

// Rows the DFS kernel walks in registers below tasks of depth d0: cfg->split (a negative value or
// no cfg: the automatic rule), never more than leaves the DFS two rows, at most MAX_SPLIT.
static int effective_split(int N, int d0, const nq_config *cfg, int *split) {
    int s = (cfg && cfg->split >= 0) ? cfg->split : auto_split(N, d0);
    if (s > MAX_SPLIT) return fail(NQ_ERR_ARG, "split must be -1 (automatic) or 0..8");
    *split = std::max(0, std::min(s, N - 2 - d0));
    return NQ_OK;
}

// The launch nq_count makes for a task list (shallowest depth d0) on a device with `cus` compute
// units: the split (cfg NULL or split -1: automatic) and the resolved configuration, for callers
// that bring their own device memory (nq_static.hip).  Needs no allocation.
int plan_count(int cus, int N, int d0, uint64_t ntasks, const nq_config *cfg, bool nodes, bool per_task,
               nq_config *out) {
    nq_config in = cfg ? *cfg : nq_config{};
    if (!cfg) in.split = -1;
    int split = 0;
    if (!(nodes && per_task)) {
        int rc = effective_split(N, d0, &in, &split);
        if (rc) return rc;
    }
    if (in.stack_depth) in.stack_depth = std::max(1, in.stack_depth - split);
    return resolve_config(cus, N, std::min(N, d0 + split), ntasks, &in, nodes, per_task, split, out);
}

int task_min_depth(int N, const uint32_t *tasks, uint64_t ntasks) {
    int d0 = N;
    for (uint64_t i = 0; i < ntasks; ++i) d0 = std::min(d0, __builtin_popcount(tasks[3 * i]));
    return d0;
}

static int ensure_tasks(Device *d, uint64_t n) {
    if (d->tasks_cap >= n && d->tasks) return NQ_OK;
    if (d->tasks) (void)hipFree(d->tasks);
    d->tasks = nullptr;
    d->tasks_cap = 0;
    const size_t cap = std::max<uint64_t>(n, 1);
    NQ_HIP(hipMalloc(&d->tasks, cap * 3 * sizeof(uint32_t)));
    d->tasks_cap = cap;
    return NQ_OK;
}

static int ensure_per_task(Device *d, uint64_t n) {
    if (d->per_task_cap >= n && d->per_task) return NQ_OK;
    if (d->per_task) (void)hipFree(d->per_task);
    d->per_task = nullptr;
    d->per_task_cap = 0;
    const size_t cap = std::max<uint64_t>(n, 1);
    NQ_HIP(hipMalloc(&d->per_task, cap * sizeof(uint64_t)));
    d->per_task_cap = cap;
    return NQ_OK;
}

static int ensure_trace(Device *d, size_t n) {
    if (d->trace_cap >= n && d->trace) return NQ_OK;
    if (d->trace) (void)hipFree(d->trace);
    d->trace = nullptr;
    d->trace_cap = 0;
    NQ_HIP(hipMalloc(&d->trace, n * sizeof(uint64_t)));
    d->trace_cap = n;
    return NQ_OK;
}

// Per-wave {start, end} stamps -> fraction of wave-time busy and the drain tail (ms).
static void trace_stats(const std::vector<uint64_t> &tr, int wall_khz, double *busy, double *tail_ms) {
    const size_t w = tr.size() / 2;
    if (!w || wall_khz <= 0) {
        *busy = 0;
        *tail_ms = 0;
        return;
    }
    uint64_t t0 = UINT64_MAX, t1 = 0, first_end = UINT64_MAX;
    double sum = 0;
    for (size_t i = 0; i < w; ++i) {
        t0 = std::min(t0, tr[2 * i]);
        t1 = std::max(t1, tr[2 * i + 1]);
        first_end = std::min(first_end, tr[2 * i + 1]);
        sum += static_cast<double>(tr[2 * i + 1] - tr[2 * i]);
    }
    const double span = static_cast<double>(t1 - t0);
    *busy = span > 0 ? sum / (span * static_cast<double>(w)) : 1.0;
    *tail_ms = static_cast<double>(t1 - first_end) / wall_khz;
}

static int ensure_xbuf(Device *d, int i, uint64_t n) {
    if (d->xcap[i] >= n && d->xbuf[i]) return NQ_OK;
    if (d->xbuf[i]) (void)hipFree(d->xbuf[i]);
    d->xbuf[i] = nullptr;
    d->xcap[i] = 0;
    NQ_HIP(hipMalloc(&d->xbuf[i], std::max<uint64_t>(n, 1) * 3 * sizeof(uint32_t)));
    d->xcap[i] = n;
    return NQ_OK;
}

// Device expansion budget: a level is expanded only if its buffer (12 B per task) fits in 40 %
// of the device memory free at the time, and never beyond 8 G tasks.
static uint64_t max_expanded_tasks() {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 2ull << 30;
    return std::min<uint64_t>(8ull << 30, static_cast<uint64_t>(free_b * 0.4) / 12);
}

// Enqueue up to `split` one-row expansions of (tasks, n) on the device (a full board is carried
// over; order is free).  The input count is n, or min(n, *n_dev_in) when n_dev_in is set.  The
// child count of every level is reduced on the device and read back first (one stream
// synchronization per row) so that each level's buffer has its exact size (the worst case,
// every task having N - depth children, is ~10x the real count per row); a level larger than
// max_expanded_tasks() ends the expansion there.  Level j writes expansion buffer
// (first + j) & 1, so a second call can continue from the first one's output.  Rows are
// expanded while the DFS keeps two rows to place (to_full: while the depth is below N).  The
// caller zeroes the overflow flag (xcount[2]).  On return *tasks_out / *n_dev / *bound describe
// the expanded list (*bound is its exact size) and *levels how many rows were expanded.
static int enqueue_split(Device *d, int N, int d0, const uint32_t *tasks, const unsigned long long *n_dev_in,
                         uint64_t n, int split, int first, bool to_full, hipStream_t s, const uint32_t **tasks_out,
                         const unsigned long long **n_dev, uint64_t *bound, int *levels) {
    const uint32_t *in = tasks;
    const unsigned long long *in_dev = n_dev_in;
    uint64_t b = n;
    int j = 0;
    for (; j < split && (to_full ? d0 + j < N : d0 + j + 2 < N); ++j) {
        const int buf = (first + j) & 1;
        hipError_t e = launch_child_count(in, b, in_dev, N, d->xcount + 3, s);
        if (e != hipSuccess) return fail_hip(e, "launch_child_count");
        unsigned long long nb = 0;
        NQ_HIP(hipMemcpyAsync(&nb, d->xcount + 3, sizeof nb, hipMemcpyDeviceToHost, s));
        NQ_HIP(hipStreamSynchronize(s));
        if (nb > max_expanded_tasks()) break;
        int rc = ensure_xbuf(d, buf, nb);
        if (rc) return rc;
        e = launch_expand(in, b, in_dev, N, d->xbuf[buf], nb, &d->xcount[buf], &d->xcount[2], s);
        if (e != hipSuccess) return fail_hip(e, "launch_expand");
        in = d->xbuf[buf];
        in_dev = &d->xcount[buf];
        b = nb;
    }
    *tasks_out = in;
    *n_dev = in_dev;
    *bound = b;
    *levels = j;
    return NQ_OK;
}


// Rows walked on the device below each task (0..MAX_SPLIT): the DFS then starts dfs_rows(N) rows
// above the last row, i.e. with dfs_rows(N) - 1 LDS frames.  Measured on one MI355X (split + DFS,
// final loop): N=16/4 -> 2 (1.1 ms), N=17/5 -> 1 or 2 (4.0 / 4.1 ms), N=18/6 -> 1 (18.0 ms; 0:
// 25.0, 2: 31.8), N=19/6 -> 2 (112.5 ms; 1: 118.2), N=20/6 -> 2 (881.6 ms; 1: 894.4, 3: 890.3),
// N=21/6 -> 3 (6,297 ms; 2: 6,335 ms; before the walk kept its frames in lanes 2 won, 7.36 s
// against 7.47 s), N=22/6 -> 3 (65.5 s; 2: 69.8 s), N=23/6 -> 4 (two slices of
// 600,000 tasks: 21.8 / 35.3 s; 3: 22.4 / 36.4 s), N=24/6 -> 5 (100,000 tasks: 27.1 s; 4: 27.8 s),
// N=25/6 -> 6 (30,000 tasks: 104.0 s; 4: 113.2 s, 10 waves/CU).  Up to N=20 shorter tasks (the
// drain tail) decide, and (N + 4) / 2 rows fits.  From N=22 on, 12 frames win: the unconditional
// push then keeps 12 waves/CU in 256-lane workgroups, while 13 frames leave the masked push at 12
// waves (about 3-7% slower), 14 and more fewer waves, and 11 frames (the masked push at 14 waves/CU
// in 128-lane workgroups, on roots with a sixth of the work) lose 0.5 % at N=22 (1M tasks, 4 rows
// walked: 9.21 s against 9.16), 2.8 % at N=23 and 6.6 % at N=24.  At N=21 the 11 frames win.
static int dfs_rows(int N) { return N <= 20 ? (N + 4) / 2 : N == 21 ? 12 : 13; }

static int auto_split(int N, int rows) {
    return std::max(0, std::min(MAX_SPLIT, N - rows - dfs_rows(N)));
}


// Whole-board count with the task list generated on the device.  The reference's tasks at
// depth `rows` (src/n_queens.cpp:74-113) are every legal continuation of its tasks at depth
// min(rows, 2): only row 0 (low half) and, for odd N, rows 0-1 (middle column, then the low
// side) are restricted.  The host generates those few seeds; the device expands them to depth
// `rows` (exact sizes, order free: the same multiset of tasks) and counts them, the DFS kernel
// walking the split rows below each task in registers.  on_generated (nullable) is called once
// the depth-`rows` list exists, with its task count and the generation time.
int solve_on_device(int N, int rows, const nq_config *cfg, uint64_t *total, nq_stats *stats,
                    void (*on_generated)(void *ctx, uint64_t ntasks, double ms), void *ctx) {
    const double t0 = now_ms();
    if (!total) return fail(NQ_ERR_ARG, "total is NULL");
    if (N < 1 || N > MAX_N) return fail(NQ_ERR_ARG, "N must be 1..28 for the GPU kernel");
    if (rows < 1 || rows > N) return fail(NQ_ERR_ARG, "rows must be 1..N");
    Device *d = nullptr;
    int rc = current_device(&d);
    if (rc) return rc;
    const int seed_rows = std::min(rows, 2);
    std::vector<uint32_t> seeds;
    generate_tasks(N, seed_rows, seeds);
    const uint64_t nseeds = seeds.size() / 3;
    // Without a configuration: the tuned one for N on this device if the tuner's cache has it,
    // else the built-in rule (nq_solve_config).
    nq_config in = cfg ? *cfg : nq_config{};
    if (!cfg) {
        rc = nq_solve_config(N, rows, &in);
        if (rc) return rc;
    }
    const int preplace = std::min(N, std::max(rows, in.preplace));

    std::lock_guard<std::mutex> g(d->mu);
    rc = ensure_tasks(d, nseeds);
    if (rc) return rc;
    const double t1 = now_ms();
    if (nseeds)
        NQ_HIP(hipMemcpyAsync(d->tasks, seeds.data(), seeds.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              d->stream));
    NQ_HIP(hipMemsetAsync(d->result, 0, 3 * sizeof(uint64_t), d->stream));
    hipError_t e = launch_zero(d->xcount + 2, 1, d->stream);
    if (e != hipSuccess) return fail_hip(e, "launch_zero");
    NQ_HIP(hipStreamSynchronize(d->stream));
    const double t2 = now_ms();

    // Generation: seeds -> depth `rows`.
    const uint32_t *t = d->tasks;
    const unsigned long long *n_dev = nullptr;
    uint64_t n = nseeds;
    int gen = 0;
    if (nseeds) {
        rc = enqueue_split(d, N, seed_rows, d->tasks, nullptr, nseeds, rows - seed_rows, 0, true, d->stream, &t,
                           &n_dev, &n, &gen);
        if (rc) return rc;
    }
    NQ_HIP(hipStreamSynchronize(d->stream));
    const uint64_t ntasks_rows = n;  // the reference's task count for (N, rows)
    // Deeper claimed tasks when the configuration asks for them (the tuner's preplace depth).
    if (nseeds && preplace > seed_rows + gen) {
        int more = 0;
        rc = enqueue_split(d, N, seed_rows + gen, t, n_dev, n, preplace - (seed_rows + gen), gen & 1, true, d->stream,
                           &t, &n_dev, &n, &more);
        if (rc) return rc;
        gen += more;
        NQ_HIP(hipStreamSynchronize(d->stream));
    }
    const double t3 = now_ms();
    if (on_generated) on_generated(ctx, ntasks_rows, t3 - t0 - (t2 - t1));

    // Count, the kernel walking `split` rows below every task in registers.  cfg->stack_depth
    // bounds the depth-`rows` tasks, as for nq_count; the DFS roots sit `split` rows lower.
    const int d_gen = std::min(N, seed_rows + gen);
    int split = 0;
    rc = effective_split(N, d_gen, &in, &split);
    if (rc) return rc;
    if (in.stack_depth) in.stack_depth = std::max(1, in.stack_depth - split);
    nq_config c;
    rc = resolve_config(d->cus, N, std::min(N, d_gen + split), n, &in, false, false, split, &c);
    if (rc) return rc;
    NQ_HIP(hipEventRecord(d->ev0, d->stream));
    const size_t nwaves = static_cast<size_t>(c.grid) * (c.block / 64);
    if (stats) {
        rc = ensure_trace(d, 2 * nwaves);
        if (rc) return rc;
    }
    LaunchArgs a{t, n, N, c.block, c.grid, c.stack_depth, c.chunk, d->state, nullptr, d->result, false,
                 stats ? d->trace : nullptr};
    a.ntasks_dev = n_dev;
    a.xflags = gen ? d->xcount + 2 : nullptr;
    a.apush = c.push == NQ_PUSH_ALWAYS;
    a.split = split;
    e = launch_count(a, d->stream);
    if (e != hipSuccess) return fail_hip(e, "launch_count");
    NQ_HIP(hipEventRecord(d->ev1, d->stream));
    uint64_t res[3] = {0, 0, 0};
    DevState st;
    NQ_HIP(hipMemcpyAsync(res, d->result, sizeof res, hipMemcpyDeviceToHost, d->stream));
    NQ_HIP(hipMemcpyAsync(&st, d->state, sizeof st, hipMemcpyDeviceToHost, d->stream));
    std::vector<uint64_t> tr;
    if (stats) {
        tr.resize(2 * nwaves);
        NQ_HIP(hipMemcpyAsync(tr.data(), d->trace, tr.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, d->stream));
    }
    NQ_HIP(hipStreamSynchronize(d->stream));
    if (res[2] & 2) return fail(NQ_ERR_STATE, "device expansion overflowed its buffer");
    if ((st.error | res[2]) & DEV_ERR_LDS) return fail(NQ_ERR_STATE, "DFS kernel LDS does not start at address 0");
    if (st.error || (res[2] & 1)) return fail(NQ_ERR_DEPTH, "a task needed more stack frames than configured");
    *total = 2 * res[0];  // symmetry fold, src/n_queens_cuda.cu:181-183
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, d->ev0, d->ev1);
        stats->ntasks = ntasks_rows;
        stats->nodes = 0;
        stats->gen_ms = t3 - t0 - (t2 - t1);
        stats->h2d_ms = t2 - t1;
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t0;
        trace_stats(tr, d->wall_khz, &stats->busy_frac, &stats->tail_ms);
        stats->waves = static_cast<uint64_t>(nwaves);
    }
    return NQ_OK;
}

}  // namespace nq

using namespace nq;

extern "C" {

const char *nq_last_error(void) { return g_err.c_str(); }

int nq_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int nq_set_device(int device) {
    int n = nq_device_count();
    if (n <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");
    if (device < 0 || device >= n) return fail(NQ_ERR_ARG, "device index out of range");
    NQ_HIP(hipSetDevice(device));
    return NQ_OK;
}

int nq_device_info(int device, int *num_cus, int *clock_khz, char *name, size_t name_len) {
    int n = nq_device_count();
    if (n <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");
    if (device < 0 || device >= n) return fail(NQ_ERR_ARG, "device index out of range");
    hipDeviceProp_t prop;
    NQ_HIP(hipGetDeviceProperties(&prop, device));
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    return NQ_OK;
}

int nq_config_default(int N, int d0, nq_config *cfg) {
    if (!cfg) return fail(NQ_ERR_ARG, "cfg is NULL");
    cfg->block = 0;  // automatic (resolve_config)
    cfg->stack_depth = std::max(1, frames_needed(N, d0));
    cfg->waves_per_cu = 0;
    cfg->chunk = 16;
    cfg->grid = 0;
    cfg->split = 0;
    cfg->push = 0;
    return NQ_OK;
}

int nq_generate(int N, int rows, uint32_t *tasks, uint64_t cap, uint64_t *ntasks) {
    if (!ntasks) return fail(NQ_ERR_ARG, "ntasks is NULL");
    if (N < 1 || N > 31) return fail(NQ_ERR_ARG, "N must be 1..31");
    if (rows < 1 || rows > N) return fail(NQ_ERR_ARG, "rows must be 1..N");
    std::vector<uint32_t> out;
    generate_tasks(N, rows, out);
    const uint64_t n = out.size() / 3;
    *ntasks = n;
    if (tasks && cap >= n) memcpy(tasks, out.data(), out.size() * sizeof(uint32_t));
    return NQ_OK;
}

int nq_generate_device(int N, int rows, uint32_t *d_tasks, uint64_t cap, uint64_t *ntasks, void *stream) {
    if (!ntasks) return fail(NQ_ERR_ARG, "ntasks is NULL");
    if (N < 1 || N > 31) return fail(NQ_ERR_ARG, "N must be 1..31");
    if (rows < 1 || rows > N) return fail(NQ_ERR_ARG, "rows must be 1..N");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");
    // Rows 0 and 1 carry the reference's restrictions: the host generates those seeds (a few
    // hundred at most); every deeper row is unrestricted and is expanded on the device.
    const int seed_rows = std::min(rows, 2);
    std::vector<uint32_t> seeds;
    generate_tasks(N, seed_rows, seeds);
    uint64_t n = 0;
    bool too_large = false;
    const hipError_t e = generate_expand(N, seeds.data(), seeds.size() / 3, rows - seed_rows, d_tasks, cap, &n,
                                         &too_large, static_cast<hipStream_t>(stream));
    if (e == hipErrorOutOfMemory) return fail(NQ_ERR_NOMEM, "device memory for the expansion rows");
    if (e != hipSuccess) return fail_hip(e, "generate_expand");
    if (too_large) return fail(NQ_ERR_NOMEM, "a row above the last holds more than 2^31-1 tasks");
    *ntasks = n;
    return NQ_OK;
}

int nq_count_device(int N, const uint32_t *d_tasks, uint64_t ntasks, const nq_config *cfg, uint64_t *d_result,
                    uint64_t *d_per_task, int count_nodes, void *stream) {
    if (N < 1 || N > MAX_N) return fail(NQ_ERR_ARG, "N must be 1..28 for the GPU kernel");
    if (ntasks && !d_tasks) return fail(NQ_ERR_ARG, "d_tasks is NULL");
    if (!d_result) return fail(NQ_ERR_ARG, "d_result is NULL");
    Device *d = nullptr;
    int rc = current_device(&d);
    if (rc) return rc;
    // The task depths are not visible from the host here: size the stack for the shallowest
    // prefix the caller allows (cfg->stack_depth), else for depth 1; the kernel flags and skips
    // any task that would need more.
    const int d0 = (cfg && cfg->stack_depth) ? std::max(0, N - 1 - cfg->stack_depth) : 1;
    nq_config in = cfg ? *cfg : nq_config{};
    if (!cfg) in.split = -1;
    int split = 0;
    if (!(count_nodes && d_per_task)) {
        rc = effective_split(N, d0, &in, &split);
        if (rc) return rc;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> g(d->mu);
    if (in.stack_depth) in.stack_depth = std::max(1, in.stack_depth - split);
    nq_config c;
    rc = resolve_config(d->cus, N, std::min(N, d0 + split), ntasks, &in, count_nodes != 0, d_per_task != nullptr,
                        split, &c);
    if (rc) return rc;
    LaunchArgs a{d_tasks, ntasks, N, c.block, c.grid, c.stack_depth, c.chunk, d->state, d_per_task, d_result,
                 count_nodes != 0};
    a.apush = c.push == NQ_PUSH_ALWAYS;
    a.split = split;
    hipError_t e = launch_count(a, s);
    if (e != hipSuccess) return fail_hip(e, "launch_count");
    return NQ_OK;
}

int nq_count(int N, const uint32_t *tasks, uint64_t ntasks, const nq_config *cfg, uint64_t *sum, uint64_t *per_task,
             uint64_t *nodes, nq_stats *stats) {
    const double t0 = now_ms();
    if (N < 1 || N > MAX_N) return fail(NQ_ERR_ARG, "N must be 1..28 for the GPU kernel");
    if (!sum) return fail(NQ_ERR_ARG, "sum is NULL");
    if (ntasks && !tasks) return fail(NQ_ERR_ARG, "tasks is NULL");
    Device *d = nullptr;
    int rc = current_device(&d);
    if (rc) return rc;
    const int d0 = task_min_depth(N, tasks, ntasks);
    nq_config in = cfg ? *cfg : nq_config{};
    if (!cfg) in.split = -1;
    int split = 0;
    if (!(nodes && per_task)) {
        rc = effective_split(N, d0, &in, &split);
        if (rc) return rc;
    }

    std::lock_guard<std::mutex> g(d->mu);
    rc = ensure_tasks(d, ntasks);
    if (rc) return rc;
    if (per_task) {
        rc = ensure_per_task(d, ntasks);
        if (rc) return rc;
    }
    const double t1 = now_ms();
    if (ntasks)
        NQ_HIP(hipMemcpyAsync(d->tasks, tasks, ntasks * 3 * sizeof(uint32_t), hipMemcpyHostToDevice, d->stream));
    NQ_HIP(hipMemsetAsync(d->result, 0, 3 * sizeof(uint64_t), d->stream));
    NQ_HIP(hipStreamSynchronize(d->stream));
    const double t2 = now_ms();
    if (in.stack_depth) in.stack_depth = std::max(1, in.stack_depth - split);
    nq_config c;
    rc = resolve_config(d->cus, N, std::min(N, d0 + split), ntasks, &in, nodes != nullptr, per_task != nullptr,
                        split, &c);
    if (rc) return rc;
    NQ_HIP(hipEventRecord(d->ev0, d->stream));
    const size_t nwaves = static_cast<size_t>(c.grid) * (c.block / 64);
    if (stats) {
        rc = ensure_trace(d, 2 * nwaves);
        if (rc) return rc;
    }
    LaunchArgs a{d->tasks, ntasks, N, c.block, c.grid, c.stack_depth, c.chunk, d->state,
                 per_task ? d->per_task : nullptr, d->result, nodes != nullptr, stats ? d->trace : nullptr};
    a.apush = c.push == NQ_PUSH_ALWAYS;
    a.split = split;
    hipError_t e = launch_count(a, d->stream);
    if (e != hipSuccess) return fail_hip(e, "launch_count");
    NQ_HIP(hipEventRecord(d->ev1, d->stream));
    uint64_t res[3] = {0, 0, 0};
    DevState st;
    NQ_HIP(hipMemcpyAsync(res, d->result, sizeof res, hipMemcpyDeviceToHost, d->stream));
    NQ_HIP(hipMemcpyAsync(&st, d->state, sizeof st, hipMemcpyDeviceToHost, d->stream));
    if (per_task && ntasks)
        NQ_HIP(hipMemcpyAsync(per_task, d->per_task, ntasks * sizeof(uint64_t), hipMemcpyDeviceToHost, d->stream));
    std::vector<uint64_t> tr;
    if (stats) {
        tr.resize(2 * nwaves);
        NQ_HIP(hipMemcpyAsync(tr.data(), d->trace, tr.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, d->stream));
    }
    NQ_HIP(hipStreamSynchronize(d->stream));
    if ((st.error | res[2]) & DEV_ERR_LDS) return fail(NQ_ERR_STATE, "DFS kernel LDS does not start at address 0");
    if (st.error || (res[2] & 1)) return fail(NQ_ERR_DEPTH, "a task needed more stack frames than configured");
    *sum = res[0];
    if (nodes) *nodes = res[1];
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, d->ev0, d->ev1);
        stats->ntasks = ntasks;
        stats->nodes = nodes ? res[1] : 0;
        stats->gen_ms = 0;
        stats->h2d_ms = t2 - t1;
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t0;
        trace_stats(tr, d->wall_khz, &stats->busy_frac, &stats->tail_ms);
        stats->waves = static_cast<uint64_t>(nwaves);
    }
    return NQ_OK;
}

int nq_solve_config(int N, int rows, nq_config *cfg) {
    if (!cfg) return fail(NQ_ERR_ARG, "cfg is NULL");
    if (N < 1 || N > MAX_N || rows < 1 || rows > N) return fail(NQ_ERR_ARG, "N must be 1..28 and rows 1..N");
    // stack_depth bounds the claimed tasks, as nq_solve and nq_count read it (they take the split
    // rows off it for the DFS after the split).
    nq_config_default(N, rows, cfg);
    cfg->split = auto_split(N, rows);
    if (cfg->split > 0) cfg->chunk = 1;  // whole tasks per wave claim
    cfg->preplace = 0;
    char arch[256];
    TunedEntry e;
    if (current_arch(arch, sizeof arch) && find_tuned(arch, N, &e)) {
        // The tuned DFS depth (preplace + split), reached from at least `rows`.
        const int g = std::min(N, std::max(rows, e.preplace));
        cfg->preplace = g > rows ? g : 0;
        cfg->split = std::max(0, std::min(MAX_SPLIT, e.preplace + e.split - g));
        cfg->block = e.block;
        cfg->chunk = e.chunk ? e.chunk : (cfg->split > 0 ? 1 : 16);
        cfg->push = e.push;
        cfg->waves_per_cu = e.waves_per_cu;
        cfg->stack_depth = std::max(1, frames_needed(N, g) + e.stack_extra);
    }
    return NQ_OK;
}

int nq_config_resolve(int N, int d0, uint64_t ntasks, const nq_config *cfg, nq_config *out) {
    if (!out) return fail(NQ_ERR_ARG, "out is NULL");
    if (N < 1 || N > MAX_N || d0 < 0 || d0 > N) return fail(NQ_ERR_ARG, "N must be 1..28 and d0 0..N");
    Device *d = nullptr;
    int rc = current_device(&d);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(d->mu);
    const int split = cfg && cfg->split > 0 ? std::min(cfg->split, MAX_SPLIT) : 0;
    return resolve_config(d->cus, N, d0, ntasks, cfg, false, false, split, out);
}

int nq_solve(int N, int rows, const nq_config *cfg, uint64_t *total, nq_stats *stats) {
    return nq::solve_on_device(N, rows, cfg, total, stats, nullptr, nullptr);
}

}  // extern "C"
