/*
 * nq.h — C ABI of libnq_hip.so, the MI355X (gfx950) N-Queens solution counter.
 *
 * Plain C: integers, pointers and sizes only.  Every function returns NQ_OK (0) or a negative
 * nq_status and never exits the process; nq_last_error() gives the message of the last failure
 * on the calling thread.  Host arrays are owned by the caller; device memory is owned by the
 * library unless a function says it takes device pointers.
 *
 * The reference (ygch/n_queens) has no FFI: its only boundary is the C++ header
 * src/n_queens.h:7-19, called from src/main.cpp:51.  include/n_queens.h keeps those C++ names
 * on top of this ABI; each entry point below cites the reference code it replaces.
 */
#ifndef NQ_H
#define NQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nq_status {
    NQ_OK = 0,
    NQ_ERR_ARG = -1,      /* invalid argument (N, rows, sizes, null pointers)            */
    NQ_ERR_NO_DEVICE = -2,/* no HIP device visible (reference: n_queens_cuda.cu:105-108)   */
    NQ_ERR_HIP = -3,      /* a HIP runtime call failed (reference: CU_SAFE_CALL, macro.h)  */
    NQ_ERR_DEPTH = -4,    /* stack depth too small for (N, prefix depth)                   */
    NQ_ERR_NOMEM = -5,    /* allocation failed, or a request larger than a server slot     */
    NQ_ERR_BUSY = -6,     /* admission refused (serving: task or LDS budget exhausted)     */
    NQ_ERR_STATE = -7     /* call not valid in the handle's current state                  */
};

/* Launch configuration of the DFS kernel.  Zero fields take the defaults chosen by
 * nq_config_default().  The reference fixes these at compile time (src/macro.h:25-35:
 * CONFIG1 128 threads / depth 24, CONFIG2 160 / 19, CONFIG3 192 / 16) plus grid 1024
 * (src/n_queens_cuda.cu:164-167). */
typedef struct nq_config {
    int block;        /* threads per workgroup, a multiple of 64, <= 256; 0 = automatic:  */
                      /* 256 when that keeps the resident waves per CU that 128 gets with   */
                      /* the same push rule (one wave of each workgroup per SIMD), else 128 */
    int stack_depth;  /* LDS frames per lane; must be >= N - 1 - (prefix depth)           */
    int waves_per_cu; /* resident-wave target that sizes the persistent grid (0 = max)     */
    int chunk;        /* tasks a wave claims per cursor grab (1..4096); with a split a     */
                      /* wave claims whole tasks, one at a time by default                  */
    int grid;         /* explicit workgroup count (0 = from waves_per_cu)                  */
    int split;        /* rows the DFS kernel walks below each task before its LDS DFS     */
                      /* (0..8; -1 = automatic): a wave claims a whole task and hands the   */
                      /* nodes `split` rows below it to its lanes as DFS roots, in registers */
                      /* (no expanded list in memory).  Totals, per-task or node counts;    */
                      /* ignored when per-task and node counts are both requested.          */
    int push;         /* DFS push form: NQ_PUSH_AUTO (0) takes NQ_PUSH_ALWAYS when its spare */
                      /* stack level costs no resident waves, else NQ_PUSH_MASKED.  Both    */
                      /* give identical counts.                                              */
    int preplace;     /* nq_solve only: depth of the task list the waves claim, generated on */
                      /* the device (0 or <= rows: rows itself); the reference's task count  */
                      /* for `rows` is still the one reported.                               */
} nq_config;

#define NQ_PUSH_AUTO 0
#define NQ_PUSH_MASKED 1   /* exec-masked write of frames that keep candidates               */
#define NQ_PUSH_ALWAYS 2   /* unconditional write, one spare LDS level per lane              */

/* Per-call statistics (all optional). */
typedef struct nq_stats {
    uint64_t ntasks;      /* tasks counted                                              */
    uint64_t nodes;       /* DFS nodes (only when requested; else 0)                    */
    double gen_ms;        /* host prefix generation                                     */
    double h2d_ms;        /* host->device copy of the task list                         */
    double kernel_ms;     /* DFS kernel, HIP events on the launch stream                */
    double total_ms;      /* the whole call                                             */
    double busy_frac;     /* sum of per-wave residency / (waves x kernel span)          */
    double tail_ms;       /* first wave finished -> last wave finished                  */
    uint64_t waves;       /* resident waves of the persistent grid                      */
} nq_stats;

/* ---- device ------------------------------------------------------------------------ */
int nq_device_count(void);            /* cudaGetDeviceCount, n_queens_cuda.cu:103-104     */
const char *nq_last_error(void);
int nq_set_device(int device);        /* cudaSetDevice, n_queens_cuda.cu:149              */
/* Compute units, reported clock (kHz) and "name (gcnArch)" of a device. */
int nq_device_info(int device, int *num_cus, int *clock_khz, char *name, size_t name_len);

/* Fill cfg with defaults for (N, shallowest prefix depth d0): automatic workgroup size (block
 * 0), the smallest legal stack depth N-1-d0 (LDS per lane decides occupancy, so no frame is
 * wasted), chunk 16, no split, push NQ_PUSH_AUTO, and a persistent grid of every workgroup that
 * fits on the device.  The reference's
 * compile-time presets only decide which (N, rows) pairs its CLI accepts; the CLI here keeps
 * that check (src/main.cpp:16-19). */
int nq_config_default(int N, int d0, nq_config *cfg);

/* The configuration nq_solve(N, rows, NULL, ...) uses, in the form nq_solve and nq_count take
 * it back.  When the autotuner's cache (python -m n_queens_amd.tune; the file $NQ_TUNE_CACHE,
 * else $HOME/.cache/n_queens_amd/tune.txt) holds an entry for N on the current device: its
 * preplace depth, split, block, chunk, push, waves per CU and stack depth, with a preplace depth
 * below `rows` raised to `rows` and the split shortened to keep the tuned DFS depth.  Otherwise
 * the built-in rule: preplace 0 (= rows), split = clamp(N - rows - R, 0, 8) so that the DFS
 * starts R rows above the last row (R = (N+4)/2 up to N = 20, R = 13 from N = 21 on: 12 LDS
 * frames), chunk 1 with a split.  stack_depth bounds the claimed tasks (depth
 * max(rows, preplace)); the count takes the split rows off it.  No device is needed for the
 * built-in rule.  nq_solve(N, rows, &cfg, ...) with this cfg is nq_solve(N, rows, NULL). */
int nq_solve_config(int N, int rows, nq_config *cfg);

/* The launch configuration nq_count takes on the current device for ntasks tasks whose
 * shallowest prefix depth is d0 (after any device split; totals): cfg's zero fields resolved, i.e. the workgroup size, the push form (never
 * NQ_PUSH_AUTO in *out) and the persistent grid.  No reference counterpart: it reports what the
 * reference fixes at compile time (src/macro.h:25-35, src/n_queens_cuda.cu:164-167). */
int nq_config_resolve(int N, int d0, uint64_t ntasks, const nq_config *cfg, nq_config *out);

/* ---- prefix/task generator (host) ------------------------------------------------------
 * Replaces partial_n_queens + partial_n_queens_for_odd (src/n_queens.cpp:74-113) as called by
 * cuda_n_queens (src/n_queens_cuda.cu:91-95): every non-attacking placement of the first
 * `rows` queens with row 0 in the low half of the columns, plus (odd N) the middle column
 * with row 1 in the low half, in the reference's emission order, as (cur, left, right)
 * triples (3 uint32 words per task).  With tasks == NULL or cap too small only the count is
 * returned in *ntasks.  Rows must be in 1..N. */
int nq_generate(int N, int rows, uint32_t *tasks, uint64_t cap, uint64_t *ntasks);

/* The same task list generated on the current device, straight into device memory: replaces the
 * host generation and the host-to-device copy of cuda_n_queens (src/n_queens_cuda.cu:91-95,
 * :156-160; partial_n_queens(_for_odd), src/n_queens.cpp:74-113).  The host places rows 0 and
 * 1 (the reference's restricted rows: a few hundred tasks at most); every deeper row is expanded
 * on the device, order-preserving (child count, scan, scatter), so d_tasks receives exactly
 * nq_generate's words in the same order.  d_tasks is a device pointer (3 * cap uint32); with
 * d_tasks == NULL or cap too small only *ntasks is set.  Intermediate rows are device memory the
 * call allocates and frees; stream is a hipStream_t (NULL = default stream), synchronised before
 * the call returns.  N 1..31, rows 1..N; NQ_ERR_NOMEM when device memory is short or a row
 * above the last holds more than 2^31-1 tasks. */
int nq_generate_device(int N, int rows, uint32_t *d_tasks, uint64_t cap, uint64_t *ntasks, void *stream);

/* ---- counting --------------------------------------------------------------------------
 * Replaces the per-GPU region of cuda_n_queens (src/n_queens_cuda.cu:146-179: H2D copy,
 * kernel n_queens :9-82, D2H of partial_sum) with one device consuming the whole list.
 * Counts all completions of every task (3 words each, all at the same N; prefixes of any
 * depth <= N).  *sum = sum of per-task counts (no symmetry fold).  per_task (nullable) gets the
 * per-task uint64 counts in task order — the reference's partial_sum vector.  nodes (nullable)
 * selects the node-counting kernel and returns the DFS node count.  stats is nullable.  cfg
 * NULL or split -1: the automatic split for the shallowest task. */
int nq_count(int N, const uint32_t *tasks, uint64_t ntasks, const nq_config *cfg, uint64_t *sum,
             uint64_t *per_task, uint64_t *nodes, nq_stats *stats);

/* Asynchronous device-pointer form of nq_count, for callers that keep the task list resident
 * in HBM: d_tasks (3*ntasks uint32), d_result (3 uint64: [0] += sum, [1] += nodes, [2] |= error
 * flags, bit 0 = a task needed more stack frames than configured; the caller zeroes it) and
 * d_per_task (nullable, ntasks uint64) are device pointers on the current device; stream is a
 * hipStream_t (NULL = default stream).  cfg->stack_depth, when set, bounds N-1-depth over the
 * tasks as given (before any split); cfg NULL or split -1 takes the automatic split for tasks of
 * depth N-1-stack_depth (depth 1 without a stack depth).  Enqueues kernels only (no host
 * synchronisation, no memset) and returns, so the call can be captured in a hipGraph; the
 * results are valid after the stream synchronises.  Calls on one device share the library's task
 * cursor and are serialised by the library's device lock; each call must be allowed to finish on
 * its stream before a call on another stream of the same device is issued. */
int nq_count_device(int N, const uint32_t *d_tasks, uint64_t ntasks, const nq_config *cfg,
                    uint64_t *d_result, uint64_t *d_per_task, int count_nodes, void *stream);

/* The whole reference pipeline: long long cuda_n_queens(int N, int rows)
 * (src/n_queens_cuda.cu:84-186): generate, count on the device, symmetry fold x2.
 * *total = the number of N-Queens solutions.  rows is used as given (the CLI applies the
 * reference's validation, src/main.cpp:10-29). */
int nq_solve(int N, int rows, const nq_config *cfg, uint64_t *total, nq_stats *stats);

/* ---- batched instances (no reference counterpart) -------------------------------------
 * An instance is a board of size N (1..28) whose first k rows (0 <= k <= N) hold queens at
 * columns cols[0..k-1]; its count is the number of completions of that prefix — what the
 * reference's n_queens(N, cur, left, right, sum) (src/n_queens.cpp:8-21) returns for the
 * prefix masks — with no symmetry fold.  k = 0 is the empty board.
 *
 * A batch plan validates every instance on the host, orders their seeds largest-first and
 * uploads them; the device expands them (a few more rows, until no task needs more than the
 * stack depth, in the order a host depth-first expansion gives), and "zero counts, reset
 * cursor, persistent batched DFS" is captured into a hipGraph; each run replays the graph on
 * the plan's own stream.  Plans on different streams run concurrently. */
typedef struct nq_instance {
    int N;
    int k;
    int8_t cols[32];
} nq_instance;

typedef struct nq_batch_stats {
    uint64_t ntasks;     /* device tasks after expansion                                */
    uint64_t total;      /* sum of all instance counts                                  */
    double expand_ms;    /* plan creation: host validation and seeds, device expansion  */
    double kernel_ms;    /* device time of the replayed graph (HIP events)              */
    double total_ms;     /* whole call (nq_batch_count only)                            */
} nq_batch_stats;

typedef struct nq_batch_plan nq_batch_plan;

/* Returns NQ_ERR_ARG (message names the instance) for N out of range, k > N, a column out of
 * range or an attacking prefix.  cfg: block, chunk, waves_per_cu, grid, push as for nq_count;
 * stack_depth (0 = 12 levels per task) bounds the per-task depth. */
int nq_batch_plan_create(const nq_instance *inst, int n, const nq_config *cfg, nq_batch_plan **plan);
int nq_batch_plan_launch(nq_batch_plan *plan);   /* enqueue one replay, asynchronous         */
int nq_batch_plan_wait(nq_batch_plan *plan, uint64_t *counts, nq_batch_stats *stats);
int nq_batch_plan_run(nq_batch_plan *plan, uint64_t *counts, nq_batch_stats *stats);
int nq_batch_plan_destroy(nq_batch_plan *plan);

/* One-shot: create, run once, destroy.  counts[i] = instance i's count. */
int nq_batch_count(const nq_instance *inst, int n, const nq_config *cfg, uint64_t *counts,
                   nq_batch_stats *stats);

/* One instance through the plain (single-N) counting path: the per-instance launch the batch
 * is measured against. */
int nq_instance_count(const nq_instance *inst, const nq_config *cfg, uint64_t *count);

/* ---- serving with continuous batching (no reference counterpart) ---------------------
 * A server keeps one ring of tasks, in pinned host memory that the device reads directly, which a
 * persistent kernel consumes while the server keeps appending to it (publishing is plain stores:
 * no copy ever waits behind the running kernel).  nq_server_submit validates a request, expands it into tasks on
 * the caller's thread (until no task keeps more than depth_target free rows, and a small request
 * further, into at least 4096 tasks unless they would keep fewer than 5 free rows) and admits it
 * while the pending work fits max_pending_tasks (else NQ_ERR_BUSY, which a later retry may
 * pass; a request of more tasks than max_pending_tasks can never fit and is refused with
 * NQ_ERR_NOMEM).  A
 * scheduler thread publishes requests into the ring in slices of at most batch_work estimated
 * DFS nodes, least estimated remaining work first (aged after a second of waiting), keeping no more than
 * two slices' work published and unclaimed, so that requests arriving while others run are
 * admitted into the running kernel within that much work.  The kernel counts every task into
 * its request's slot and publishes a request's count as soon as its last task is counted.  A
 * request's count equals nq_instance_count's for the same instance.  Calls on one server are
 * thread-safe. */
typedef struct nq_server_config {
    int waves_per_cu;            /* resident kernel waves per CU (default 0: all that fit) */
    uint64_t slot_tasks;         /* ring capacity in tasks (default 8M = 128 MB pinned host  */
                                 /* memory; rounded up to 2^k)                               */
    int slot_requests;           /* published slices not yet counted (default 4096)          */
    uint64_t max_pending_tasks;  /* admission budget of queued tasks (default 64M)         */
    int depth_target;            /* most free rows a task keeps (default 12): bounds both  */
                                 /* the LDS stack and one task's run time on its lane      */
    int block;                   /* workgroup size (default 128)                           */
    int chunk;                   /* tickets a wave claims at once while published work is  */
                                 /* plentiful, 1..64 (default 32); else one per idle lane  */
    double batch_work;           /* estimated DFS nodes per published slice (default 5e8)  */
} nq_server_config;

typedef struct nq_request_stats {
    double queue_ms;             /* submit -> its batch launched                           */
    double latency_ms;           /* submit -> result ready on the host                     */
    uint64_t ntasks;             /* tasks the request was expanded into                    */
} nq_request_stats;

typedef struct nq_server nq_server;

int nq_server_config_default(nq_server_config *cfg);
int nq_server_open(const nq_server_config *cfg, nq_server **server);
int nq_server_submit(nq_server *server, const nq_instance *inst, uint64_t *ticket);
/* NQ_OK and the count when done (the ticket is then released), NQ_ERR_STATE while pending. */
int nq_server_poll(nq_server *server, uint64_t ticket, uint64_t *count, nq_request_stats *stats);
/* Block until done or timeout_ms (< 0: no timeout); NQ_ERR_STATE on timeout. */
int nq_server_wait(nq_server *server, uint64_t ticket, int timeout_ms, uint64_t *count,
                   nq_request_stats *stats);
/* batches: slices published into the ring. */
int nq_server_stats(nq_server *server, uint64_t *completed, uint64_t *batches, uint64_t *solutions);
/* Time with at least one request in the ring (its first slice published -> count observed on
 * the host), and time since open. */
int nq_server_timing(nq_server *server, double *device_ms, double *open_ms);
/* Finish every admitted request, stop the scheduler and free the device queues. */
int nq_server_close(nq_server *server);

/* ---- counting without device allocation (no reference counterpart) ---------------------
 * For callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc): nq_count's counts (sum, per_task, nodes, same cfg rules) with every device buffer
 * nq_count would allocate — the device state, the result words and the staging of the task list
 * and of the per-task counts — in one static device array of NQ_STATIC_SCRATCH_BYTES (about 20
 * MiB) that each device the library runs on holds until the process ends.  At most
 * NQ_STATIC_MAX_TASKS tasks per call (NQ_ERR_ARG beyond).  Calls on one device share its array
 * and are serialised by the library; each call resets what must start at zero. */
#define NQ_STATIC_MAX_TASKS (1u << 20)
#define NQ_STATIC_SCRATCH_BYTES (256u + 20u * NQ_STATIC_MAX_TASKS)
int nq_count_static(int N, const uint32_t *tasks, uint64_t ntasks, const nq_config *cfg, uint64_t *sum,
                    uint64_t *per_task, uint64_t *nodes);

/* ---- resumable counts (no reference counterpart) --------------------------------------
 * For runs that take days (the reference's N=27 run took 28.4 days with no way to resume,
 * README.md:46).  The (N, rows) task list is the host generator's (reference order), counted
 * in consecutive ranges of range_tasks tasks (0 = 1 % of the list) with nq_count.  After every
 * range the text file `path` is rewritten atomically (written to path.tmp, flushed, renamed)
 * with N, rows, the task count, the first uncounted task and the partial sum.  A call that finds
 * the file resumes after its last completed range; a file for another (N, rows) or task count,
 * or one that does not parse, is NQ_ERR_ARG before any device work.  max_ranges > 0 returns
 * after that many ranges (state->done = 0 if tasks remain).  cfg as for nq_count; split < 0 or
 * cfg NULL takes nq_solve's split. */
typedef struct nq_resume_state {
    uint64_t ntasks;   /* tasks in the (N, rows) list                                       */
    uint64_t next;     /* first task not yet counted                                        */
    uint64_t sum;      /* sum of the per-task counts of tasks [0, next)                     */
    uint64_t total;    /* 2 * sum once done (symmetry fold), else 0                         */
    int done;          /* 1 when every task is counted                                       */
    int ranges;        /* ranges counted by this call                                        */
    double gen_ms;     /* host generation of the task list                                  */
    double kernel_ms;  /* device time of this call's ranges                                 */
} nq_resume_state;

int nq_solve_resumable(int N, int rows, const nq_config *cfg, const char *path, uint64_t range_tasks,
                       int max_ranges, nq_resume_state *state);

#ifdef __cplusplus
}
#endif

#endif /* NQ_H */
