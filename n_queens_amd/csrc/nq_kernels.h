// Internal interface between the HIP kernels (nq_dfs.hip) and the host runtime (nq_api.cpp).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace nq {

// Largest board.  The DFS keeps the columns word with every bit from N up set (the
// candidates are ~(left | right | cur)); the batched task word packs N into 5 bits.
constexpr int MAX_N = 28;
constexpr int MAX_BLOCK = 256;
constexpr int MAX_WAVES_PER_BLOCK = MAX_BLOCK / 64;
// Rows the DFS kernel can enumerate in registers below each task (split_kernel).
constexpr int MAX_SPLIT = 4;

// Device-resident per-launch state, owned by the library (one per device).
struct DevState {
    unsigned long long cursor;  // next unclaimed task index
    unsigned long long total;   // sum of per-task completion counts
    unsigned long long nodes;   // DFS nodes (node-counting kernel only)
    unsigned long long error;   // DEV_ERR_* bits
};

// DevState::error bits (folded into result[2] beside the expansion's overflow bit 2).
constexpr unsigned long long DEV_ERR_DEPTH = 1;  // a task needed more stack frames than configured
constexpr unsigned long long DEV_ERR_LDS = 4;    // dynamic LDS did not start at address 0

struct LaunchArgs {
    const uint32_t *tasks;  // 3 words per task: cur, left, right
    uint64_t ntasks;
    int N;
    int block;        // threads per workgroup (multiple of 64, <= MAX_BLOCK)
    int grid;         // workgroups
    int stack_depth;  // frames per lane
    int chunk;        // tasks per cursor grab
    DevState *state;
    uint64_t *per_task;   // nullable
    uint64_t *result;     // nullable: [0] += total, [1] += nodes, [2] |= error flags at the end
    bool count_nodes;
    uint64_t *trace = nullptr;  // nullable: per wave {start, end} wall-clock ticks (grid * block/64 * 2)
    const unsigned long long *ntasks_dev = nullptr;  // nullable: task count on the device (<= ntasks)
    const unsigned long long *xflags = nullptr;      // nullable: expansion flags folded into result[2]
    bool apush = false;  // unconditional push (the launch allocates stack_depth + 1 levels)
    // Rows enumerated in registers below every task before its DFS (0..MAX_SPLIT; plain totals
    // only: ignored with count_nodes or per_task).  stack_depth then bounds the tasks' roots.
    int split = 0;
};

struct BatchArgs {
    const uint4 *tasks;  // {cur, left, right, inst << 6 | twice << 5 | N}
    uint64_t ntasks;
    uint64_t *counts;    // per instance, zeroed by the launch
    uint64_t ninst;
    int block, grid, stack_depth, chunk;
    DevState *state;
    bool apush = false;  // unconditional push (stack_depth + 1 levels)
};

// Serving: a persistent kernel consuming a ring of batched tasks that the host keeps appending
// to while it runs (nq_serve.cpp).  The host publishes tasks by raising `tail` in ServeCtl
// (pinned host memory) after their copy into the ring has completed; waves claim published
// tasks with a compare-and-swap on the device `head`, never beyond `tail`.  A task's tag
// carries its request slot (tag >> 6); every finished task adds its count to count[slot] and
// takes one off remaining[slot], and whoever takes the last one publishes {count, seq[slot]}
// into done[slot] (pinned host memory).  Each wave exits after `idle_ticks` wall-clock ticks
// without work, or at once when ServeCtl::stop is set and nothing is claimable; a wave that exits
// stores its launch epoch in ServeCtl::exited.  From then on the instance retires: its waves
// claim nothing new and exit once their tasks are counted, and the host launches a fresh
// instance for the work that follows.  Every wait in the kernel is bounded by the idle timeout.
struct ServeCtl {
    unsigned long long tail;     // tasks published so far (written by the host)
    unsigned long long claimed;  // head after a recent claim (device; a scheduling hint only)
    unsigned int stop;           // exit once nothing is claimable (written by the host)
    unsigned int exited;         // epoch of the last instance a wave of which has exited (device)
};

struct ServeDone {
    unsigned long long count;   // the request's count once seq is stored
    unsigned long long seq;     // the request's sequence number, stored last
};

struct ServeArgs {
    const uint4 *ring;              // device: {cur, left, right, slot << 6 | twice << 5 | N}
    unsigned long long ring_mask;   // ring capacity - 1 (a power of two)
    unsigned long long *head;       // device: next unclaimed task index
    ServeCtl *ctl;                  // pinned host memory
    ServeDone *done;                // pinned host memory, one per slot
    unsigned long long *remaining;  // device, one per slot: tasks not yet counted
    unsigned long long *count;      // device, one per slot
    const unsigned long long *seq;  // device, one per slot
    unsigned long long *error;      // device: DEV_ERR_* bits
    unsigned long long idle_ticks;  // wall-clock ticks a wave waits for work before it exits
    unsigned int epoch;             // this launch's number (stored into ctl->exited on exit)
    int block, grid, stack_depth, chunk;
    bool apush;
};

// Enqueue one instance of the serving kernel.  serve_fit: push form and workgroups per CU.
hipError_t launch_serve(const ServeArgs &a, hipStream_t stream);
hipError_t serve_fit(int block, int stack_depth, int push, bool *apush, int *blocks_per_cu);

// Expand a batched task list (uint4 {cur, left, right, tag}, N in tag bits 0..4) on the device,
// up to `levels` rows, until no task has more than `target` free rows, keeping the order a
// depth-first expansion of each task gives.  *tasks (hipMalloc'd) is replaced by a new buffer
// and *n by its count; synchronizes the stream (nq_expand.hip).
hipError_t batch_expand(uint4 **tasks, uint64_t *n, uint32_t target, int levels, hipStream_t stream);

// Enqueue: zero counts, reset state, batched DFS.  Capturable into a hipGraph (kernels only).
hipError_t launch_batch(const BatchArgs &a, hipStream_t stream);
hipError_t batch_occupancy(int block, int stack_depth, bool apush, int *blocks_per_cu);
// Push form for push 0 (automatic) from the resident workgroups per CU of each form (same
// block).  At equal occupancy the unconditional push is about 7% faster (N=20 and N=21, 6 rows),
// but the masked push with 14 waves/CU beats it with 12 by about 1%; so it is taken only when
// its spare stack level costs no workgroup.
inline bool prefer_always_push(int occ_masked, int occ_always) {
    return occ_always > 0 && occ_always >= occ_masked;
}

// Push form and resident workgroups per CU for a batched launch: push 0 automatic
// (prefer_always_push), 1 masked, 2 unconditional; waves_per_cu > 0 caps the workgroups.
hipError_t batch_fit(int block, int stack_depth, int waves_per_cu, int push, bool *apush, int *blocks_per_cu);

// Enqueue: reset state, DFS, (optional) fold into result.  Returns a hipError_t.
hipError_t launch_count(const LaunchArgs &a, hipStream_t stream);

// Expand tasks (3 words each) by one row into out (children in arbitrary order); *out_n is
// zeroed and receives the child count.  The input count is n_max, or min(n_max, *n_dev) when
// n_dev is set.  Nothing is written past cap tasks: overflow sets bit 1 of *flags.
hipError_t launch_expand(const uint32_t *in, uint64_t n_max, const unsigned long long *n_dev, int N, uint32_t *out,
                         uint64_t cap, unsigned long long *out_n, unsigned long long *flags, hipStream_t stream);

// *out = the number of tasks launch_expand would produce from these (zeroed by the launch).
hipError_t launch_child_count(const uint32_t *in, uint64_t n_max, const unsigned long long *n_dev, int N,
                              unsigned long long *out, hipStream_t stream);

// Zero n 64-bit words with a kernel (graph-capturable, unlike a memset node).
hipError_t launch_zero(unsigned long long *p, uint64_t n, hipStream_t stream);

// Resident workgroups per CU for the DFS kernel at (block, stack_depth, push form, split rows).
hipError_t dfs_occupancy(int block, int stack_depth, bool count_nodes, bool per_task, bool apush, int split,
                         int *blocks_per_cu);

// Dynamic LDS of a launch with `levels` stack levels.
size_t dfs_lds_bytes(int block, int levels);

}  // namespace nq
