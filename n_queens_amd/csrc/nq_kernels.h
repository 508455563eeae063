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
constexpr int MAX_SPLIT = 8;

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
    // Rows enumerated in registers below every task before its DFS (0..MAX_SPLIT; totals,
    // per-task or node counts, ignored with both).  stack_depth then bounds the tasks' roots.
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
// to while it runs (nq_serve.cpp).  Everything the host writes while the kernel runs lives in
// pinned, coherent host memory that the kernel reads directly (a copy would be a kernel waiting
// for a CU behind the persistent one): the ring, ServeCtl, and the completion records the
// kernel writes back.
//
// Ring entries are 16 bytes: {cur, left, right, tag}, tag = slot << 7 | TERM | twice << 5 | N.
// A published slice is its tasks (at most SERVE_SLICE_MAX) followed by one terminator entry
// (tag bit SERVE_TAG_TERM, x = the slice's task count).
//
// Claims are tickets: a wave takes tickets with one fetch-and-add on the device `head` (a claim
// never fails or retries, however many waves claim at once) and loads their entries with one
// coalesced read into its LDS buffer; ticket t is ring position base + t, and is published once
// base + t < tail[parity].  A wave whose tickets are not yet published waits for the host to
// raise the tail.  Each wave records the end of its latest claim in its `claimed` entry (epoch in
// bits 48..63, end - base below), which the host reads as the claim cursor.
//
// Each slot has one 64-bit accumulator acc[slot] = count << SERVE_FIN_BITS | entries: a lane
// adds a run of finished entries of one slot with one fetch-and-add of (their count << FIN_BITS |
// their weight), where a task weighs 1 and the terminator SERVE_FIN_DONE - tasks, so the entry
// field reaches SERVE_FIN_DONE exactly when the slice's last entry is added (and never passes it:
// no carry into the count).  The lane whose add brings it there takes the slice's count from
// the same word, clears it (waiting for that before it writes the record), numbers the
// completion with gen[slot] and writes {count, generation} into done[slot].  The host sizes slices
// so that their count fits the field (SERVE_COUNT_BITS).
//
// The host stops an instance only with every published task counted: it stores the instance's
// epoch into stop_epoch and from then on publishes into the other tail, so the stopped instance
// sees no new task; its waves exit once idle.  The next instance starts on the same stream
// (after the stopped one has exited) with its head zeroed, the other parity and base = the tail
// at the stop.  A wave without work for `orphan_ticks` wall-clock ticks also exits, so that an
// instance whose host has gone away ends by itself.
constexpr uint32_t SERVE_TAG_TERM = 1u << 6;
constexpr int SERVE_TAG_SLOT_SHIFT = 7;
constexpr int SERVE_CLAIM_EPOCH_SHIFT = 48;  // claimed[] entries: epoch << 48 | (end - base)
constexpr int SERVE_BUF = 64;                // tickets a wave claims and loads at most at once
constexpr int SERVE_FIN_BITS = 18;
constexpr unsigned long long SERVE_SLICE_MAX = 1ull << 16;           // tasks per slice
constexpr unsigned long long SERVE_FIN_DONE = 2 * SERVE_SLICE_MAX;   // entry field of a counted slice
constexpr int SERVE_COUNT_BITS = 64 - SERVE_FIN_BITS;

struct ServeCtl {
    unsigned long long tail[2];  // tasks published so far, per instance parity (host)
    unsigned int stop_epoch;     // the instance asked to stop (host)
    unsigned int pad;
};

struct ServeDone {
    unsigned long long count;  // the slice's count once gen is stored
    unsigned long long gen;    // completions of this slot so far, stored last
};

struct ServeArgs {
    const uint4 *ring;              // host
    unsigned long long ring_mask;   // ring capacity - 1 (a power of two)
    unsigned long long base;        // ring position of ticket 0
    unsigned long long *head;       // device: next ticket (zero at launch)
    const ServeCtl *ctl;            // host
    ServeDone *done;                // host, one per slot
    unsigned long long *claimed;    // host, one per wave: epoch and end of its latest claim
    unsigned long long *acc;        // device, one per slot: count << FIN_BITS | entries (zero between slices)
    unsigned long long *gen;        // device, one per slot: completions so far
    unsigned long long *error;      // device: DEV_ERR_* bits
    unsigned long long orphan_ticks;  // wall-clock ticks a wave waits for work before it exits
    unsigned int epoch;             // this launch's number
    int parity;                     // which tail this instance reads
    int chunk;                      // tickets a wave claims while published work is plentiful
    int block, grid, stack_depth;
    bool apush;
};

size_t serve_lds_bytes(int block, int levels);

// Enqueue one instance of the serving kernel.  serve_fit: push form and workgroups per CU.
hipError_t launch_serve(const ServeArgs &a, hipStream_t stream);
hipError_t serve_fit(int block, int stack_depth, int push, bool *apush, int *blocks_per_cu);

// Expand a batched task list (uint4 {cur, left, right, tag}, N in tag bits 0..4) on the device,
// up to `levels` rows, until no task has more than `target` free rows, keeping the order a
// depth-first expansion of each task gives.  *tasks (hipMalloc'd) is replaced by a new buffer
// and *n by its count; synchronizes the stream (nq_expand.hip).
hipError_t batch_expand(uint4 **tasks, uint64_t *n, uint32_t target, int levels, hipStream_t stream);

// The reference's task list `levels` rows below host seeds (nseeds 12-byte tasks, all at one
// depth, rows >= 2 unrestricted), order-preserving (child count, scan, scatter per row), into
// d_out (device, 3 * cap words) when it holds the result; *n_out = the task count either way.
// Intermediate rows are library-owned device buffers; synchronizes the stream.  *too_large:
// a row above the last holds more than 2^31 - 1 tasks (one scan's limit); nothing was written
// (nq_expand.hip).
hipError_t generate_expand(int N, const uint32_t *seeds, uint64_t nseeds, int levels, uint32_t *d_out, uint64_t cap,
                           uint64_t *n_out, bool *too_large, hipStream_t stream);

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
