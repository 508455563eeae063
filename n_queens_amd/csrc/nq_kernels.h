// Internal interface between the HIP kernels (nq_dfs.hip) and the host runtime (nq_api.cpp).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace nq {

// Bit layout of the "cur" word the DFS keeps per lane: board columns in bits 0..N-1 and the
// number of queens placed (the depth) in bits DEPTH_SHIFT..31, so that placing a queen is one
// add (cur + p + DEPTH_ONE) and the last-row test is one unsigned compare.
constexpr int DEPTH_SHIFT = 27;
constexpr int MAX_N = 27;  // board bits must stay below DEPTH_SHIFT
constexpr int MAX_BLOCK = 256;
constexpr int MAX_WAVES_PER_BLOCK = MAX_BLOCK / 64;

// Device-resident per-launch state, owned by the library (one per device).
struct DevState {
    unsigned long long cursor;  // next unclaimed task index
    unsigned long long total;   // sum of per-task completion counts
    unsigned long long nodes;   // DFS nodes (node-counting kernel only)
    unsigned long long error;   // nonzero: a task needed more stack frames than configured
};

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
};

struct BatchArgs {
    const uint4 *tasks;  // {cur, left, right, inst << 5 | N}
    uint64_t ntasks;
    uint64_t *counts;    // per instance, zeroed by the launch
    uint64_t ninst;
    int block, grid, stack_depth, chunk;
    DevState *state;
};

// Enqueue: zero counts, reset state, batched DFS.  Capturable into a hipGraph (kernels only).
hipError_t launch_batch(const BatchArgs &a, hipStream_t stream);
hipError_t batch_occupancy(int block, int stack_depth, int *blocks_per_cu);

// Enqueue: reset state, DFS, (optional) fold into result.  Returns a hipError_t.
hipError_t launch_count(const LaunchArgs &a, hipStream_t stream);

// Expand tasks (3 words each) by one row into out (children in arbitrary order); *out_n is
// zeroed and receives the child count.  The input count is n_max, or min(n_max, *n_dev) when
// n_dev is set.  Nothing is written past cap tasks: overflow sets bit 1 of *flags.
hipError_t launch_expand(const uint32_t *in, uint64_t n_max, const unsigned long long *n_dev, int N, uint32_t *out,
                         uint64_t cap, unsigned long long *out_n, unsigned long long *flags, hipStream_t stream);

// Zero n 64-bit words with a kernel (graph-capturable, unlike a memset node).
hipError_t launch_zero(unsigned long long *p, uint64_t n, hipStream_t stream);

// Resident workgroups per CU for the DFS kernel at (block, stack_depth).
hipError_t dfs_occupancy(int block, int stack_depth, bool count_nodes, bool per_task, int *blocks_per_cu);

size_t dfs_lds_bytes(int block, int stack_depth);

}  // namespace nq
