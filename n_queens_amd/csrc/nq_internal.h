// Internal host-side declarations shared by the runtime sources.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "nq.h"

namespace nq {

// Task list of the reference generators for (N, rows): 3 words per task (nq_prefix.cpp).
void generate_tasks(int N, int rows, std::vector<uint32_t> &out);

// One generator part from an arbitrary start state (the reference's recursive signature):
// odd_part selects partial_n_queens_for_odd's row restrictions.
void generate_from(int N, bool odd_part, uint32_t cur, uint32_t left, uint32_t right, int rows,
                   std::vector<uint32_t> &out);

// Modes of expand_tagged (nq_prefix.cpp; declared where HIP's uint4 is known): EXPAND_PLAIN
// places queens anywhere legal, EXPAND_EVEN and EXPAND_ODD apply the row restrictions of the
// reference's two generators (generate_tasks' parts).
enum { EXPAND_PLAIN = 0, EXPAND_EVEN = 1, EXPAND_ODD = 2 };

// nq_solve with the task list generated on the device (nq_api.cpp); on_generated (nullable)
// receives the reference's task count for (N, rows) and the generation time in ms.
int solve_on_device(int N, int rows, const nq_config *cfg, uint64_t *total, nq_stats *stats,
                    void (*on_generated)(void *ctx, uint64_t ntasks, double ms), void *ctx);

// nq_count's launch plan for a task list of shallowest depth d0 on a device with `cus` compute
// units (split resolved from cfg, cfg NULL = automatic), without touching device memory
// (nq_api.cpp).  out->split holds the split rows the DFS kernel walks.
int plan_count(int cus, int N, int d0, uint64_t ntasks, const nq_config *cfg, bool nodes, bool per_task,
               nq_config *out);

// Shallowest prefix depth (queens placed) over a task list.
int task_min_depth(int N, const uint32_t *tasks, uint64_t ntasks);

// One entry of the autotuner's cache (nq_tune.cpp): the fastest DFS launch for whole boards of
// size N on one device type.
struct TunedEntry {
    int N = 0, preplace = 0, split = 0, block = 0, chunk = 0, push = 0, waves_per_cu = 0, stack_extra = 0;
};

// The cache file's path ($NQ_TUNE_CACHE, else $HOME/.cache/n_queens_amd/tune.txt; "" = none).
std::string tune_cache_path();

// The last cache entry for (device arch, N), if any.
bool find_tuned(const char *arch, int N, TunedEntry *out);

// Thread-local error reporting for the C ABI.
int fail(int status, const std::string &msg);
int fail_hip(int hip_error, const char *what);

double now_ms();

}  // namespace nq
