// Internal host-side declarations shared by the runtime sources.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace nq {

// Task list of the reference generators for (N, rows): 3 words per task (nq_prefix.cpp).
void generate_tasks(int N, int rows, std::vector<uint32_t> &out);

// One generator part from an arbitrary start state (the reference's recursive signature):
// odd_part selects partial_n_queens_for_odd's row restrictions.
void generate_from(int N, bool odd_part, uint32_t cur, uint32_t left, uint32_t right, int rows,
                   std::vector<uint32_t> &out);

// Thread-local error reporting for the C ABI.
int fail(int status, const std::string &msg);
int fail_hip(int hip_error, const char *what);

double now_ms();

}  // namespace nq
