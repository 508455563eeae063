// n_queens.h — C++ drop-in for the reference's counting entry points.
//
// Same names, signatures and C++ linkage as the reference header src/n_queens.h:7-19, so that a
// main.cpp-style caller (src/main.cpp:51, and the CPU cross-checks commented out at :35-47)
// links unchanged against libnq_hip.so (with include/utils.h, the reference's logging pair).
// cuda_n_queens keeps its name as "the GPU count" even though it runs HIP on an MI355X; it is
// implemented on top of the C ABI in nq.h.  The CPU counters run on the host only and are never
// used by the GPU path.
#pragma once

#include <vector>

using namespace std;  // as the reference header does (src/n_queens.h:5)

// n_queens (src/n_queens.cpp:8-21): add the number of completions of the prefix
// (cur, left, right) to sum (which the caller initialises).
void n_queens(int N, int cur, int left, int right, long long &sum);

// n_queens_iterative (src/n_queens.cpp:24-72): the reference's explicit-stack variant of
// n_queens.  It adds the same number to sum, except that a prefix already holding N-1 or N
// queens adds nothing, as the reference's loop does (it counts a board when a placement leaves
// one free column that is still a candidate, :62-66).
void n_queens_iterative(int N, int cur, int left, int right, long long &sum);

// partial_n_queens (src/n_queens.cpp:74-91): append the (cur, left, right) triple of every
// placement of `rows` more queens below (cur, left, right) to tot; row 0 (cur == 0) is limited
// to columns < N/2.
void partial_n_queens(int N, int cur, int left, int right, vector<int> &tot, int rows);

// partial_n_queens_for_odd (src/n_queens.cpp:93-113): the same with row 0 limited to the middle
// column and the row after a single queen limited to columns < (N-2)/2.
void partial_n_queens_for_odd(int N, int cur, int left, int right, vector<int> &tot, int rows);

// cuda_n_queens (src/n_queens_cuda.cu:84-186): the number of N-Queens solutions, counted on the
// GPU with the first `rows` rows pre-placed.  Prints the reference's progress lines.  Returns -1
// when no GPU is present (src/n_queens_cuda.cu:105-108) and, unlike the reference (which calls
// exit(-1), src/macro.h:40-49), also -1 after any other failure, with the reason on stderr.
long long cuda_n_queens(int N, int rows);

// serial_n_queens / parallel_n_queens (src/n_queens.cpp:115-163): the same total on the host,
// task by task over the generator's list, on one thread or on 32 threads.
long long serial_n_queens(int N, int rows);
long long parallel_n_queens(int N, int rows);
