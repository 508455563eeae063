/*
 * CPU oracle for the N-Queens counting path — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker or the timed CPU baseline.  The product library
 * (n_queens_amd/) never links or calls it.
 *
 * Every function restates one piece of the reference (ygch/n_queens) with uint32_t masks
 * instead of the reference's signed int (left shifts of a signed int with bit 31 set are
 * undefined behaviour in C/C++; only the low N bits are ever read, so results are identical).
 */
#ifndef NQ_ORACLE_H
#define NQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Task-list generator: partial_n_queens (src/n_queens.cpp:74-91) followed, for odd N, by
 * partial_n_queens_for_odd (src/n_queens.cpp:93-113), in the order cuda_n_queens calls them
 * (src/n_queens_cuda.cu:91-95).  Writes (cur, left, right) triples into out[3*i..3*i+2] for
 * i < cap and returns the total number of tasks (also when it exceeds cap; out may be NULL). */
uint64_t oracle_gen_tasks(int N, int rows, uint32_t *out, uint64_t cap);

/* Completions of one prefix: n_queens(N, cur, left, right, sum) (src/n_queens.cpp:8-21).
 * Returns the number of full boards below (cur, left, right); 1 when cur already fills all N
 * columns. */
uint64_t oracle_count_subtree(int N, uint32_t cur, uint32_t left, uint32_t right);

/* The reference's n_queens_iterative (src/n_queens.cpp:24-72): completions of one prefix,
 * except that a prefix already holding N-1 or N queens adds nothing (the reference's own
 * behaviour).  The reference's 16-frame stack overflows for prefixes more than 16 rows above
 * the last; this one holds 32 frames. */
uint64_t oracle_count_iterative(int N, uint32_t cur, uint32_t left, uint32_t right);

/* DFS nodes below one prefix in the GPU kernel's sense (src/n_queens_cuda.cu:39-72): one node
 * per queen placed at depths popcount(cur)+1 .. N-1 (the last row is counted by popcount and
 * is not a node). */
uint64_t oracle_count_nodes(int N, uint32_t cur, uint32_t left, uint32_t right);

/* Per-task counts over a task list (3 words per task), `threads` OpenMP threads (0 = the
 * runtime default).  per_task may be NULL.  Returns the plain sum (no symmetry fold). */
uint64_t oracle_count_tasks(int N, const uint32_t *tasks, uint64_t ntasks, uint64_t *per_task,
                            int threads);

/* Nodes over a task list, same threading. */
uint64_t oracle_count_tasks_nodes(int N, const uint32_t *tasks, uint64_t ntasks, int threads);

/* serial_n_queens (src/n_queens.cpp:115-137): generator + per-task count + x2 fold. */
uint64_t oracle_serial_n_queens(int N, int rows);

/* parallel_n_queens (src/n_queens.cpp:139-163): the same with `#pragma omp parallel for`
 * (default static schedule).  The reference fixes 32 threads (src/n_queens.cpp:152);
 * threads <= 0 selects that default. */
uint64_t oracle_parallel_n_queens(int N, int rows, int threads);

/* Instance prefix encoder: queens in rows 0..k-1 at columns cols[0..k-1], composed row by row
 * exactly as the recursion does (cur|p, (left|p)<<1, (right|p)>>1; src/n_queens.cpp:19).
 * Returns 0 and the masks, or -1 when a queen attacks an earlier one or a column is >= N. */
int oracle_encode_prefix(int N, int k, const int *cols, uint32_t *cur, uint32_t *left,
                         uint32_t *right);

#ifdef __cplusplus
}
#endif

#endif
