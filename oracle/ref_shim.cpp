// C-linkage wrappers over the reference's own CPU functions — TEST INFRASTRUCTURE ONLY.
//
// oracle/Makefile compiles this file together with the reference's src/n_queens.cpp and
// src/utils.cpp, where they lie under the reference checkout, into oracle/_ref/libnq_ref.so.
// Nothing from the reference is copied: the reference header is found through -I.  Tests use
// the library to pin the oracle restatement against the reference itself.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "n_queens.h"  // reference: src/n_queens.h:7-19

extern "C" {

// partial_n_queens + partial_n_queens_for_odd in the order of src/n_queens_cuda.cu:91-95.
uint64_t ref_gen_tasks(int N, int rows, uint32_t *out, uint64_t cap) {
    std::vector<int> tot;
    partial_n_queens(N, 0, 0, 0, tot, rows);
    if (N & 1) partial_n_queens_for_odd(N, 0, 0, 0, tot, rows);
    uint64_t n = tot.size() / 3;
    if (out) {
        uint64_t m = n < cap ? n : cap;
        memcpy(out, tot.data(), sizeof(uint32_t) * 3 * m);
    }
    return n;
}

// n_queens (src/n_queens.cpp:8-21); it accumulates into sum, so start from zero.
uint64_t ref_count_subtree(int N, uint32_t cur, uint32_t left, uint32_t right) {
    long long s = 0;
    n_queens(N, (int)cur, (int)left, (int)right, s);
    return (uint64_t)s;
}

// n_queens_iterative (src/n_queens.cpp:24-72), also accumulating into sum.
uint64_t ref_count_iterative(int N, uint32_t cur, uint32_t left, uint32_t right) {
    long long s = 0;
    n_queens_iterative(N, (int)cur, (int)left, (int)right, s);
    return (uint64_t)s;
}

uint64_t ref_serial_n_queens(int N, int rows) { return (uint64_t)serial_n_queens(N, rows); }

uint64_t ref_parallel_n_queens(int N, int rows) { return (uint64_t)parallel_n_queens(N, rows); }
}
