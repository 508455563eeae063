// CPU entry points of the reference header (src/n_queens.h:7-17), for callers that use them
// beside cuda_n_queens and for the CLI's --serial / --openmp cross-checks (the blocks
// commented out at src/main.cpp:35-47).  They run on the host only; nothing on the GPU path
// calls them.
//
//   n_queens(N, cur, left, right, sum)  adds the completions of the prefix (cur, left, right)
//                                       to sum (src/n_queens.cpp:8-21; sum accumulates)
//   n_queens_iterative(N, cur, left, right, sum)
//                                       the reference's explicit-stack variant
//                                       (src/n_queens.cpp:24-72): the same sum, except that a
//                                       prefix already holding N-1 or N queens adds nothing
//   serial_n_queens(N, rows)            generator + every task + x2 (src/n_queens.cpp:115-135)
//   parallel_n_queens(N, rows)          the same over 32 threads (src/n_queens.cpp:137-163)
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "n_queens.h"
#include "nq_internal.h"

namespace {

// Completions below one prefix: a bitmask DFS over an explicit stack of remaining candidate
// sets (a completed board is cur == last, as in the reference's recursion).
uint64_t count_prefix(int N, uint32_t cur, uint32_t left, uint32_t right) {
    const uint32_t last = (1u << N) - 1u;
    if (cur == last) return 1;  // src/n_queens.cpp:10-13
    struct Frame {
        uint32_t c, l, r, v;
    };
    Frame st[33];
    int top = 0;
    st[0] = {cur, left, right, last & ~(cur | left | right)};
    uint64_t count = 0;
    while (top >= 0) {
        Frame &f = st[top];
        if (f.v == 0) {
            --top;
            continue;
        }
        const uint32_t p = f.v & (0u - f.v);
        f.v ^= p;
        const uint32_t c = f.c | p;
        if (c == last) {
            ++count;
            continue;
        }
        const uint32_t l = (f.l | p) << 1, r = (f.r | p) >> 1;
        const uint32_t v = last & ~(c | l | r);
        if (v) st[++top] = {c, l, r, v};
    }
    return count;
}

long long fold_tasks(int N, int rows, int threads) {
    std::vector<uint32_t> tasks;
    nq::generate_tasks(N, rows, tasks);
    const size_t n = tasks.size() / 3;
    std::vector<uint64_t> partial(n, 0);
    auto work = [&](std::atomic<size_t> *next) {
        for (size_t i = next->fetch_add(1); i < n; i = next->fetch_add(1))
            partial[i] = count_prefix(N, tasks[3 * i], tasks[3 * i + 1], tasks[3 * i + 2]);
    };
    std::atomic<size_t> next{0};
    if (threads <= 1) {
        work(&next);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) pool.emplace_back(work, &next);
        for (auto &th : pool) th.join();
    }
    long long sum = 0;
    for (uint64_t x : partial) sum += static_cast<long long>(x) * 2;  // src/n_queens.cpp:132-134
    return sum;
}

}  // namespace

void n_queens(int N, int cur, int left, int right, long long &sum) {
    sum += static_cast<long long>(
        count_prefix(N, static_cast<uint32_t>(cur), static_cast<uint32_t>(left), static_cast<uint32_t>(right)));
}

// The reference's loop counts a board when a placement leaves one free column that is still a
// candidate (src/n_queens.cpp:62-66), so the placement that fills the board is never reached
// from a prefix of N-1 queens, and a full board has no candidate to start from (:29-31).  Any
// other prefix gets its exact completion count.  Its 16-frame stack (int stack[64], :26)
// overflows more than 16 rows above the last row; the count here has no such limit.
void n_queens_iterative(int N, int cur, int left, int right, long long &sum) {
    const uint32_t last = (1u << N) - 1u;
    const uint32_t c = static_cast<uint32_t>(cur);
    if (__builtin_popcount(c & last) + 1 >= N) return;
    sum += static_cast<long long>(count_prefix(N, c, static_cast<uint32_t>(left), static_cast<uint32_t>(right)));
}

long long serial_n_queens(int N, int rows) { return fold_tasks(N, rows, 1); }

long long parallel_n_queens(int N, int rows) { return fold_tasks(N, rows, 32); }  // omp_set_num_threads(32)
