// C++ drop-in entry points of include/n_queens.h, on top of the C ABI (nq.h).
#include "n_queens.h"

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <sys/time.h>
#include <time.h>

#include <vector>

#include "nq.h"
#include "nq_internal.h"

namespace nq {

// "[YYYY-mm-dd HH:MM:SS.mmm] " + message, the reference's log line format
// (print_with_time, src/utils.cpp:14-30).
void log_line(const char *fmt, ...) {
    timeval tv;
    gettimeofday(&tv, nullptr);
    time_t secs = tv.tv_sec;
    tm t;
    localtime_r(&secs, &t);
    char stamp[40];
    strftime(stamp, sizeof stamp, "[%Y-%m-%d %H:%M:%S.", &t);
    fprintf(stdout, "%s%03d] ", stamp, static_cast<int>(tv.tv_usec / 1000));
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stdout, fmt, ap);
    va_end(ap);
}

}  // namespace nq

static void append(vector<int> &tot, const std::vector<uint32_t> &w) {
    tot.reserve(tot.size() + w.size());
    for (uint32_t x : w) tot.push_back(static_cast<int>(x));
}

void partial_n_queens(int N, int cur, int left, int right, vector<int> &tot, int rows) {
    std::vector<uint32_t> w;
    nq::generate_from(N, false, static_cast<uint32_t>(cur), static_cast<uint32_t>(left),
                      static_cast<uint32_t>(right), rows, w);
    append(tot, w);
}

void partial_n_queens_for_odd(int N, int cur, int left, int right, vector<int> &tot, int rows) {
    std::vector<uint32_t> w;
    nq::generate_from(N, true, static_cast<uint32_t>(cur), static_cast<uint32_t>(left),
                      static_cast<uint32_t>(right), rows, w);
    append(tot, w);
}

long long cuda_n_queens(int N, int rows) {
    const double t0 = nq::now_ms();
    std::vector<uint32_t> tasks;
    nq::generate_tasks(N, rows, tasks);
    const uint64_t cnt = tasks.size() / 3;
    const double t1 = nq::now_ms();
    nq::log_line("Use %.2fms to generate %lld subproblems!\n", t1 - t0, static_cast<long long>(cnt));
    if (nq_device_count() <= 0) {
        printf("Failed to find any gpu!\n");  // src/n_queens_cuda.cu:105-108
        return -1;
    }
    nq::log_line("gpu [%d] start job, with %lld(%.2f) subproblems.\n", 0, static_cast<long long>(cnt), 1.0);
    uint64_t sum = 0;
    // The same device split as nq_solve (totals only; every child keeps its exact count).
    nq_config solve_cfg;
    nq_config c{};
    if (nq_solve_config(N, rows, &solve_cfg) == NQ_OK) c.split = solve_cfg.split;
    int rc = nq_count(N, tasks.data(), cnt, &c, &sum, nullptr, nullptr, nullptr);
    if (rc != NQ_OK) {
        fprintf(stderr, "cuda_n_queens: %s\n", nq_last_error());
        return -1;
    }
    nq::log_line("gpu [%d] finish job.\n", 0);
    return static_cast<long long>(2 * sum);  // src/n_queens_cuda.cu:181-183
}
