// C++ drop-in entry points of include/n_queens.h and include/utils.h, on top of the C ABI
// (nq.h).
#include "n_queens.h"
#include "utils.h"

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <sys/time.h>
#include <time.h>

#include <vector>

#include "nq.h"
#include "nq_internal.h"

// src/utils.cpp:10-12: milliseconds between two gettimeofday stamps.
float time_diff_ms(struct timeval &start, struct timeval &end) {
    return static_cast<float>((end.tv_sec - start.tv_sec) * 1000.0 + (end.tv_usec - start.tv_usec) / 1000.0);
}

// src/utils.cpp:14-30: the message behind a "[YYYY-mm-dd HH:MM:SS.mmm] " local-time stamp.
void print_with_time(const char *format, ...) {
    timeval tv;
    gettimeofday(&tv, nullptr);
    time_t secs = tv.tv_sec;
    tm t;
    localtime_r(&secs, &t);
    char stamp[40];
    strftime(stamp, sizeof stamp, "[%Y-%m-%d %H:%M:%S.", &t);
    fprintf(stdout, "%s%03d] ", stamp, static_cast<int>(tv.tv_usec / 1000));
    va_list ap;
    va_start(ap, format);
    vfprintf(stdout, format, ap);
    va_end(ap);
}

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

static void log_generated(void *, uint64_t ntasks, double ms) {
    print_with_time("Use %.2fms to generate %lld subproblems!\n", ms, static_cast<long long>(ntasks));
    print_with_time("gpu [%d] start job, with %lld(%.2f) subproblems.\n", 0, static_cast<long long>(ntasks), 1.0);
}

long long cuda_n_queens(int N, int rows) {
    if (nq_device_count() <= 0) {
        // The reference generates first and then finds no device (src/n_queens_cuda.cu:91-108).
        const double t0 = nq::now_ms();
        std::vector<uint32_t> tasks;
        nq::generate_tasks(N, rows, tasks);
        print_with_time("Use %.2fms to generate %lld subproblems!\n", nq::now_ms() - t0,
                     static_cast<long long>(tasks.size() / 3));
        printf("Failed to find any gpu!\n");  // src/n_queens_cuda.cu:105-108
        return -1;
    }
    // Tasks generated on the device (the same task set and count as the reference's host
    // generator), split as nq_solve splits them, counted, folded x2.
    uint64_t total = 0;
    int rc = nq::solve_on_device(N, rows, nullptr, &total, nullptr, log_generated, nullptr);
    if (rc != NQ_OK) {
        fprintf(stderr, "cuda_n_queens: %s\n", nq_last_error());
        return -1;
    }
    print_with_time("gpu [%d] finish job.\n", 0);
    return static_cast<long long>(total);
}
