// Resumable whole-board counts (nq_solve_resumable): the reference's task list is counted in
// consecutive ranges of tasks, and a small text checkpoint is rewritten atomically after each
// range.  The reference has no counterpart: its N=27 run took 28.4 days with no way to resume
// (README.md:46).
//
// The task list is the host generator's (reference order, src/n_queens.cpp:74-113), so a range
// [first, first + n) names the same tasks in every process; each range is counted by nq_count
// (device split included), and the sum of the ranges is the sum over the list.
#include <errno.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <string>
#include <vector>

#include "nq.h"
#include "nq_internal.h"
#include "nq_kernels.h"

namespace nq {
namespace {

constexpr const char *kMagic = "nq-checkpoint 1";

struct Checkpoint {
    int N = 0, rows = 0;
    uint64_t ntasks = 0, next = 0, sum = 0;
};

// Reads a checkpoint; returns 1 if the file exists and parses, 0 if it does not exist, or an
// error status (the file exists but is not a checkpoint of this format).
int read_checkpoint(const char *path, Checkpoint *cp) {
    FILE *f = fopen(path, "r");
    if (!f) return errno == ENOENT ? 0 : fail(NQ_ERR_ARG, std::string("cannot open checkpoint ") + path);
    char magic[64] = {0};
    unsigned long long ntasks = 0, next = 0, sum = 0;
    int N = 0, rows = 0;
    const bool ok = fgets(magic, sizeof magic, f) && strncmp(magic, kMagic, strlen(kMagic)) == 0 &&
                    fscanf(f, " N %d rows %d ntasks %llu next %llu sum %llu", &N, &rows, &ntasks, &next, &sum) == 5;
    fclose(f);
    if (!ok) return fail(NQ_ERR_ARG, std::string("not a checkpoint file: ") + path);
    cp->N = N;
    cp->rows = rows;
    cp->ntasks = ntasks;
    cp->next = next;
    cp->sum = sum;
    return 1;
}

// Writes the checkpoint to path.tmp, flushes it to disk and renames it over path, so that a
// crash leaves either the previous or the new checkpoint.
int write_checkpoint(const char *path, const Checkpoint &cp) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return fail(NQ_ERR_ARG, "cannot write checkpoint " + tmp);
    fprintf(f, "%s\nN %d rows %d ntasks %llu next %llu sum %llu\n", kMagic, cp.N, cp.rows,
            static_cast<unsigned long long>(cp.ntasks), static_cast<unsigned long long>(cp.next),
            static_cast<unsigned long long>(cp.sum));
    const bool ok = fflush(f) == 0 && fsync(fileno(f)) == 0;
    if (fclose(f) != 0 || !ok) return fail(NQ_ERR_ARG, "cannot write checkpoint " + tmp);
    if (rename(tmp.c_str(), path) != 0) return fail(NQ_ERR_ARG, std::string("cannot replace checkpoint ") + path);
    return NQ_OK;
}

}  // namespace
}  // namespace nq

using namespace nq;

extern "C" {

int nq_solve_resumable(int N, int rows, const nq_config *cfg, const char *path, uint64_t range_tasks, int max_ranges,
                       nq_resume_state *state) {
    if (!path || !*path) return fail(NQ_ERR_ARG, "checkpoint path is empty");
    if (!state) return fail(NQ_ERR_ARG, "state is NULL");
    if (N < 1 || N > MAX_N || rows < 1 || rows > N) return fail(NQ_ERR_ARG, "N must be 1..28 and rows 1..N");
    if (max_ranges < 0) return fail(NQ_ERR_ARG, "max_ranges must be >= 0");
    memset(state, 0, sizeof *state);
    Checkpoint cp;
    const int have = read_checkpoint(path, &cp);
    if (have < 0) return have;
    if (have && (cp.N != N || cp.rows != rows))
        return fail(NQ_ERR_ARG, "checkpoint is for a different problem (N, rows)");
    // The task count is checked once the list exists; an out-of-range cursor is rejected here.
    if (have && cp.next > cp.ntasks) return fail(NQ_ERR_ARG, "checkpoint cursor is past its task count");

    const double t0 = now_ms();
    std::vector<uint32_t> tasks;
    generate_tasks(N, rows, tasks);
    const uint64_t n = tasks.size() / 3;
    if (have && cp.ntasks != n) return fail(NQ_ERR_ARG, "checkpoint task count does not match this generator");
    if (!have) {
        cp.N = N;
        cp.rows = rows;
        cp.ntasks = n;
        cp.next = 0;
        cp.sum = 0;
        int rc = write_checkpoint(path, cp);
        if (rc) return rc;
    }
    state->ntasks = n;
    state->gen_ms = now_ms() - t0;
    const uint64_t step = range_tasks ? range_tasks : std::max<uint64_t>(1, (n + 99) / 100);
    nq_config c = cfg ? *cfg : nq_config{};
    if (!cfg || c.split < 0) {
        nq_config sc;
        int rc = nq_solve_config(N, rows, &sc);
        if (rc) return rc;
        c.split = sc.split;
    }
    int ranges = 0;
    while (cp.next < n && (max_ranges == 0 || ranges < max_ranges)) {
        const uint64_t len = std::min(step, n - cp.next);
        uint64_t sum = 0;
        nq_stats st;
        int rc = nq_count(N, tasks.data() + 3 * cp.next, len, &c, &sum, nullptr, nullptr, &st);
        if (rc) return rc;
        cp.next += len;
        cp.sum += sum;
        rc = write_checkpoint(path, cp);
        if (rc) return rc;
        state->kernel_ms += st.kernel_ms;
        ++ranges;
    }
    state->next = cp.next;
    state->sum = cp.sum;
    state->ranges = ranges;
    state->done = cp.next == n ? 1 : 0;
    state->total = state->done ? 2 * cp.sum : 0;  // symmetry fold, src/n_queens_cuda.cu:181-183
    return NQ_OK;
}

}  // extern "C"
