// `n_queens N rows [--config 1|2|3] [--checkpoint PATH [--range-tasks K]]` — drop-in for the
// reference CLI (src/main.cpp:8-57).
//
// Same positional arguments, same validation order and messages (src/main.cpp:10-29), same
// timed call and result line (src/main.cpp:49-54).  The reference picks a (threads, stack
// depth) preset at compile time (-D_USE_CONFIGn_, src/macro.h:25-35; compile.sh uses CONFIG2);
// here --config selects the preset at run time and only its stack depth is used, for the
// legality check.  Missing arguments print a usage line (the reference reads argv unchecked).
// --checkpoint (no reference counterpart) counts through nq_solve_resumable: the run can be
// stopped at any time and resumed by running the same command again.  --serial and --openmp
// first run the reference's host cross-checks (commented out at src/main.cpp:35-47), printing
// its "serial ..." / "parallel ..." lines.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "n_queens.h"
#include "nq.h"
#include "utils.h"

int main(int argc, char *argv[]) {
    int config = 2;
    const char *checkpoint = nullptr;
    bool serial = false, openmp = false;
    unsigned long long range_tasks = 0;
    int pos[2];
    int npos = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--config") && i + 1 < argc) {
            config = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--checkpoint") && i + 1 < argc) {
            checkpoint = argv[++i];
        } else if (!strcmp(argv[i], "--serial")) {
            serial = true;
        } else if (!strcmp(argv[i], "--openmp")) {
            openmp = true;
        } else if (!strcmp(argv[i], "--range-tasks") && i + 1 < argc) {
            range_tasks = strtoull(argv[++i], nullptr, 10);
        } else if (npos < 2) {
            pos[npos++] = atoi(argv[i]);
        }
    }
    if (npos < 2 || config < 1 || config > 3) {
        fprintf(stderr, "usage: %s N rows [--config 1|2|3] [--serial] [--openmp] [--checkpoint PATH [--range-tasks K]]\n",
                argv[0]);
        return 2;
    }
    static const int kStack[4] = {0, 24, 19, 16};  // STACKSIZE of CONFIG1/2/3, src/macro.h:25-35
    const int STACKSIZE = kStack[config];

    int N = pos[0];
    if (N < 2 || N > 28) {
        printf("Not allowed %d queens problem!\n", N);
        return -1;
    }
    int rows = pos[1];
    if (STACKSIZE + 1 + rows < N) {
        printf("This configure cannot process %d-queens with pre-placing %d rows. Pre-placed rows should be >= %d.\n",
               N, rows, N - STACKSIZE - 1);
        return -1;
    }
    if (N - rows < 2) {
        printf("Pre-placed rows should be <= %d.\n", N - 2);
        return -1;
    }
    if (rows < 2) {
        printf("Pre-placed rows shold be at least 2, so set to 2!\n");
        rows = 2;
    }

    timeval start, end;
    long long cpu_sum = 0;
    if (serial) {
        gettimeofday(&start, nullptr);
        cpu_sum = serial_n_queens(N, rows);
        gettimeofday(&end, nullptr);
        printf("serial %d queens result %lld, calc time: [%.2fms]\n", N, cpu_sum, time_diff_ms(start, end));
    }
    if (openmp) {
        gettimeofday(&start, nullptr);
        cpu_sum = parallel_n_queens(N, rows);
        gettimeofday(&end, nullptr);
        printf("parallel %d queens result %lld, calc time: [%.2fms]\n", N, cpu_sum, time_diff_ms(start, end));
    }
    print_with_time("===============================================================\n");
    gettimeofday(&start, nullptr);
    long long sum = -1;
    if (checkpoint) {
        // One range per call, so that every counted range prints a progress line (a long run
        // would otherwise be silent for hours); each call resumes from the file the last one
        // wrote, after regenerating the host task list.
        nq_resume_state st;
        int rc = NQ_OK, ranges = 0;
        do {
            rc = nq_solve_resumable(N, rows, nullptr, checkpoint, range_tasks, 1, &st);
            if (rc != NQ_OK) break;
            ranges += st.ranges;
            if (st.ranges) {
                print_with_time("checkpoint %s: tasks %llu of %llu done\n", checkpoint,
                                static_cast<unsigned long long>(st.next), static_cast<unsigned long long>(st.ntasks));
                fflush(stdout);  // seen at once also when stdout is a file or a pipe
            }
        } while (!st.done);
        if (rc == NQ_OK) {
            print_with_time("checkpoint %s: %d ranges counted in this run, tasks %llu of %llu done\n", checkpoint,
                            ranges, static_cast<unsigned long long>(st.next), static_cast<unsigned long long>(st.ntasks));
            sum = static_cast<long long>(st.total);
        } else {
            fprintf(stderr, "n_queens: %s\n", nq_last_error());
        }
    } else {
        sum = cuda_n_queens(N, rows);
    }
    gettimeofday(&end, nullptr);
    print_with_time("cuda %d queens result %lld, calc time: [%.2fms]\n", N, sum, time_diff_ms(start, end));
    print_with_time("===============================================================\n");
    return 0;
}
