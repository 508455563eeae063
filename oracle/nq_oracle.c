/*
 * CPU oracle for the N-Queens counting path — TEST INFRASTRUCTURE ONLY (see nq_oracle.h).
 *
 * A plain C restatement of the reference's CPU algorithms.  It is deliberately written the
 * way the reference computes (recursive bitmask backtracking, lowest-set-bit-first order) so
 * that it can be read side by side with src/n_queens.cpp; it is never the thing measured on
 * the GPU and never shipped in the product library.
 */
#include "nq_oracle.h"

#include <omp.h>
#include <stdlib.h>

static inline uint32_t board_mask(int N) { return N >= 32 ? 0xFFFFFFFFu : ((1u << N) - 1u); }

/* ---- completions: src/n_queens.cpp:8-21 ---------------------------------------------- */
static void count_rec(uint32_t last, uint32_t cur, uint32_t left, uint32_t right, uint64_t *sum) {
    if (cur == last) { /* n_queens.cpp:10-13 */
        (*sum)++;
        return;
    }
    uint32_t valid = last & ~(cur | left | right); /* n_queens.cpp:15 */
    while (valid) {                                /* n_queens.cpp:16-20 */
        uint32_t p = valid & (0u - valid);
        valid -= p;
        count_rec(last, cur | p, (left | p) << 1, (right | p) >> 1, sum);
    }
}

uint64_t oracle_count_subtree(int N, uint32_t cur, uint32_t left, uint32_t right) {
    uint64_t s = 0;
    count_rec(board_mask(N), cur, left, right, &s);
    return s;
}

/* ---- iterative counter: src/n_queens.cpp:24-72 ---------------------------------------
 * The reference's explicit-stack variant.  It counts a board when a placement leaves exactly
 * one free column that is still a candidate (:62-66), so a prefix that already holds N-1 or
 * N queens adds nothing (its first placement fills the board and finds no candidate, :56-60),
 * unlike n_queens above, which counts a full board as 1.  The reference's stack holds 16
 * frames (int stack[64], :26); this restatement's holds 32, so it also covers the deeper
 * prefixes on which the reference would overflow. */
uint64_t oracle_count_iterative(int N, uint32_t cur, uint32_t left, uint32_t right) {
    uint32_t last = board_mask(N);
    uint32_t st[4 * 33];
    int top = 0;
    uint64_t sum = 0;
    uint32_t valid = last & ~(cur | left | right); /* :29 */
    if (valid == 0) return 0;                      /* :31 */
    st[top++] = cur;
    st[top++] = left;
    st[top++] = right;
    st[top++] = valid;
    while (top != 0) { /* :38 */
        valid = st[top - 1];
        right = st[top - 2];
        left = st[top - 3];
        cur = st[top - 4];
        uint32_t p = valid & (0u - valid); /* :44-45 */
        valid -= p;
        if (valid == 0) /* :47-51 */
            top -= 4;
        else
            st[top - 1] = valid;
        cur |= p; /* :53-56 */
        left = (left | p) << 1;
        right = (right | p) >> 1;
        valid = last & ~(cur | left | right);
        if (valid == 0) continue; /* :58-60 */
        p = cur ^ last;           /* :62-66 */
        if ((p & (p - 1)) == 0) {
            sum++;
            continue;
        }
        st[top++] = cur; /* :68-71 */
        st[top++] = left;
        st[top++] = right;
        st[top++] = valid;
    }
    return sum;
}

/* ---- kernel node count: src/n_queens_cuda.cu:39-72 ------------------------------------
 * Every loop iteration of the reference kernel places one queen at depth popc(cur)+1; a
 * placement that reaches depth N-1 ends the branch and adds popc(valid) (the last row), so
 * nodes are the placements at depths rows+1 .. N-1. */
static void nodes_rec(int N, uint32_t last, uint32_t cur, uint32_t left, uint32_t right, int depth,
                      uint64_t *nodes) {
    if (depth >= N - 1) return;
    uint32_t valid = last & ~(cur | left | right);
    while (valid) {
        uint32_t p = valid & (0u - valid);
        valid -= p;
        (*nodes)++;
        nodes_rec(N, last, cur | p, (left | p) << 1, (right | p) >> 1, depth + 1, nodes);
    }
}

uint64_t oracle_count_nodes(int N, uint32_t cur, uint32_t left, uint32_t right) {
    uint64_t n = 0;
    nodes_rec(N, board_mask(N), cur, left, right, __builtin_popcount(cur), &n);
    return n;
}

/* ---- generator: src/n_queens.cpp:74-113 ----------------------------------------------- */
typedef struct {
    uint32_t *out;
    uint64_t cap;
    uint64_t n;
} task_sink;

static inline void emit(task_sink *s, uint32_t c, uint32_t l, uint32_t r) {
    if (s->out && s->n < s->cap) {
        s->out[3 * s->n + 0] = c;
        s->out[3 * s->n + 1] = l;
        s->out[3 * s->n + 2] = r;
    }
    s->n++;
}

/* partial_n_queens, src/n_queens.cpp:74-91: row 0 restricted to columns < N/2. */
static void gen_even(int N, uint32_t cur, uint32_t left, uint32_t right, int rows, task_sink *s) {
    uint32_t last = board_mask(N);
    if (cur == 0) last = (1u << (N / 2)) - 1u; /* :76-78 */
    uint32_t valid = last & ~(cur | left | right);
    while (valid) {
        uint32_t p = valid & (0u - valid);
        valid -= p;
        if (rows == 1) { /* :83-87 */
            emit(s, cur | p, (left | p) << 1, (right | p) >> 1);
            continue;
        }
        gen_even(N, cur | p, (left | p) << 1, (right | p) >> 1, rows - 1, s);
    }
}

/* partial_n_queens_for_odd, src/n_queens.cpp:93-113: row 0 = middle column, row 1 restricted
 * to columns < (N-2)/2. */
static void gen_odd(int N, uint32_t cur, uint32_t left, uint32_t right, int rows, task_sink *s) {
    uint32_t last = board_mask(N);
    if (cur == 0)
        last = 1u << (N / 2); /* :95-96 */
    else if ((cur & (cur - 1)) == 0)
        last = (1u << ((N - 2) / 2)) - 1u; /* :97-98 */
    uint32_t valid = last & ~(cur | left | right);
    while (valid) {
        uint32_t p = valid & (0u - valid);
        valid -= p;
        if (rows == 1) {
            emit(s, cur | p, (left | p) << 1, (right | p) >> 1);
            continue;
        }
        gen_odd(N, cur | p, (left | p) << 1, (right | p) >> 1, rows - 1, s);
    }
}

uint64_t oracle_gen_tasks(int N, int rows, uint32_t *out, uint64_t cap) {
    task_sink s = {out, cap, 0};
    if (N < 1 || N > 31 || rows < 1) return 0;
    gen_even(N, 0, 0, 0, rows, &s); /* n_queens_cuda.cu:91 */
    if (N & 1) gen_odd(N, 0, 0, 0, rows, &s); /* n_queens_cuda.cu:93-95 */
    return s.n;
}

/* ---- per-task drivers: src/n_queens.cpp:115-163 ---------------------------------------- */
uint64_t oracle_count_tasks(int N, const uint32_t *tasks, uint64_t ntasks, uint64_t *per_task,
                            int threads) {
    uint64_t sum = 0;
    uint32_t last = board_mask(N);
    if (threads > 0) omp_set_num_threads(threads);
    long long n = (long long)ntasks;
    /* default static schedule, as n_queens.cpp:153; the loop index is 64-bit (the reference's
     * int index overflows above 2^31/3 tasks, n_queens.cpp:128,154). */
#pragma omp parallel for reduction(+ : sum)
    for (long long i = 0; i < n; i++) {
        uint64_t s = 0;
        count_rec(last, tasks[3 * i], tasks[3 * i + 1], tasks[3 * i + 2], &s);
        if (per_task) per_task[i] = s;
        sum += s;
    }
    return sum;
}

uint64_t oracle_count_tasks_nodes(int N, const uint32_t *tasks, uint64_t ntasks, int threads) {
    uint64_t sum = 0;
    if (threads > 0) omp_set_num_threads(threads);
    long long n = (long long)ntasks;
#pragma omp parallel for reduction(+ : sum) schedule(dynamic, 16)
    for (long long i = 0; i < n; i++)
        sum += oracle_count_nodes(N, tasks[3 * i], tasks[3 * i + 1], tasks[3 * i + 2]);
    return sum;
}

static uint64_t drive(int N, int rows, int threads) {
    uint64_t n = oracle_gen_tasks(N, rows, NULL, 0);
    uint32_t *t = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (n ? n : 1));
    if (!t) return 0;
    oracle_gen_tasks(N, rows, t, n);
    uint64_t s = oracle_count_tasks(N, t, n, NULL, threads);
    free(t);
    return 2 * s; /* n_queens.cpp:132-134 / 158-160 */
}

uint64_t oracle_serial_n_queens(int N, int rows) { return drive(N, rows, 1); }

uint64_t oracle_parallel_n_queens(int N, int rows, int threads) {
    return drive(N, rows, threads > 0 ? threads : 32); /* n_queens.cpp:152 */
}

/* ---- instance prefixes ----------------------------------------------------------------- */
int oracle_encode_prefix(int N, int k, const int *cols, uint32_t *cur, uint32_t *left,
                         uint32_t *right) {
    uint32_t c = 0, l = 0, r = 0, last = board_mask(N);
    for (int i = 0; i < k; i++) {
        if (cols[i] < 0 || cols[i] >= N) return -1;
        uint32_t p = 1u << cols[i];
        if (!(p & last & ~(c | l | r))) return -1;
        c |= p;
        l = (l | p) << 1;
        r = (r | p) >> 1;
    }
    *cur = c;
    *left = l;
    *right = r;
    return 0;
}
