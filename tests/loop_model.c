/* CPU model of the DFS loop of n_queens_amd/csrc/nq_dfs.hip (NQ_HOT_ASM and its callers), test
 * infrastructure only.
 *
 * One workgroup of B lanes runs the kernel's per-lane iteration in lockstep trips of TRIP (16)
 * iterations over an LDS byte array with the kernel's layout (lds_layout(): a dead frame at
 * address 0, then stack[levels][B] of 16-B frames).  Between trips, lanes whose sp is 0 (parked
 * on the dead frame) take the next task, as the kernel's refill does.  It checks:
 *   - every LDS write and read is the lane's own slot at a level below `levels`, or the dead
 *     frame at address 0 (and a write there leaves the dead frame unchanged);
 *   - every lane that ends a task reaches sp == 0 with v == 0;
 *   - every task's count equals a plain recursion (src/n_queens.cpp:8-21 semantics), and so
 *     does its node count (expansions minus completed boards: the node-counting kernel's sum);
 *   - a lane completes at most one board per iteration (the kernel counts once per iteration).
 * Both push forms: masked (write only when candidates remain) and unconditional (always write,
 * advance by min(v, 1), one spare level).  Each iteration makes the forced move: a child with
 * exactly one candidate has it placed at once, with no push.
 *
 * Build flags plant a mistake, to show that the checks catch it:
 *   -DPLANT_NO_CLAMP     the pop subtracts without saturating (reads below the stack);
 *   -DPLANT_FULL_DEAD    the dead frame has c == ~0 (parked lanes count boards);
 *   -DPLANT_FORCED_DEAD  the forced move is also made from a child with no candidate;
 *   -DPLANT_COUNT_FIRST  boards are counted before the forced move (its boards are missed).
 * Exit status 0 when every case passes. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t c, l, r, v;
} frame;

#define TRIP 16 /* iterations per trip of the hot loop (NQ_HOT_ASM) */

#ifdef PLANT_FULL_DEAD
static const frame DEAD = {0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u};
#else
static const frame DEAD = {0xFFFFFFFEu, 0u, 0xFFFFFFFFu, 0u};
#endif

static uint64_t completions(int N, uint32_t c, uint32_t l, uint32_t r) {
    const uint32_t last = (1u << N) - 1u;
    if (c == last) return 1;
    uint64_t s = 0;
    for (uint32_t v = last & ~(c | l | r); v; v &= v - 1) {
        const uint32_t p = v & (0u - v);
        s += completions(N, c | p, (l | p) << 1, (r | p) >> 1);
    }
    return s;
}

/* DFS nodes below a prefix in the reference kernel's sense: one per queen placed at depths
 * popcount(c)+1 .. N-1 (src/n_queens_cuda.cu:39-72). */
static uint64_t nodes_below(int N, uint32_t c, uint32_t l, uint32_t r) {
    const uint32_t last = (1u << N) - 1u;
    uint64_t s = 0;
    for (uint32_t v = last & ~(c | l | r); v; v &= v - 1) {
        const uint32_t p = v & (0u - v);
        if ((c | p) != last) s += 1 + nodes_below(N, c | p, (l | p) << 1, (r | p) >> 1);
    }
    return s;
}

static uint32_t *tasks;
static size_t ntasks, cap;

static void prefixes(int N, int rows, int row, uint32_t c, uint32_t l, uint32_t r) {
    if (row == rows) {
        if (ntasks == cap) {
            cap = cap ? 2 * cap : 1024;
            tasks = realloc(tasks, cap * 3 * sizeof(uint32_t));
        }
        tasks[3 * ntasks] = c;
        tasks[3 * ntasks + 1] = l;
        tasks[3 * ntasks + 2] = r;
        ntasks++;
        return;
    }
    const uint32_t last = (1u << N) - 1u;
    for (uint32_t v = last & ~(c | l | r); v; v &= v - 1) {
        const uint32_t p = v & (0u - v);
        prefixes(N, rows, row + 1, c | p, (l | p) << 1, (r | p) >> 1);
    }
}

/* Address a of lane i: 0 (dead frame) or 16 + level * stride + 16 * i with level < levels. */
static int owned(uint32_t a, int i, uint32_t stride, int levels) {
    if (a == 0) return 1;
    if (a < 16) return 0;
    const uint32_t off = a - 16;
    return off % stride == 16u * (uint32_t)i && off / stride < (uint32_t)levels;
}

static int run(int N, int rows, int B, int apush) {
    ntasks = 0;
    prefixes(N, rows, 0, 0, 0, 0);
    const uint32_t last = (1u << N) - 1u;
    const int depth = N - 1 - rows; /* frames a task of this depth needs */
    const int levels = depth + apush;
    const uint32_t stride = 16u * (uint32_t)B;
    const size_t lds_bytes = 16 + (size_t)B * (size_t)levels * 16;
    uint8_t *lds = calloc(lds_bytes, 1);
    memcpy(lds, &DEAD, 16);
    frame *f = calloc((size_t)B, sizeof(frame));
    uint32_t *sp = calloc((size_t)B, sizeof(uint32_t));
    uint64_t *cnt = calloc((size_t)B, sizeof(uint64_t));
    uint64_t *nd = calloc((size_t)B, sizeof(uint64_t)); /* expansions (NQ_NODE_INC, NQ_NODE_FORCED) */
    uint64_t *gotn = calloc(ntasks ? ntasks : 1, sizeof(uint64_t));
    long *tid = malloc((size_t)B * sizeof(long));
    int *has = calloc((size_t)B, sizeof(int));
    uint64_t *got = calloc(ntasks ? ntasks : 1, sizeof(uint64_t));
    size_t cursor = 0;
    long trips = 0;
    int rc = 0;
    for (int i = 0; i < B; i++) {
        tid[i] = -1;
        has[i] = 1;
    }
    for (;;) {
        int any = 0;
        for (int i = 0; i < B; i++) { /* refill (dfs_kernel's claim loop) */
            if (has[i] && f[i].v == 0) {
                if (tid[i] >= 0) {
                    got[tid[i]] = cnt[i];
                    gotn[tid[i]] = nd[i] - cnt[i];
                }
                cnt[i] = 0;
                nd[i] = 0;
                has[i] = 0;
                tid[i] = -1;
                while (cursor < ntasks) {
                    const size_t t = cursor++;
                    const uint32_t c0 = tasks[3 * t], l0 = tasks[3 * t + 1], r0 = tasks[3 * t + 2];
                    const uint32_t v0 = last & ~(c0 | l0 | r0);
                    const uint32_t d0 = (uint32_t)__builtin_popcount(c0);
                    if (d0 + 1 >= (uint32_t)N || v0 == 0) {
                        got[t] = d0 >= (uint32_t)N ? (c0 == last)
                                                   : (d0 + 1 == (uint32_t)N ? (uint64_t)__builtin_popcount(v0) : 0);
                        continue;
                    }
                    f[i] = (frame){c0 | ~last, l0, r0, v0};
                    sp[i] = 16u + 16u * (uint32_t)i;
                    tid[i] = (long)t;
                    has[i] = 1;
                    break;
                }
            }
            any |= has[i];
        }
        if (!any) break;
        for (;;) { /* trips of TRIP iterations until some lane is parked */
            if (++trips > 50000000L) {
                fprintf(stderr, "N=%d rows=%d B=%d apush=%d: no termination\n", N, rows, B, apush);
                return 1;
            }
            for (int it = 0; it < TRIP; it++)
                for (int i = 0; i < B; i++) {
                    if (!has[i]) continue;
                    frame *x = &f[i];
                    if (sp[i] != 0) nd[i]++; /* NQ_NODE_INC */
                    const uint32_t p = x->v & (0u - x->v); /* NQ_EXTRACT */
                    x->v ^= p;
                    if (apush || x->v != 0) { /* NQ_PUSH_ALWAYS / NQ_PUSH_MASKED */
                        const uint32_t a = sp[i];
                        if (!owned(a, i, stride, levels) || a + 16 > lds_bytes ||
                            (a == 0 && memcmp(x, &DEAD, 16) != 0)) {
                            fprintf(stderr, "N=%d rows=%d B=%d apush=%d: bad write lane %d at %u\n", N, rows, B,
                                    apush, i, a);
                            return 1;
                        }
                        memcpy(lds + a, x, 16);
                        sp[i] += apush ? stride * (x->v != 0) : stride;
                    }
                    x->c |= p; /* NQ_CHILD */
                    x->l = (x->l + p) << 1;
                    x->r = (x->r | p) >> 1;
                    x->v = ~(x->l | x->r | x->c);
                    const int full_first = x->c == 0xFFFFFFFFu;
#ifdef PLANT_COUNT_FIRST
                    if (full_first) cnt[i]++;
#endif
#ifdef PLANT_FORCED_DEAD
                    if ((x->v & (x->v - 1)) == 0) { /* NQ_FORCED with the empty set let in */
#else
                    if (__builtin_popcount(x->v) == 1) { /* NQ_FORCED: v_bcnt, v_cmpx_eq 1 */
#endif
                        const uint32_t q = x->v;
                        nd[i]++; /* NQ_NODE_FORCED */
                        x->c |= q;
                        x->l = (x->l + q) << 1;
                        x->r = (x->r | q) >> 1;
                        x->v = ~(x->l | x->r | x->c);
                        if (full_first && x->c == 0xFFFFFFFFu) {
                            fprintf(stderr, "lane %d: two boards in one iteration\n", i);
                            return 1;
                        }
                    }
#ifndef PLANT_COUNT_FIRST
                    if (x->c == 0xFFFFFFFFu) cnt[i]++; /* NQ_COUNT_POP */
#endif
                    if (x->v == 0) {
#ifdef PLANT_NO_CLAMP
                        sp[i] -= stride;
#else
                        sp[i] = sp[i] >= stride ? sp[i] - stride : 0; /* v_sub_u32 ... clamp */
#endif
                        const uint32_t a = sp[i];
                        if (!owned(a, i, stride, levels) || a + 16 > lds_bytes) {
                            fprintf(stderr, "N=%d rows=%d B=%d apush=%d: bad read lane %d at %u\n", N, rows, B,
                                    apush, i, a);
                            return 1;
                        }
                        memcpy(x, lds + a, 16);
                    }
                }
            int parked = 0;
            for (int i = 0; i < B; i++)
                if (has[i] && sp[i] == 0) {
                    parked = 1;
                    if (f[i].v != 0) {
                        fprintf(stderr, "lane %d parked with candidates\n", i);
                        return 1;
                    }
                }
            if (parked) break;
        }
    }
    uint64_t sum = 0;
    for (size_t t = 0; t < ntasks; t++) {
        const uint64_t want = completions(N, tasks[3 * t], tasks[3 * t + 1], tasks[3 * t + 2]);
        if (got[t] != want) {
            fprintf(stderr, "N=%d rows=%d B=%d apush=%d: task %zu counted %llu, expected %llu\n", N, rows, B, apush,
                    t, (unsigned long long)got[t], (unsigned long long)want);
            rc = 1;
            break;
        }
        /* Tasks counted at claim time (no DFS) have no nodes below them in the kernel's sense. */
        const uint32_t c0 = tasks[3 * t], l0 = tasks[3 * t + 1], r0 = tasks[3 * t + 2];
        const uint32_t d0 = (uint32_t)__builtin_popcount(c0);
        const int direct = d0 + 1 >= (uint32_t)N || (last & ~(c0 | l0 | r0)) == 0;
        const uint64_t wantn = direct ? 0 : nodes_below(N, c0, l0, r0);
        if (gotn[t] != wantn) {
            fprintf(stderr, "N=%d rows=%d B=%d apush=%d: task %zu nodes %llu, expected %llu\n", N, rows, B, apush, t,
                    (unsigned long long)gotn[t], (unsigned long long)wantn);
            rc = 1;
            break;
        }
        sum += got[t];
    }
    if (!rc) printf("N=%d rows=%d B=%d apush=%d tasks=%zu sum=%llu ok\n", N, rows, B, apush, ntasks,
                    (unsigned long long)sum);
    free(lds);
    free(f);
    free(sp);
    free(cnt);
    free(nd);
    free(gotn);
    free(tid);
    free(has);
    free(got);
    return rc;
}

int main(int argc, char **argv) {
    const int nmax = argc > 1 ? atoi(argv[1]) : 13;
    for (int N = 4; N <= nmax; N++)
        for (int rows = 1; rows <= 4 && rows < N - 1; rows++)
            for (int B = 64; B <= 256; B += 64)
                for (int ap = 0; ap < 2; ap++)
                    if (run(N, rows, B, ap)) return 1;
    return 0;
}
