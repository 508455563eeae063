// CPU model of split_kernel's control flow (n_queens_amd/csrc/nq_dfs.hip), built around the
// kernel's own TaskWalk (n_queens_amd/csrc/nq_walk.h, compiled here for the host).
//
// Waves share one task cursor and claim guided shares of the tasks left (at least `chunk`); each
// wave walks its task's top `split` rows and hands every root to its lowest lane that is parked
// (no root, or the last one exhausted), exactly as the kernel's refill loop does.  A lane's DFS is modelled by its work:
// the root's expansions, consumed 16 per trip, and its count, the root's completions (both from
// a plain recursion).  Waves advance round-robin, one refill + trips-until-a-lane-parks at a
// time, so lanes finish at different times and waves interleave their claims.
//
// The walk is TaskWalk (scalar frames, split <= 4) or LaneWalk (frames in lanes, split 5..8), as
// split_kernel picks them; a seventh argument `lanes` runs LaneWalk at every split.
//
// Per-task counts follow split_kernel's PT protocol: a lane adds the counts of its current task's
// roots when it takes a root of another task and at exit, the walk's last-row counts go to their
// task when the walk moves on, edge tasks' counts when they are claimed.
//
// Checks, reported on stdout as one line "total flags roots claims_ok frames_ok terminated
// cursor_claims nodes" (nodes: split_kernel's node count, the walks' placements plus each root's
// expansions less its completed boards):
//   * every task index is claimed exactly once;
//   * every handed-out root needs at most stack_depth frames (N - 1 - its depth);
//   * every wave terminates.
// The total is compared with the oracle by tests/test_walk_model.py.
//
//   walk_model N split chunk waves stack_depth taskfile [lanes|regs [per_task_out]]
//   (taskfile: uint32 triples; per_task_out: the per-task counts as uint64, in task order)
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "nq_walk.h"

namespace {

struct Sub {
    uint64_t count = 0, expansions = 0;
};

// Completions and loop expansions (one per queen placed, the last row included) below a prefix.
void subtree(uint32_t last, uint32_t c, uint32_t l, uint32_t r, Sub *s) {
    if (c == last) {
        s->count++;
        return;
    }
    uint32_t v = last & ~(c | l | r);
    while (v) {
        const uint32_t p = v & (0u - v);
        v ^= p;
        s->expansions++;
        subtree(last, c | p, (l | p) << 1, (r | p) >> 1, s);
    }
}

struct Lane {
    bool has = true;
    bool parked = true;  // v == 0 in the kernel: no root, or the root is exhausted
    uint64_t work = 0;
    uint64_t my_task = 0, acc = 0;  // per-task counts: the kernel's lane accumulator (PT)
};

struct Model {
    uint32_t N, last, chunk, stack_depth;
    const std::vector<uint32_t> *tasks;
    uint64_t ntasks;
    uint64_t cursor = 0, nwaves = 1, claims = 0;
    std::vector<int> claimed;
    uint64_t total = 0, flags = 0, roots = 0;
    uint64_t nodes = 0;  // split_kernel's node count: the walks' placements + each root's own nodes
    bool frames_ok = true;
    std::vector<uint64_t> per_task;  // split_kernel<..., PT>'s per-task counts, as it adds them
    const char *per_task_out = nullptr;
};

template <typename W>
struct Wave {
    static constexpr int S = W::rows;
    Lane lane[64];
    W walk;
    uint64_t next = 0, end = 0;
    bool drained = false, finished = false;
    uint64_t wsum = 0;
    uint64_t cur_t = 0, tdirect = 0;  // PT: the task being walked, its last-row counts not yet added

    // One pass of the kernel's outer loop: refill, then trips until some lane parks.
    void step(Model &m) {
        uint64_t need = 0;
        for (int i = 0; i < 64; ++i)
            if (lane[i].has && lane[i].parked) need |= 1ull << i;
        while (need) {
            uint32_t dc = 0, dl = 0, dr = 0, dv = 0;
            bool got = false;
            while (!drained) {
                uint64_t direct = 0;
                got = walk.next(m.N, m.last, &dc, &dl, &dr, &dv, &direct);
                wsum += direct;
                tdirect += direct;
                if (got) break;
                if (tdirect) {  // the walked task is exhausted: its last-row counts
                    m.per_task[cur_t] += tdirect;
                    tdirect = 0;
                }
                if (next >= end) {
                    // split_kernel's guided claim: 1 / (GUIDED_SHARES = 4 x waves) of the tasks
                    // left at the current cursor, at least `chunk`.
                    const uint64_t left = m.ntasks > m.cursor ? m.ntasks - m.cursor : 0;
                    uint64_t want = left / (4 * m.nwaves);
                    if (want < m.chunk) want = m.chunk;
                    next = m.cursor;
                    m.cursor += want;
                    end = next + want;
                    m.claims++;
                }
                if (next >= m.ntasks) {
                    drained = true;
                    break;
                }
                const uint64_t t = next++;
                m.claimed[t]++;
                const uint32_t c0 = (*m.tasks)[3 * t], l0 = (*m.tasks)[3 * t + 1], r0 = (*m.tasks)[3 * t + 2];
                const uint32_t v0 = m.last & ~(c0 | l0 | r0);
                const uint32_t d0 = __builtin_popcount(c0);
                if (d0 + 1 >= m.N || v0 == 0) {
                    const uint32_t e = d0 >= m.N ? (c0 == m.last ? 1u : 0u) : (d0 + 1 == m.N ? __builtin_popcount(v0) : 0u);
                    wsum += e;
                    m.per_task[t] += e;
                    continue;
                }
                if (d0 + S + 1 < m.N && m.N - 1 - (d0 + S) > m.stack_depth) {
                    m.flags |= 1;
                    continue;
                }
                walk.start(c0, l0, r0, v0, d0);
                cur_t = t;
            }
            if (!got) {
                for (int i = 0; i < 64; ++i)
                    if ((need >> i) & 1) lane[i].has = false;
                break;
            }
            const int sel = __builtin_ctzll(need);
            need &= need - 1;
            // The root dc|dl|dr with candidates dv: its DFS needs N-1-depth frames.
            const uint32_t depth = __builtin_popcount(dc);
            if (m.N - 1 - depth > m.stack_depth) m.frames_ok = false;
            if (dv != (m.last & ~(dc | dl | dr)) || dv == 0) m.frames_ok = false;
            Sub sub;
            subtree(m.last, dc, dl, dr, &sub);
            wsum += sub.count;
            m.nodes += sub.expansions - sub.count;  // the hot loop's expansions less its boards
            lane[sel].parked = false;
            lane[sel].work = sub.expansions;
            if (lane[sel].my_task != cur_t) {  // the lane's roots of my_task are all counted
                m.per_task[lane[sel].my_task] += lane[sel].acc;
                lane[sel].acc = 0;
                lane[sel].my_task = cur_t;
            }
            lane[sel].acc += sub.count;  // folded from the lane's counter while it runs the root
            m.roots++;
        }
        bool live = false;
        for (auto &ln : lane) live |= ln.has;
        if (!live) {
            finished = true;
            m.total += wsum;
            m.nodes += walk.placed;
            for (auto &ln : lane) m.per_task[ln.my_task] += ln.acc;  // the kernel's exit flush
            return;
        }
        // Trips of 16 iterations until some live lane is parked at a trip's end.
        for (;;) {
            bool any_parked = false;
            for (auto &ln : lane) {
                if (!ln.has) continue;
                if (!ln.parked) {
                    ln.work = ln.work > 16 ? ln.work - 16 : 0;
                    if (ln.work == 0) ln.parked = true;
                }
                any_parked |= ln.parked;
            }
            if (any_parked) break;
        }
    }
};

template <typename W>
int run(Model &m, int nwaves) {
    std::vector<Wave<W>> waves(nwaves);
    m.nwaves = static_cast<uint64_t>(nwaves);
    // Waves take turns; a wave's turn count is bounded by its claims plus its roots, so a
    // generous cap detects a wave that never terminates.
    const uint64_t cap = 64 * (m.ntasks + 1) * 1000 + 1000000;
    uint64_t turns = 0;
    bool all = false;
    while (!all && turns < cap) {
        all = true;
        for (auto &w : waves) {
            if (w.finished) continue;
            w.step(m);
            all = false;
            ++turns;
        }
    }
    bool claims_ok = true;
    for (int c : m.claimed) claims_ok &= c == 1;
    printf("%llu %llu %llu %d %d %d %llu %llu\n", (unsigned long long)m.total, (unsigned long long)m.flags,
           (unsigned long long)m.roots, claims_ok ? 1 : 0, m.frames_ok ? 1 : 0, all ? 1 : 0,
           (unsigned long long)m.claims, (unsigned long long)m.nodes);
    if (m.per_task_out) {
        FILE *f = fopen(m.per_task_out, "wb");
        if (f) {
            fwrite(m.per_task.data(), sizeof(uint64_t), m.per_task.size(), f);
            fclose(f);
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 7 || argc > 9) {
        fprintf(stderr, "usage: %s N split chunk waves stack_depth taskfile [lanes|regs [per_task_out]]\n", argv[0]);
        return 2;
    }
    const bool lanes = argc >= 8 && !strcmp(argv[7], "lanes");
    Model m;
    m.N = static_cast<uint32_t>(atoi(argv[1]));
    const int split = atoi(argv[2]);
    m.chunk = static_cast<uint32_t>(atoi(argv[3]));
    const int nwaves = atoi(argv[4]);
    m.stack_depth = static_cast<uint32_t>(atoi(argv[5]));
    m.last = (1u << m.N) - 1u;
    std::vector<uint32_t> tasks;
    FILE *f = fopen(argv[6], "rb");
    if (!f) return 2;
    uint32_t w[3];
    while (fread(w, sizeof w, 1, f) == 1) tasks.insert(tasks.end(), w, w + 3);
    fclose(f);
    m.tasks = &tasks;
    m.ntasks = tasks.size() / 3;
    m.claimed.assign(m.ntasks, 0);
    m.per_task.assign(m.ntasks, 0);
    if (argc == 9) m.per_task_out = argv[8];
    if (lanes) switch (split) {
            case 1: return run<nq::LaneWalk<1>>(m, nwaves);
            case 2: return run<nq::LaneWalk<2>>(m, nwaves);
            case 3: return run<nq::LaneWalk<3>>(m, nwaves);
            case 4: return run<nq::LaneWalk<4>>(m, nwaves);
        }
    switch (split) {
        case 1: return run<nq::WalkFor<1>::type>(m, nwaves);
        case 2: return run<nq::WalkFor<2>::type>(m, nwaves);
        case 3: return run<nq::WalkFor<3>::type>(m, nwaves);
        case 4: return run<nq::WalkFor<4>::type>(m, nwaves);
        case 5: return run<nq::WalkFor<5>::type>(m, nwaves);
        case 6: return run<nq::WalkFor<6>::type>(m, nwaves);
        case 7: return run<nq::WalkFor<7>::type>(m, nwaves);
        case 8: return run<nq::WalkFor<8>::type>(m, nwaves);
    }
    return 2;
}
