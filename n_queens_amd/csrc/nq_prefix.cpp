// Host prefix/task generator.
//
// Produces exactly the task list of the reference's partial_n_queens (src/n_queens.cpp:74-91)
// followed, for odd N, by partial_n_queens_for_odd (src/n_queens.cpp:93-113), in the same
// order (cuda_n_queens calls them in that order, src/n_queens_cuda.cu:91-95):
//   * row 0 of the even part is restricted to columns < N/2;
//   * the odd part puts row 0 in the middle column and restricts row 1 to columns < (N-2)/2;
//   * children are visited lowest column first; a task is (cur, left<<, right>>) after `rows`
//     queens.
// The masks are uint32_t (the reference shifts signed ints; only the low N bits are read).
//
// Implementation: an explicit-stack iterative walk (no recursion).  The subtrees under the
// row-0 choices are generated by separate threads into their own buffers and concatenated
// in order, so the output is identical to the sequential order.
//
// expand_tagged (the serving and batching paths' expansion into 16-byte tasks) splits a walk
// into seed subtrees a few rows down, counts every seed's leaves on all threads, and then has
// each thread write its seeds' leaves straight to their final positions: the same order as one
// sequential walk, with no buffer growth or copy.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "nq_internal.h"
#include "nq_tasks.h"

namespace nq {

namespace {

inline uint32_t full_mask(int N) { return N >= 32 ? 0xFFFFFFFFu : ((1u << N) - 1u); }

// Row restriction of the reference generators, evaluated on the state the row starts from.
inline uint32_t row_mask(int N, bool odd_part, uint32_t cur) {
    if (!odd_part) return cur == 0 ? ((1u << (N / 2)) - 1u) : full_mask(N);
    if (cur == 0) return 1u << (N / 2);
    if ((cur & (cur - 1u)) == 0) return (1u << ((N - 2) / 2)) - 1u;
    return full_mask(N);
}

// Row restriction for expand_tagged's modes.
inline uint32_t mode_mask(int N, int mode, uint32_t cur) {
    return mode == EXPAND_PLAIN ? full_mask(N) : row_mask(N, mode == EXPAND_ODD, cur);
}

// Visit every leaf `rows` rows below (cur, left, right), lowest column first.
template <class Leaf>
void walk_leaves(int N, int mode, uint32_t cur, uint32_t left, uint32_t right, int rows, Leaf &&leaf) {
    if (rows <= 0) {
        leaf(cur, left, right);
        return;
    }
    struct Frame {
        uint32_t c, l, r, v;
    };
    Frame st[33];
    int top = 0;
    st[0] = {cur, left, right, mode_mask(N, mode, cur) & ~(cur | left | right)};
    while (top >= 0) {
        Frame &f = st[top];
        if (f.v == 0) {
            --top;
            continue;
        }
        const uint32_t p = f.v & (0u - f.v);
        f.v ^= p;
        const uint32_t c = f.c | p, l = (f.l | p) << 1, r = (f.r | p) >> 1;
        if (top + 1 == rows) {
            leaf(c, l, r);
            continue;
        }
        ++top;
        st[top] = {c, l, r, mode_mask(N, mode, c) & ~(c | l | r)};
    }
}

// Walk `rows` rows below (cur, left, right) and append every leaf triple to out.
void walk(int N, bool odd_part, uint32_t cur, uint32_t left, uint32_t right, int rows,
          std::vector<uint32_t> &out) {
    if (rows <= 0) return;
    struct Frame {
        uint32_t c, l, r, v;
    };
    Frame st[33];
    int top = 0;
    st[0] = {cur, left, right, row_mask(N, odd_part, cur) & ~(cur | left | right)};
    while (top >= 0) {
        Frame &f = st[top];
        if (f.v == 0) {
            --top;
            continue;
        }
        const uint32_t p = f.v & (0u - f.v);
        f.v ^= p;
        const uint32_t c = f.c | p, l = (f.l | p) << 1, r = (f.r | p) >> 1;
        if (top + 1 == rows) {
            out.push_back(c);
            out.push_back(l);
            out.push_back(r);
            continue;
        }
        ++top;
        st[top] = {c, l, r, row_mask(N, odd_part, c) & ~(c | l | r)};
    }
}

// The first row's candidates from (cur, left, right), each generated in parallel.
void generate_part(int N, bool odd_part, uint32_t cur, uint32_t left, uint32_t right, int rows,
                   std::vector<uint32_t> &out) {
    uint32_t v = row_mask(N, odd_part, cur) & ~(cur | left | right);
    std::vector<uint32_t> firsts;
    while (v) {
        const uint32_t p = v & (0u - v);
        v ^= p;
        firsts.push_back(p);
    }
    if (rows == 1 || firsts.size() <= 1 || rows < 5) {
        walk(N, odd_part, cur, left, right, rows, out);
        return;
    }
    std::vector<std::vector<uint32_t>> parts(firsts.size());
    std::vector<std::thread> th;
    for (size_t i = 0; i < firsts.size(); ++i) {
        th.emplace_back([&, i] {
            const uint32_t p = firsts[i];
            walk(N, odd_part, cur | p, (left | p) << 1, (right | p) >> 1, rows - 1, parts[i]);
        });
    }
    for (auto &t : th) t.join();
    size_t total = out.size();
    for (auto &p : parts) total += p.size();
    out.reserve(total);
    for (auto &p : parts) out.insert(out.end(), p.begin(), p.end());
}

}  // namespace

void generate_from(int N, bool odd_part, uint32_t cur, uint32_t left, uint32_t right, int rows,
                   std::vector<uint32_t> &out) {
    generate_part(N, odd_part, cur, left, right, rows, out);
}

void generate_tasks(int N, int rows, std::vector<uint32_t> &out) {
    out.clear();
    generate_part(N, false, 0, 0, 0, rows, out);
    if (N & 1) generate_part(N, true, 0, 0, 0, rows, out);
}

void expand_tagged(int N, const ExpandPart *parts, int nparts, int rows, uint32_t tag, TaskVec &out) {
    // Seeds: the nodes SEED_ROWS rows down (short walks run on the calling thread).
    constexpr int SEED_ROWS = 3;
    const int threads = static_cast<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    if (rows < SEED_ROWS + 3 || threads == 1) {
        for (int i = 0; i < nparts; ++i)
            walk_leaves(N, parts[i].mode, parts[i].cur, parts[i].left, parts[i].right, rows,
                        [&](uint32_t c, uint32_t l, uint32_t r) { out.push_back(make_uint4(c, l, r, tag)); });
        return;
    }
    std::vector<ExpandPart> seeds;
    for (int i = 0; i < nparts; ++i)
        walk_leaves(N, parts[i].mode, parts[i].cur, parts[i].left, parts[i].right, SEED_ROWS,
                    [&](uint32_t c, uint32_t l, uint32_t r) { seeds.push_back({parts[i].mode, c, l, r}); });
    const int below = rows - SEED_ROWS;
    std::vector<uint64_t> start(seeds.size() + 1, 0);
    auto on_threads = [&](auto &&body) {
        std::atomic<size_t> next{0};
        std::vector<std::thread> th;
        for (int t = 0; t < threads; ++t)
            th.emplace_back([&] {
                for (size_t i = next++; i < seeds.size(); i = next++) body(i);
            });
        for (auto &t : th) t.join();
    };
    on_threads([&](size_t i) {
        uint64_t n = 0;
        const ExpandPart &e = seeds[i];
        walk_leaves(N, e.mode, e.cur, e.left, e.right, below, [&](uint32_t, uint32_t, uint32_t) { ++n; });
        start[i + 1] = n;
    });
    for (size_t i = 0; i < seeds.size(); ++i) start[i + 1] += start[i];
    const size_t at = out.size();
    out.resize(at + start.back());  // uninitialized: every element is written below
    uint4 *dst = out.data() + at;
    on_threads([&](size_t i) {
        uint4 *o = dst + start[i];
        const ExpandPart &e = seeds[i];
        walk_leaves(N, e.mode, e.cur, e.left, e.right, below,
                    [&](uint32_t c, uint32_t l, uint32_t r) { *o++ = make_uint4(c, l, r, tag); });
    });
}

}  // namespace nq
