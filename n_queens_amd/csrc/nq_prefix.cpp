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
#include <stdint.h>
#include <string.h>

#include <thread>
#include <vector>

#include "nq_internal.h"

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

}  // namespace nq
