// 16-byte task lists of the batching and serving paths (host side).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <memory>
#include <utility>
#include <vector>

namespace nq {

// An allocator whose resize() leaves new elements uninitialized (default-initialized), so that
// a list about to be overwritten is not first cleared: for the large lists of whole boards the
// clearing, not the walk, would dominate.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};

using TaskVec = std::vector<uint4, NoInitAlloc<uint4>>;

// One walk of expand_tagged: a start state and the row restrictions to apply (EXPAND_* modes).
struct ExpandPart {
    int mode;
    uint32_t cur, left, right;
};

// Every placement `rows` rows below each part's start, as 16-byte tasks {cur, left, right, tag},
// appended to out part by part, each in the order of one depth-first walk (lowest column
// first).  Large walks are split into seed subtrees a few rows down whose leaves are counted on
// all threads, then written by all threads straight to their final positions (nq_prefix.cpp).
void expand_tagged(int N, const ExpandPart *parts, int nparts, int rows, uint32_t tag, TaskVec &out);

// A request's or batch instance's tasks: its prefix expanded until no task keeps more than
// depth_target free rows, or for an empty board the reference's mirror-symmetric task list
// (tag bit 5 set: weight 2).  Returns the rows placed in every task (nq_batch.cpp).
int append_instance_tasks(int N, int k, uint32_t c, uint32_t l, uint32_t r, int depth_target, uint32_t idx,
                          TaskVec &out);

// Estimated DFS nodes per task of a prefix (board mask in c not set; k queens placed) whose
// tasks lie `split` rows below it: Knuth's random-probe estimate of the node count at every
// depth of the subtree (D. E. Knuth, "Estimating the efficiency of backtrack programs", Math.
// Comp. 29 (1975)), from a fixed number of probes with a generator seeded by the prefix, so the
// same prefix always gets the same estimate.  Orders work only; never affects a count.
double est_nodes_per_task(int N, int k, uint32_t c, uint32_t l, uint32_t r, int split);

}  // namespace nq
