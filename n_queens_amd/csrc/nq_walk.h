// Enumeration of a task's descendants a fixed number of rows below it: the device split that the
// DFS kernel does in registers (split_kernel), with no expanded task list in memory.
//
// A wave claims a whole task and walks the top S levels of its subtree with this state, which
// is the same on every lane of the wave (TaskWalk: scalar registers on the device; LaneWalk: one
// word of frame k in lane k of four vector registers; WalkFor picks one by depth).  Each
// call to next() yields the next node exactly S rows below the task that still has a legal
// next queen, in column order (lowest column first, the order the reference's recursion
// visits them, src/n_queens.cpp:15-20); the wave hands it to one of its lanes as that lane's
// DFS root.  A node met on the way that sits on the last row (one free column left) is not
// handed out: its count, the popcount of its candidates, is returned through *direct, as the
// reference kernel counts the last row (src/n_queens_cuda.cu:59-67).  The counts of all
// handed-out nodes plus *direct sum to the task's count.  `placed` counts the queens the walk
// places (each one a DFS node in the sense of src/n_queens_cuda.cu:39-72: a placement from a
// frame above the last row), so that a node count is `placed` plus the roots' own nodes.
//
// The header is plain C++ so that tests/walk_model.cpp runs the same code on the CPU.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define NQ_HD __host__ __device__ __forceinline__
#else
#define NQ_HD inline
#endif

namespace nq {

NQ_HD uint32_t walk_popc(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __popc(x);
#else
    return static_cast<uint32_t>(__builtin_popcount(x));
#endif
}

template <int S>
struct TaskWalk {
    static_assert(S >= 1 && S <= 4, "1..4 rows are enumerated in registers");
    static constexpr int rows = S;
    // Frame k is the node k rows below the task (frame 0 is the task): its columns, diagonals and
    // the candidates not yet taken.  Frames 0..top are valid.
    uint32_t c[S], l[S], r[S], v[S];
    int top = -1;
    uint32_t d0 = 0;      // rows the task has placed
    uint64_t placed = 0;  // queens placed by the walk so far, over every task (DFS nodes)

    // A task at depth d (d <= N - 2) whose candidate set v0 is not empty.
    NQ_HD void start(uint32_t c0, uint32_t l0, uint32_t r0, uint32_t v0, uint32_t d) {
        c[0] = c0;
        l[0] = l0;
        r[0] = r0;
        v[0] = v0;
        top = 0;
        d0 = d;
    }

    // Next node S rows below the task with candidates (its masks and candidates in *oc..*ov);
    // false once the task is exhausted.  *direct accumulates the last-row counts met on the way.
    NQ_HD bool next(uint32_t N, uint32_t last, uint32_t *oc, uint32_t *ol, uint32_t *orr, uint32_t *ov,
                    uint64_t *direct) {
        for (;;) {
            // Drop exhausted frames (one pass from the deepest: the frames below a dropped one
            // are checked after it).
#pragma unroll
            for (int k = S - 1; k >= 0; --k)
                if (top == k && v[k] == 0) top = k - 1;
            if (top < 0) return false;
            // Lowest candidate of the deepest frame and the child it makes.
            uint32_t fc = 0, fl = 0, fr = 0, p = 0;
#pragma unroll
            for (int k = 0; k < S; ++k)
                if (top == k) {
                    p = v[k] & (0u - v[k]);
                    v[k] ^= p;
                    fc = c[k];
                    fl = l[k];
                    fr = r[k];
                }
            ++placed;
            const uint32_t cc = fc | p, cl = (fl | p) << 1, cr = (fr | p) >> 1;
            const uint32_t cv = last & ~(cc | cl | cr);
            if (cv == 0) continue;
            if (d0 + static_cast<uint32_t>(top) + 2u == N) {  // the child is on the last row
                *direct += walk_popc(cv);
                continue;
            }
            if (top + 1 == S) {
                *oc = cc;
                *ol = cl;
                *orr = cr;
                *ov = cv;
                return true;
            }
#pragma unroll
            for (int k = 1; k < S; ++k)
                if (top + 1 == k) {
                    c[k] = cc;
                    l[k] = cl;
                    r[k] = cr;
                    v[k] = cv;
                }
            ++top;
        }
    }
};

// One 32-bit word of every frame of a deep walk.  On the device frame k's word sits in lane k of
// one register of the wave (x[0]): read with readlane at a wave-uniform index, written by a
// select on the lane id (a compare and a v_cndmask).  Each access costs the same at any depth,
// where TaskWalk's selects over S scalar registers grow with S (and S = 4 already fills the
// scalar register file).  On the host (the CPU model) frame k is x[k].
struct LaneWord {
    uint32_t x[8];
    NQ_HD uint32_t get(int k) const {
#if defined(__HIP_DEVICE_COMPILE__)
        return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x[0]), k));
#else
        return x[k];
#endif
    }
    NQ_HD void set(int k, uint32_t val) {
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        x[0] = lane == static_cast<uint32_t>(k) ? val : x[0];
#else
        x[k] = val;
#endif
    }
};

// TaskWalk with the frames in lanes (LaneWord), used for 3..8 rows: the same nodes in the same
// order, the same *direct counts.  Called by every lane of the wave with the same arguments (the
// state is uniform; only frame k's word lives in lane k).
template <int S>
struct LaneWalk {
    static_assert(S >= 1 && S <= 8, "1..8 rows are enumerated with the frames in lanes");
    static constexpr int rows = S;
    LaneWord c, l, r, v;
    int top = -1;
    uint32_t d0 = 0;
    uint64_t placed = 0;  // queens placed by the walk so far, over every task (DFS nodes)

    NQ_HD void start(uint32_t c0, uint32_t l0, uint32_t r0, uint32_t v0, uint32_t d) {
        c.x[0] = l.x[0] = r.x[0] = v.x[0] = 0;
        c.set(0, c0);
        l.set(0, l0);
        r.set(0, r0);
        v.set(0, v0);
        top = 0;
        d0 = d;
    }

    NQ_HD bool next(uint32_t N, uint32_t last, uint32_t *oc, uint32_t *ol, uint32_t *orr, uint32_t *ov,
                    uint64_t *direct) {
        for (;;) {
            uint32_t fv = 0;
            while (top >= 0 && (fv = v.get(top)) == 0) --top;  // drop exhausted frames
            if (top < 0) return false;
            const uint32_t p = fv & (0u - fv);
            v.set(top, fv ^ p);
            ++placed;
            const uint32_t cc = c.get(top) | p, cl = (l.get(top) | p) << 1, cr = (r.get(top) | p) >> 1;
            const uint32_t cv = last & ~(cc | cl | cr);
            if (cv == 0) continue;
            if (d0 + static_cast<uint32_t>(top) + 2u == N) {  // the child is on the last row
                *direct += walk_popc(cv);
                continue;
            }
            if (top + 1 == S) {
                *oc = cc;
                *ol = cl;
                *orr = cr;
                *ov = cv;
                return true;
            }
            ++top;
            c.set(top, cc);
            l.set(top, cl);
            r.set(top, cr);
            v.set(top, cv);
        }
    }
};

// The walk split_kernel uses for S rows: scalar registers for 1 and 2 rows, frames in lanes from 3
// on (N=21 rows 6: 3 rows 6,297 ms against 6,338 with TaskWalk<3>; N=23 rows 6, 300,000 tasks: 4
// rows 15.42 s against 15.47; 2 rows alike either way).
template <int S, bool DEEP = (S > 2)>
struct WalkFor {
    using type = TaskWalk<S>;
};
template <int S>
struct WalkFor<S, true> {
    using type = LaneWalk<S>;
};

}  // namespace nq
