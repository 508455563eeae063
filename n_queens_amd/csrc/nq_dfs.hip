// DFS counting kernel for gfx950 (CDNA4, wave64).
//
// Work: every task is a prefix (cur, left, right) of queens already placed in the first rows;
// the kernel counts all completions of it.  Reference semantics: src/n_queens.cpp:8-21 (what
// is counted) and src/n_queens_cuda.cu:9-82 (the GPU kernel it replaces).
//
// Design (one lane = one independent DFS, 64 lanes per wave):
//  * The frame being expanded lives in VGPRs: c (occupied columns), l, r (diagonals), v
//    (candidates not yet tried).  One loop iteration places one queen on every active lane,
//    and a second one where the first queen's child has exactly one candidate (the forced
//    move: that child needs no push); there is no per-iteration reload of the top frame.
//  * Before a frame is replaced by its child it is pushed to LDS only if it still has
//    untried candidates (tail-call elimination); a lane whose frame runs empty pops.
//    Frames are 16 B {c, l, r, v} at stack[level][threadIdx.x], behind a dead frame at LDS
//    address 0 that the pop of an empty stack lands on (no stack-empty test; see the loop):
//    lanes of a wave touch 64 consecutive 16-B slots, so ds_write_b128 / ds_read_b128 of real
//    frames are bank-conflict free whatever depth each lane is at.
//  * The columns word c has every bit from N up set: the candidates of a frame are then
//    ~(l | r | c) (one v_bitop3, no board mask) and a completed board is c == ~0.  It is
//    counted with a compare and either an add-with-carry into a per-lane counter (per-task
//    counts, node counts, batches) or a scalar popcount of the compare mask into a per-wave
//    counter (plain totals: one VALU fewer per node).  Nothing in the iteration depends on N,
//    so lanes of one wave may work on different board sizes (batch_kernel).
//  * Two push forms: exec-masked (only frames with candidates left are written), or an
//    unconditional write with sp advanced by stride * min(v, 1) (no exec switch, one spare
//    stack level); the host takes the second when the spare level costs no resident
//    workgroup (prefer_always_push).
//  * A lane whose task is exhausted waits on the dead frame until the end of its 16-iteration
//    trip, then claims the next task (lane-granular refill) from a global cursor; claims are
//    aggregated per wave and handed out in chunks through a wave-private LDS range.
//  * Per-lane counts are 32-bit (a lane adds at most one board and two expansions per
//    iteration) and folded into 64-bit at most every 2^31 increments and at task end; the
//    64-bit sums are reduced per wave with DPP shuffles, per workgroup through LDS, and added
//    with one atomic per workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "nq_kernels.h"
#include "nq_walk.h"

namespace nq {

namespace {

// Hot-loop trips (16 iterations each) between folds of the 32-bit per-lane counters: a lane
// adds at most 2 per iteration (the node counter: a placement and a forced move), so at most
// 2 * 16 * 2^26 = 2^31 between folds.
constexpr uint32_t FLUSH_EVERY = 1u << 26;
// Trips between folds of the wave-uniform 32-bit counter (64 lanes add to it).
constexpr uint32_t FLUSH_EVERY_WAVE = 1u << 21;
// split_kernel's guided claims: a wave takes 1 / (GUIDED_SHARES x waves) of the tasks left.
constexpr uint64_t GUIDED_SHARES = 4;

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t first_lane_u64(uint64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Claim one task index for every active lane of the wave.  wq[0..1] is the wave's claimed but
// unused range [next, end) in LDS (kept in memory, not registers, because lanes leave the
// claiming loop at different times and must all see the same range afterwards).  Lane 0 of the
// claiming lanes writes it and every lane reads it, so the accesses are volatile: the compiler
// must not keep a lane's earlier read in a register (that handed out task indices twice).  LDS
// operations of one wave complete in order, so the next claim sees the last write.
__device__ __forceinline__ uint64_t claim(unsigned long long *wq_, unsigned long long *cursor,
                                          uint32_t chunk) {
    // The explicit LDS address space keeps these ds_read/ds_write (a plain volatile pointer
    // would become flat accesses).
    volatile __attribute__((address_space(3))) unsigned long long *wq =
        (volatile __attribute__((address_space(3))) unsigned long long *)wq_;
    const uint64_t mask = __ballot(1);
    const uint32_t k = __popcll(mask);
    const uint32_t rank = lane_rank(mask);
    const uint64_t next = wq[0];
    const uint64_t end = wq[1];
    const uint64_t avail = end - next;
    uint64_t mine;
    if (avail >= k) {
        mine = next + rank;
        if (rank == 0) wq[0] = next + k;
    } else {
        const uint64_t want = (k - avail) > chunk ? (k - avail) : chunk;
        uint64_t g = 0;
        if (rank == 0) g = atomicAdd(cursor, static_cast<unsigned long long>(want));
        g = first_lane_u64(g);
        mine = rank < avail ? next + rank : g + (rank - avail);
        if (rank == 0) {
            wq[0] = g + (k - avail);
            wq[1] = g + want;
        }
    }
    return mine;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// One DFS iteration for every live lane of the wave, hand-scheduled (gfx950).  The frame
// {c, l, r, v} lives in v[20:23] (bound by the asm constraints below) so that it is both the
// ds_write_b128 source and the ds_read_b128 destination without register moves.  c holds the
// occupied columns with every bit from N up set, so the candidates are ~(l | r | c) with no
// board mask and a full board is c == ~0; the iteration does not depend on N.
//
//   p = v & -v; v ^= p                          lowest candidate, v = candidates left
//   push {c,l,r,v} if v != 0                    (masked, or unconditional: see below)
//   c |= p; l = (l|p) << 1; r = (r|p) >> 1      (l & p == 0: one v_add_lshl_u32)
//   v = ~(l | r | c)                            (v_bitop3 table 0x01)
//   if (popcount(v) == 1) {                     forced move: the child's only candidate is
//       c |= v; l = (l|v) << 1; r = (r|v) >> 1  placed at once; that node keeps nothing to
//       v = ~(l | r | c) }                      come back to, so it needs no push
//   if (c == ~0) cnt += 1                       (compare + add with carry, or a scalar
//                                                popcount of the compare mask per wave)
//   if (v == 0) { sp = max(sp - stride, 0); {c,l,r,v} = lds[sp] }
//
// The forced move places a second queen in 37 % of iterations (the share of children with
// exactly one candidate) for seven VALU and no LDS access, so one push and one pop serve 1.37
// placements on average: the 16-byte push write bounds the plain iteration, and the VALU has
// room.  A lane completes at most one board per iteration (a completed board has no candidate,
// so no forced move follows it), so one count after the forced move counts both placements.
//
// The pop needs no stack-empty test.  LDS address 0 holds a dead frame DEAD_FRAME and every
// lane's level 0 lies less than one stride above it, so the saturating subtract (v_sub_u32
// with clamp) sends the pop of an empty stack to address 0.  A lane on the dead frame places
// nothing (p = 0), its child is dead again (~(c | l<<1 | r>>1) == 0, c != ~0: nothing is
// counted) and it pops the dead frame again, until the end of the trip (16 iterations),
// where lanes with sp == 0 are found and the asm statement returns for their refill.  Done
// lanes thus idle until the end of the trip, as they would out of exec.
//
// The push and the pop take their lanes with v_cmpx (the compare writes exec), and exec goes
// back to the live mask with one s_mov: exec outside those two regions is always `live`, so
// there is no VALU -> SGPR -> SALU mask round trip on the chain.  No DPP, readlane or VMEM
// follows an exec write, and every LDS read is waited for before the iteration ends, so no
// hazard wait states are needed.
#define NQ_EXTRACT                                                                            \
    "v_sub_u32 %[t], 0, v23\n\t"                                                              \
    "v_and_b32 %[p], v23, %[t]\n\t"                                                           \
    "v_xor_b32 v23, v23, %[p]\n\t"

// Push only the lanes whose frame keeps untried candidates (exec-masked by v_cmpx).
#define NQ_PUSH_MASKED                                                                        \
    "v_cmpx_ne_u32_e64 %[m1], 0, v23\n\t"                                                     \
    "ds_write_b128 %[sp], v[20:23]\n\t"                                                       \
    "v_add_u32 %[sp], %[stride], %[sp]\n\t"                                                   \
    "s_mov_b64 exec, %[live]\n\t"

// Write every lane's frame and keep it (advance sp by stride * min(v, 1)) only where
// candidates remain: no exec switch.  A lane whose every level is in use writes one level
// above its top frame, so this form needs one spare level.  A lane on the dead frame writes
// the dead frame back to address 0 unchanged (v == 0 after a placement of p = 0).
#define NQ_PUSH_ALWAYS                                                                        \
    "ds_write_b128 %[sp], v[20:23]\n\t"                                                       \
    "v_min_u32 %[t], 1, v23\n\t"                                                              \
    "v_mad_u32_u24 %[sp], %[t], %[stride], %[sp]\n\t"

// The child's right diagonal is started first so that no VALU waits on the one before it.
#define NQ_CHILD                                                                              \
    "v_or_b32 v22, v22, %[p]\n\t"                                                             \
    "v_or_b32 v20, v20, %[p]\n\t"                                                             \
    "v_lshrrev_b32 v22, 1, v22\n\t"                                                           \
    "v_add_lshl_u32 v21, v21, %[p], 1\n\t"                                                    \
    "v_bitop3_b32 v23, v21, v22, v20 bitop3:1\n\t"

// The forced move, for the lanes whose child has exactly one candidate (v_cmpx on its
// popcount): the candidate is v itself.  NODES_FORCED counts the extra expansion.
#define NQ_FORCED(NODES_FORCED)                                                               \
    "v_bcnt_u32_b32 %[t], v23, 0\n\t"                                                         \
    "v_cmpx_eq_u32_e64 %[m1], 1, %[t]\n\t" NODES_FORCED                                       \
    "v_or_b32 v22, v22, v23\n\t"                                                              \
    "v_or_b32 v20, v20, v23\n\t"                                                              \
    "v_lshrrev_b32 v22, 1, v22\n\t"                                                           \
    "v_add_lshl_u32 v21, v21, v23, 1\n\t"                                                     \
    "v_bitop3_b32 v23, v21, v22, v20 bitop3:1\n\t"                                            \
    "s_mov_b64 exec, %[live]\n\t"

// Count, then pop.  A completed board has no candidates, so its lane pops: LANE_CNT adds the
// compare (vcc) to the lane's counter inside the pop's exec region.  WAVE_CNT adds the
// popcount of vcc to a wave-uniform SGPR counter in the shadow of the pop's LDS read.
#define NQ_COUNT_POP(LANE_CNT, WAVE_CNT)                                                      \
    "v_cmp_eq_u32 vcc, -1, v20\n\t"                                                           \
    "v_cmpx_eq_u32_e64 %[m1], 0, v23\n\t" LANE_CNT                                            \
    "v_sub_u32_e64 %[sp], %[sp], %[stride] clamp\n\t"                                         \
    "ds_read_b128 v[20:23], %[sp]\n\t"                                                        \
    "s_mov_b64 exec, %[live]\n\t" WAVE_CNT                                                    \
    "s_waitcnt lgkmcnt(0)\n\t"

#define NQ_CNT_LANE "v_addc_co_u32 %[cnt], vcc, 0, %[cnt], vcc\n\t"
#define NQ_CNT_WAVE                                                                           \
    "s_bcnt1_i32_b64 %[bc], vcc\n\t"                                                          \
    "s_add_u32 %[wcnt], %[wcnt], %[bc]\n\t"

#define NQ_ITER(PUSH, NODES_INSN, NODES_FORCED, LANE_CNT, WAVE_CNT)                           \
    NODES_INSN NQ_EXTRACT PUSH NQ_CHILD NQ_FORCED(NODES_FORCED) NQ_COUNT_POP(LANE_CNT, WAVE_CNT)
#define NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT) \
    NQ_ITER(PUSH, NI, NF, LANE_CNT, WAVE_CNT) NQ_ITER(PUSH, NI, NF, LANE_CNT, WAVE_CNT)

// The whole hot loop: 16 iterations per trip, leaving after a trip that ends with some lane
// on the dead frame (sp == 0), or after [iters] trips (the caller folds the 32-bit counters
// then).  [nd] counts expansions (node-counting variant only).
#define NQ_HOT_ASM(PUSH, NI, NF, LANE_CNT, WAVE_CNT)                                      \
    __asm__ volatile(                                                                         \
        "s_mov_b64 %[save], exec\n\t"                                                         \
        "s_mov_b64 exec, %[live]\n"                                                           \
        "nq_hot_%=:\n\t" NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT)                        \
        NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT) NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT) \
        NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT) NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT) \
        NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT) NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT) \
        NQ_ITER2(PUSH, NI, NF, LANE_CNT, WAVE_CNT)                                         \
        "v_cmp_eq_u32 %[m1], 0, %[sp]\n\t"                                                    \
        "s_cmp_lg_u64 %[m1], 0\n\t"                                                           \
        "s_cbranch_scc1 nq_exit_%=\n\t"                                                       \
        "s_sub_u32 %[iters], %[iters], 1\n\t"                                                 \
        "s_cmp_lg_u32 %[iters], 0\n\t"                                                        \
        "s_cbranch_scc1 nq_hot_%=\n"                                                          \
        "nq_exit_%=:\n\t"                                                                     \
        "s_mov_b64 exec, %[save]\n\t"                                                         \
        : "+{v20}"(c), "+{v21}"(l), "+{v22}"(r), "+{v23}"(v), [sp] "+v"(sp), [cnt] "+v"(cnt),   \
          [nd] "+v"(nd), [iters] "+s"(iters), [wcnt] "+s"(wcnt), [t] "=&v"(t), [p] "=&v"(p),   \
          [m1] "=&s"(m1), [save] "=&s"(save), [bc] "=&s"(bc)                                   \
        : [live] "s"(live), [stride] "s"(stride)                                              \
        : "memory", "scc", "vcc")

// Expansions of the lanes not on the dead frame (sp != 0 at the start of the iteration), and
// the forced move's expansion (inside its exec region).
#define NQ_NODE_INC                                                                           \
    "v_min_u32 %[t], 1, %[sp]\n\t"                                                            \
    "v_add_u32 %[nd], %[t], %[nd]\n\t"
#define NQ_NODE_FORCED "v_add_u32 %[nd], 1, %[nd]\n\t"

// Runs the hot loop for the lanes in `live` until one of them finishes its task (or the fold
// interval ends).  APUSH selects the unconditional push (the stack has a spare level).  WAVE
// counts completed boards per wave into wcnt (SALU) instead of per lane into cnt; its fold
// interval is 2^21 trips, at most 2^21 * 16 * 64 = 2^31 boards (one per lane and iteration).
template <bool NODES, bool APUSH, bool WAVE>
__device__ __forceinline__ void hot_loop(uint32_t &c, uint32_t &l, uint32_t &r, uint32_t &v, uint32_t &sp,
                                         uint32_t &cnt, uint32_t &wcnt, uint32_t &nd, uint64_t live,
                                         uint32_t stride, uint32_t max_trips = 0) {
    // At most the fold interval of the counters (a smaller bound returns sooner).
    const uint32_t fold = WAVE ? FLUSH_EVERY_WAVE : FLUSH_EVERY;
    uint32_t t, p, bc, iters = (max_trips && max_trips < fold) ? max_trips : fold;
    uint64_t m1, save;
    if (WAVE) {
        static_assert(!WAVE || !NODES, "the node-counting variant counts per lane");
        if (APUSH)
            NQ_HOT_ASM(NQ_PUSH_ALWAYS, "", "", "", NQ_CNT_WAVE);
        else
            NQ_HOT_ASM(NQ_PUSH_MASKED, "", "", "", NQ_CNT_WAVE);
    } else if (APUSH) {
        if (NODES)
            NQ_HOT_ASM(NQ_PUSH_ALWAYS, NQ_NODE_INC, NQ_NODE_FORCED, NQ_CNT_LANE, "");
        else
            NQ_HOT_ASM(NQ_PUSH_ALWAYS, "", "", NQ_CNT_LANE, "");
    } else {
        if (NODES)
            NQ_HOT_ASM(NQ_PUSH_MASKED, NQ_NODE_INC, NQ_NODE_FORCED, NQ_CNT_LANE, "");
        else
            NQ_HOT_ASM(NQ_PUSH_MASKED, "", "", NQ_CNT_LANE, "");
    }
}

// Dynamic LDS of dfs_kernel and batch_kernel (the kernels have no static LDS, so it starts at
// address 0): the dead frame, the frame stack [levels][blockDim.x], then per wave the claim
// range wq[2] and the reduction slot red[2].
struct LdsLayout {
    uint4 *stack;
    unsigned long long (*wq)[2];
    unsigned long long (*red)[2];
};

__device__ __forceinline__ LdsLayout lds_layout(uint4 *lds, uint32_t levels) {
    LdsLayout L;
    L.stack = lds + 1;
    auto *tail = reinterpret_cast<unsigned long long(*)[2]>(lds + 1 + levels * blockDim.x);
    L.wq = tail;
    L.red = tail + MAX_WAVES_PER_BLOCK;
    return L;
}

// {c, l, r, v} of the dead frame: p = 0 leaves it as is, its child ~(c | l<<1 | r>>1) is 0 and
// c != ~0.  Written by lane 0 of every wave before the wave's first pop (identical values).
__device__ __forceinline__ void put_dead_frame(uint4 *lds) {
    if ((threadIdx.x & 63u) == 0) lds[0] = make_uint4(0xFFFFFFFEu, 0u, 0xFFFFFFFFu, 0u);
}

// The dead frame must sit at LDS address 0 (the target of the saturating pop).  Nonzero only if
// the kernel were given static LDS; then it flags DEV_ERR_LDS and does nothing.
__device__ __forceinline__ bool lds_base_ok(const uint4 *lds, DevState *st) {
    if (static_cast<uint32_t>(reinterpret_cast<uintptr_t>(lds)) == 0u) return true;
    if (threadIdx.x == 0) atomicOr(&st->error, DEV_ERR_LDS);
    return false;
}

template <bool NODES, bool PER_TASK, bool APUSH>
__global__ void __launch_bounds__(MAX_BLOCK)
dfs_kernel(const uint32_t *__restrict__ tasks, uint64_t ntasks, uint32_t N, uint32_t chunk,
           uint32_t stack_depth, DevState *__restrict__ st, uint64_t *__restrict__ per_task,
           uint64_t *__restrict__ trace, const unsigned long long *__restrict__ ntasks_dev) {
    // Task count produced on the device (expansion), never beyond the buffer bound `ntasks`.
    if (ntasks_dev) ntasks = min(ntasks, static_cast<uint64_t>(*ntasks_dev));
    extern __shared__ uint4 lds[];  // lds_layout(): dead frame, stack, wq, red
    if (!lds_base_ok(lds, st)) return;
    const LdsLayout L = lds_layout(lds, stack_depth + (APUSH ? 1u : 0u));
    auto *wq = L.wq;
    auto *red = L.red;

    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t t_start = trace ? wall_clock64() : 0;
    if ((threadIdx.x & 63u) == 0) {
        wq[wave][0] = 0;
        wq[wave][1] = 0;
    }
    put_dead_frame(lds);
    const uint32_t last = (1u << N) - 1u;
    const uint32_t stride = blockDim.x * static_cast<uint32_t>(sizeof(uint4));
    // Level 0 of this lane; a pop from it saturates to the dead frame at address 0.
    const uint32_t base = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(L.stack + threadIdx.x));
    uint32_t sp = base;

    uint32_t c = 0, l = 0, r = 0, v = 0;  // the frame being expanded
    // Per-task counts and node counts need a counter per lane; a plain total is counted per
    // wave (SALU popcount of the completed-board mask, off the VALU).
    constexpr bool WAVE = !NODES && !PER_TASK;
    uint32_t cnt = 0;     // completed boards since the last fold (at most one per iteration)
    uint32_t wcnt = 0;    // the same for the whole wave (WAVE variant), wave-uniform
    uint64_t wsum = 0;    // folded wcnt, wave-uniform
    uint32_t nd = 0;      // expansions since the last fold (node-counting variant)
    uint64_t tsum = 0;    // folded count of the current task
    uint64_t acc = 0;     // counts of finished tasks
    uint64_t nodes = 0;   // expansions minus completed boards, i.e. DFS nodes
    uint64_t tid = ~0ull;
    bool has = true;      // lane still has (or may claim) work

    for (;;) {
        // Lanes whose task is exhausted (parked on the dead frame: v == 0) claim the next one.
        if (has && v == 0) {
            const uint64_t done = tsum;
            acc += done;
            if (PER_TASK && tid != ~0ull) per_task[tid] = done;
            tsum = 0;
            has = false;
            for (;;) {
                const uint64_t t = claim(wq[wave], &st->cursor, chunk);
                if (t >= ntasks) break;
                const uint32_t c0 = tasks[3 * t + 0];
                const uint32_t l0 = tasks[3 * t + 1];
                const uint32_t r0 = tasks[3 * t + 2];
                const uint32_t v0 = last & ~(c0 | l0 | r0);
                const uint32_t d0 = __popc(c0);
                if (d0 + 1u >= N || v0 == 0) {
                    // Prefix already at the last row (or beyond), or no legal next queen:
                    // its count is immediate (src/n_queens.cpp:10-13 counts a full board).
                    const uint64_t direct =
                        d0 >= N ? (c0 == last ? 1u : 0u) : (d0 + 1u == N ? __popc(v0) : 0u);
                    acc += direct;
                    if (PER_TASK) per_task[t] = direct;
                    continue;
                }
                if (N - 1u - d0 > stack_depth) {
                    // Not enough frames for this prefix: flag and skip (host checks first).
                    atomicOr(&st->error, DEV_ERR_DEPTH);
                    if (PER_TASK) per_task[t] = 0;
                    continue;
                }
                c = c0 | ~last;
                l = l0;
                r = r0;
                v = v0;
                sp = base;
                tid = t;
                has = true;
                break;
            }
        }
        const uint64_t live = __ballot(has);
        if (live == 0) break;
        // Hot loop until a lane needs a new task.
        hot_loop<NODES, APUSH, WAVE>(c, l, r, v, sp, cnt, wcnt, nd, live, stride);
        tsum += cnt;
        if (NODES) nodes += static_cast<uint64_t>(nd) - cnt;
        wsum += wcnt;
        cnt = 0;
        wcnt = 0;
        nd = 0;
    }

    if (trace && (threadIdx.x & 63u) == 0) {
        // Per-wave residency (constant-rate wall clock) for the tail-idle statistic.
        const uint64_t w = static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + wave;
        trace[2 * w] = t_start;
        trace[2 * w + 1] = wall_clock64();
    }
    // Reduce: wave (shuffles), workgroup (LDS), device (one atomic per workgroup).
    acc = wave_sum_u64(acc) + wsum;
    if (NODES) nodes = wave_sum_u64(nodes);
    if ((threadIdx.x & 63u) == 0) {
        red[wave][0] = acc;
        red[wave][1] = nodes;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0, n = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
            s += red[w][0];
            n += red[w][1];
        }
        atomicAdd(&st->total, s);
        if (NODES) atomicAdd(&st->nodes, n);
    }
}

// The first active lane of the wave (all lanes are active wherever this is called).
__device__ __forceinline__ uint32_t wave_leader() {
    return static_cast<uint32_t>(__builtin_ctzll(__ballot(1)));
}

// Plain totals with the device split done in registers: each wave claims whole tasks from the
// cursor (guided: a share of the remaining tasks, at least `chunk`) and walks the top S rows of a
// task's subtree (the same state on every lane: TaskWalk in scalar registers for 1-2 rows,
// LaneWalk with frame k in lane k beyond); every node S rows below the task becomes one lane's
// DFS root when that lane's previous root is exhausted.  The expanded list never exists in
// memory, the launch needs no host synchronisation, and each task is read once.  A task's
// subtree stays on the wave that claimed it (spread over its 64 lanes).  The DFS is
// dfs_kernel's hot loop with the per-wave board counter.
//
// MODE 0: totals.  MODE 1 (PT): per-task counts as well (the reference kernel's partial_sum,
// src/n_queens_cuda.cu:79), into per_task[], zeroed before the launch.  The hot loop then counts
// per lane; a lane keeps the task of its current root and the completions of that task's roots it
// has run, and adds them to per_task[] with one atomic when it takes a root of another task (a
// wave walks its tasks one after another, so a lane's roots come in task order) and at exit.  The
// walk's last-row counts go to their task when the wave moves on to the next task, edge tasks'
// counts when they are claimed.  MODE 2 (NODES): DFS node counts as well: the walk's placements
// plus, per lane, the hot loop's expansions less its completed boards (dfs_kernel's count).
template <bool APUSH, int S, int MODE>
__global__ void __launch_bounds__(MAX_BLOCK)
split_kernel(const uint32_t *__restrict__ tasks, uint64_t ntasks, uint32_t N, uint32_t chunk, uint32_t stack_depth,
             DevState *__restrict__ st, uint64_t *__restrict__ per_task, uint64_t *__restrict__ trace,
             const unsigned long long *__restrict__ ntasks_dev) {
    constexpr bool PT = MODE == 1, NODES = MODE == 2, LANE_COUNT = MODE != 0;
    if (ntasks_dev) ntasks = min(ntasks, static_cast<uint64_t>(*ntasks_dev));
    extern __shared__ uint4 lds[];  // lds_layout(): dead frame, stack, wq (unused), red
    if (!lds_base_ok(lds, st)) return;
    const LdsLayout L = lds_layout(lds, stack_depth + (APUSH ? 1u : 0u));
    auto *red = L.red;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t t_start = trace ? wall_clock64() : 0;
    put_dead_frame(lds);
    const uint32_t last = (1u << N) - 1u;
    const uint32_t stride = blockDim.x * static_cast<uint32_t>(sizeof(uint4));
    const uint32_t base = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(L.stack + threadIdx.x));
    uint32_t sp = base;
    uint32_t c = 0, l = 0, r = 0, v = 0;  // the frame being expanded
    uint32_t cnt = 0, nd = 0;             // per-lane counters of hot_loop (unused: per-wave count)
    uint32_t wcnt = 0;                    // completed boards of the wave since the last fold
    uint64_t wsum = 0;                    // the wave's count: folded wcnt + last-row and edge counts
    bool has = true;                      // lane still has (or may get) a DFS root
    // PT, per lane: the task of the current root, the completions of that task's roots not yet
    // added to per_task[], and (PT and NODES) all completions the lane counted, for the total;
    // NODES, per lane: the DFS nodes of its roots.
    uint64_t my_task = 0, acc = 0, lsum = 0, nodes = 0;
    // PT, wave-uniform: the task being walked and its last-row counts not yet added.
    uint64_t cur_t = 0, tdirect = 0;

    // Wave-uniform state: the walk over the current task and the wave's claimed task range.
    typename nq::WalkFor<S>::type walk;
    uint64_t next = 0, end = 0;
    bool drained = false;  // no task left to claim
    const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (blockDim.x >> 6);

    for (;;) {
        // Lanes parked on the dead frame take the next roots, one each, lowest lane first.
        uint64_t need = __ballot(has && v == 0);
        while (need) {
            uint32_t dc = 0, dl = 0, dr = 0, dv = 0;
            bool got = false;
            while (!drained) {
                uint64_t direct = 0;
                got = walk.next(N, last, &dc, &dl, &dr, &dv, &direct);
                wsum += direct;
                if (PT) tdirect += direct;
                if (got) break;
                if (PT && tdirect) {  // the walked task is exhausted: its last-row counts
                    if (lane == wave_leader()) atomicAdd(reinterpret_cast<unsigned long long *>(per_task + cur_t),
                                                         static_cast<unsigned long long>(tdirect));
                    tdirect = 0;
                }
                if (next >= end) {
                    // Guided claims: a share of the tasks left, 1 / (GUIDED_SHARES x waves) of them,
                    // never less than `chunk`, so single tasks near the end.  The cursor is read
                    // first (a wave that ran long would otherwise size its share from an old
                    // reading).  Each cursor access is a device-scope atomic on one address (about
                    // 100 B of memory traffic: the XCDs' L2s are not coherent), so there are few.
                    const uint32_t leader = wave_leader();
                    uint64_t g = 0;
                    if (lane == leader) {
                        const uint64_t cur = __hip_atomic_load(&st->cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const uint64_t left = ntasks > cur ? ntasks - cur : 0;
                        const uint64_t want = max(static_cast<uint64_t>(chunk), left / (GUIDED_SHARES * nwaves));
                        g = atomicAdd(&st->cursor, static_cast<unsigned long long>(want));
                        end = g + want;  // the leader's copy; made uniform below
                    }
                    g = first_lane_u64(g);
                    end = first_lane_u64(end);
                    next = g;
                }
                if (next >= ntasks) {
                    drained = true;
                    break;
                }
                const uint64_t t = next++;
                const uint32_t c0 = __builtin_amdgcn_readfirstlane(tasks[3 * t + 0]);
                const uint32_t l0 = __builtin_amdgcn_readfirstlane(tasks[3 * t + 1]);
                const uint32_t r0 = __builtin_amdgcn_readfirstlane(tasks[3 * t + 2]);
                const uint32_t v0 = last & ~(c0 | l0 | r0);
                const uint32_t d0 = __popc(c0);
                if (d0 + 1u >= N || v0 == 0) {
                    // Prefix at the last row (or full), or no legal next queen: counted here
                    // (src/n_queens.cpp:10-13 counts a full board).
                    const uint32_t e = d0 >= N ? (c0 == last ? 1u : 0u) : (d0 + 1u == N ? __popc(v0) : 0u);
                    wsum += e;
                    if (PT && e && lane == wave_leader())
                        atomicAdd(reinterpret_cast<unsigned long long *>(per_task + t), static_cast<unsigned long long>(e));
                    continue;
                }
                if (d0 + S + 1u < N && N - 1u - (d0 + S) > stack_depth) {
                    // Its roots would need more frames than configured: flag and skip.
                    if (lane == wave_leader()) atomicOr(&st->error, DEV_ERR_DEPTH);
                    continue;
                }
                walk.start(c0, l0, r0, v0, d0);
                cur_t = t;
            }
            if (!got) {
                if ((need >> lane) & 1ull) has = false;
                break;
            }
            const uint32_t sel = static_cast<uint32_t>(__builtin_ctzll(need));
            need &= need - 1;
            if (lane == sel) {
                c = dc | ~last;
                l = dl;
                r = dr;
                v = dv;
                sp = base;
                if (PT && my_task != cur_t) {  // the lane's roots of my_task are all counted
                    if (acc) atomicAdd(reinterpret_cast<unsigned long long *>(per_task + my_task),
                                       static_cast<unsigned long long>(acc));
                    acc = 0;
                    my_task = cur_t;
                }
            }
        }
        const uint64_t live = __ballot(has);
        if (live == 0) break;
        if (LANE_COUNT) {
            uint32_t unused = 0;  // the per-wave counter of the other variant
            hot_loop<NODES, APUSH, false>(c, l, r, v, sp, cnt, unused, nd, live, stride);
            if (PT) acc += cnt;
            if (NODES) nodes += static_cast<uint64_t>(nd) - cnt;
            lsum += cnt;
            cnt = 0;
            nd = 0;
        } else {
            hot_loop<false, APUSH, true>(c, l, r, v, sp, cnt, wcnt, nd, live, stride);
            wsum += wcnt;
            wcnt = 0;
        }
    }
    if (PT && acc) atomicAdd(reinterpret_cast<unsigned long long *>(per_task + my_task), static_cast<unsigned long long>(acc));
    if (LANE_COUNT) wsum += wave_sum_u64(lsum);
    uint64_t wnodes = 0;
    if (NODES) wnodes = wave_sum_u64(nodes) + walk.placed;

    if (trace && lane == 0) {
        const uint64_t w = static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + wave;
        trace[2 * w] = t_start;
        trace[2 * w + 1] = wall_clock64();
    }
    if (lane == 0) {
        red[wave][0] = wsum;
        red[wave][1] = wnodes;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0, n = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
            s += red[w][0];
            n += red[w][1];
        }
        atomicAdd(&st->total, s);
        if (NODES) atomicAdd(&st->nodes, n);
    }
}

// Batched instances: tasks are {cur, left, right, inst << 6 | twice << 5 | N}; N varies per
// lane (the iteration does not depend on it), and every finished task's count (doubled when
// `twice` is set: a mirror-symmetric half of an empty board) goes to counts[inst].  Tasks of one
// instance are contiguous in the list and a lane claims increasing indices, so each lane keeps
// a running (instance, sum) and adds it with one 64-bit atomic only when it moves on to another
// instance, and at exit; a per-task atomic made the counters the bottleneck (about 9e7
// atomics/s) once tasks are small.  Same iteration as dfs_kernel.
template <bool APUSH>
__global__ void __launch_bounds__(MAX_BLOCK)
batch_kernel(const uint4 *__restrict__ tasks, uint64_t ntasks, uint32_t chunk, uint32_t stack_depth,
             DevState *__restrict__ st, unsigned long long *__restrict__ counts) {
    extern __shared__ uint4 lds[];  // lds_layout(): dead frame, stack, wq, red
    if (!lds_base_ok(lds, st)) return;
    const LdsLayout L = lds_layout(lds, stack_depth + (APUSH ? 1u : 0u));
    auto *wq = L.wq;
    auto *red = L.red;

    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        wq[wave][0] = 0;
        wq[wave][1] = 0;
    }
    put_dead_frame(lds);
    const uint32_t stride = blockDim.x * static_cast<uint32_t>(sizeof(uint4));
    const uint32_t base = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(L.stack + threadIdx.x));
    uint32_t sp = base;
    uint32_t c = 0, l = 0, r = 0, v = 0, cnt = 0, nd = 0;
    uint32_t inst = 0, weight = 1;
    uint64_t tsum = 0, acc = 0;
    bool has = true, own = false;
    uint32_t run_inst = 0;  // instance of the pending run_sum
    uint64_t run_sum = 0;
    auto add = [&](uint32_t i, uint64_t x) {
        if (i != run_inst) {
            if (run_sum) atomicAdd(&counts[run_inst], static_cast<unsigned long long>(run_sum));
            run_inst = i;
            run_sum = 0;
        }
        run_sum += x;
        acc += x;
    };

    // The lane's next task, claimed and loaded one task ahead: its load completes while the
    // current task runs, so a refill does not wait for memory (tasks average about 1e4 nodes,
    // and every refill stalls the whole wave).
    uint4 ahead = make_uint4(0u, 0u, 0u, 0u);
    bool have_ahead = false;

    for (;;) {
        if (has && v == 0) {
            if (own) add(inst, tsum * weight);
            tsum = 0;
            own = false;
            has = false;
            for (;;) {
                uint4 w;
                if (have_ahead) {
                    w = ahead;
                } else {
                    const uint64_t t = claim(wq[wave], &st->cursor, chunk);
                    if (t >= ntasks) break;
                    w = tasks[t];
                }
                const uint64_t t_next = claim(wq[wave], &st->cursor, chunk);
                have_ahead = t_next < ntasks;
                if (have_ahead) ahead = tasks[t_next];
                const uint32_t N = w.w & 31u;
                weight = 1u + ((w.w >> 5) & 1u);
                inst = w.w >> 6;
                const uint32_t last = (1u << N) - 1u;
                const uint32_t v0 = last & ~(w.x | w.y | w.z);
                const uint32_t d0 = __popc(w.x);
                if (d0 + 1u >= N || v0 == 0) {
                    const uint64_t direct =
                        d0 >= N ? (w.x == last ? 1u : 0u) : (d0 + 1u == N ? __popc(v0) : 0u);
                    add(inst, direct * weight);
                    continue;
                }
                if (N - 1u - d0 > stack_depth) {
                    atomicOr(&st->error, DEV_ERR_DEPTH);
                    continue;
                }
                c = w.x | ~last;
                l = w.y;
                r = w.z;
                v = v0;
                sp = base;
                has = true;
                own = true;
                break;
            }
        }
        const uint64_t live = __ballot(has);
        if (live == 0) break;
        uint32_t wcnt = 0;  // unused: a batch counts per lane (lanes work on different instances)
        hot_loop<false, APUSH, false>(c, l, r, v, sp, cnt, wcnt, nd, live, stride);
        tsum += cnt;
        cnt = 0;
    }
    if (run_sum) atomicAdd(&counts[run_inst], static_cast<unsigned long long>(run_sum));
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63u) == 0) red[wave][0] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) s += red[w][0];
        atomicAdd(&st->total, s);
    }
}

// ---- serving: persistent kernel over a ring the host appends to (ServeArgs in nq_kernels.h) ----

// Trips a wave with waiting lanes keeps running its busy lanes before it looks for new tasks.
constexpr uint32_t SERVE_IDLE_TRIPS = 64;
// A wave with no task at all waits 1, 2, 4, ... 16 s_sleep(127) (about 3.4 us each at 2.4 GHz)
// between looks at the host's tail, so that idle waves do not flood the bus with reads.
constexpr uint32_t SERVE_IDLE_SLEEP_LOG2 = 4;
// Published tickets beyond head that let a wave claim `chunk` tickets at once; with fewer, it
// claims one per waiting lane, so that the last tasks spread over the waves.
constexpr uint64_t SERVE_PLENTY = 64ull * SERVE_BUF;

// The published tickets of this instance: tail[parity] - base (acquire: the entries below it
// are written).  Every active lane reads the same word (one request per wave).
__device__ __forceinline__ uint64_t serve_tail(const ServeArgs &a) {
    const uint64_t t = __hip_atomic_load(&a.ctl->tail[a.parity], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    return first_lane_u64(t > a.base ? t - a.base : 0);
}

// A ring entry in host memory: one 16-byte load past every cache (volatile: sc0 sc1).
__device__ __forceinline__ uint4 load_ring(const uint4 *p) {
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    const v4u x = *reinterpret_cast<const volatile v4u *>(p);
    return make_uint4(x.x, x.y, x.z, x.w);
}

template <bool APUSH>
__global__ void __launch_bounds__(MAX_BLOCK) serve_kernel(ServeArgs a) {
    extern __shared__ uint4 lds[];  // lds_layout(), then SERVE_BUF entries per wave
    if (static_cast<uint32_t>(reinterpret_cast<uintptr_t>(lds)) != 0u) {
        if (threadIdx.x == 0) atomicOr(a.error, DEV_ERR_LDS);
        return;
    }
    const LdsLayout L = lds_layout(lds, static_cast<uint32_t>(a.stack_depth) + (APUSH ? 1u : 0u));
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t gwave = blockIdx.x * (blockDim.x >> 6) + wave;
    uint4 *buf = reinterpret_cast<uint4 *>(L.red + MAX_WAVES_PER_BLOCK) + wave * SERVE_BUF;
    put_dead_frame(lds);
    const uint32_t stride = blockDim.x * static_cast<uint32_t>(sizeof(uint4));
    const uint32_t base = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(L.stack + threadIdx.x));
    const uint32_t sd = static_cast<uint32_t>(a.stack_depth);
    // Wave-uniform claim state: tickets [bfirst, bend) claimed, [bfirst, bload) loaded into buf,
    // [bfirst, bnext) handed to lanes; lim = published tickets as last read.
    uint64_t bfirst = 0, bend = 0, bload = 0, bnext = 0, lim = 0;
    uint32_t sp = base;
    uint32_t c = 0, l = 0, r = 0, v = 0, cnt = 0, nd = 0;
    uint32_t weight = 1;
    uint64_t tsum = 0;
    bool own = false;  // the lane holds a task it is counting
    // The lane's run of counted entries of one slot (their count and weight), added with one
    // fetch-and-add when the lane takes an entry of another slot or waits, so that a slice
    // completes with its last entry.
    uint32_t run_slot = 0;
    uint64_t run_sum = 0, run_n = 0;
    auto flush = [&]() {
        if (!run_n) return;
        const uint64_t add = (run_sum << SERVE_FIN_BITS) | run_n;
        const uint64_t now = __hip_atomic_fetch_add(&a.acc[run_slot], static_cast<unsigned long long>(add),
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + add;
        if ((now & ((1ull << SERVE_FIN_BITS) - 1)) == SERVE_FIN_DONE) {
            // The slice's last entry: `now` holds its whole count (one word, so nothing else to
            // wait for).  Clear the word before the record can let the host reuse the slot (the
            // exchange returns once done, and the record's release follows it).
            (void)__hip_atomic_exchange(&a.acc[run_slot], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t g = __hip_atomic_fetch_add(&a.gen[run_slot], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
            __hip_atomic_store(&a.done[run_slot].count, now >> SERVE_FIN_BITS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&a.done[run_slot].gen, g, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        run_sum = 0;
        run_n = 0;
    };
    uint64_t idle_since = 0;
    uint32_t idle_polls = 0;

    for (;;) {
        // (Every lane of the wave is active here.)  Account the tasks that have finished.
        if (v == 0 && own) {
            run_sum += tsum * weight;
            run_n += 1;
            own = false;
            tsum = 0;
        }
        uint64_t need = __ballot(v == 0);
        while (need) {
            if (bnext == bload) {
                if (bload == bend) {
                    // Claim: `chunk` tickets while published work is plentiful, else one per
                    // waiting lane.
                    const uint32_t k = __popcll(need);
                    uint64_t first = 0, take = 0;
                    if (lane == 0) {
                        const uint64_t h = __hip_atomic_load(a.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        take = (lim > h && lim - h >= SERVE_PLENTY) ? max(k, static_cast<uint32_t>(a.chunk)) : k;
                        first = __hip_atomic_fetch_add(a.head, static_cast<unsigned long long>(take), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
                        const uint64_t rec = (static_cast<uint64_t>(a.epoch & 0xFFFFu) << SERVE_CLAIM_EPOCH_SHIFT) |
                                             ((first + take) & ((1ull << SERVE_CLAIM_EPOCH_SHIFT) - 1));
                        __hip_atomic_store(&a.claimed[gwave], rec, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    }
                    first = first_lane_u64(first);
                    take = first_lane_u64(take);
                    bfirst = bload = bnext = first;
                    bend = first + take;
                }
                if (bload >= lim) lim = serve_tail(a);
                const uint64_t n = (lim > bload ? min(bend, lim) : bload) - bload;  // published, not loaded
                if (n == 0) break;
                if (lane < n) buf[bload - bfirst + lane] = load_ring(a.ring + ((a.base + bload + lane) & a.ring_mask));
                bload += n;
            }
            // Hand the loaded entries to the waiting lanes, lowest rank first.
            const uint64_t avail = bload - bnext;
            const uint32_t rank = lane_rank(need);
            if (((need >> lane) & 1u) && rank < avail) {
                const uint4 w = buf[bnext - bfirst + rank];
                const uint32_t s = w.w >> SERVE_TAG_SLOT_SHIFT;
                if (run_n && s != run_slot) flush();
                run_slot = s;
                const uint32_t N = w.w & 31u;
                weight = 1u + ((w.w >> 5) & 1u);
                const uint32_t last = (1u << N) - 1u;
                const uint32_t v0 = last & ~(w.x | w.y | w.z);
                const uint32_t d0 = __popc(w.x);
                if (w.w & SERVE_TAG_TERM) {
                    // The slice's terminator weighs what completes the entry field.
                    run_n += SERVE_FIN_DONE - w.x;
                } else if (d0 + 1u >= N || v0 == 0) {
                    const uint64_t direct =
                        d0 >= N ? (w.x == last ? 1u : 0u) : (d0 + 1u == N ? __popc(v0) : 0u);
                    run_sum += direct * weight;
                    run_n += 1;
                } else if (N - 1u - d0 > sd) {
                    // More frames than the server sized: flagged, counted as done (with 0).
                    atomicOr(a.error, DEV_ERR_DEPTH);
                    run_n += 1;
                } else {
                    c = w.x | ~last;
                    l = w.y;
                    r = w.z;
                    v = v0;
                    sp = base;
                    own = true;
                }
            }
            bnext += min(avail, static_cast<uint64_t>(__popcll(need)));
            need = __ballot(v == 0);
        }
        // Lanes left waiting hand in their counts, so that finished slices complete now.
        if (v == 0) flush();
        const uint64_t live = __ballot(v != 0);
        if (live == 0) {
            // Nothing to run: every lane waits for its ticket to be published.  Leave when the
            // host has asked this instance to stop (it does so only with every published task
            // counted, and publishes no more for it) or after orphan_ticks.
            const uint32_t stop_ep = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&a.ctl->stop_epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
            if (stop_ep == a.epoch) break;
            const uint64_t now = wall_clock64();
            if (idle_polls == 0) idle_since = now;
            if (now - idle_since > a.orphan_ticks) break;
            for (uint32_t i = 0; i < (1u << min(idle_polls, SERVE_IDLE_SLEEP_LOG2)); ++i) __builtin_amdgcn_s_sleep(127);
            ++idle_polls;
            continue;
        }
        idle_polls = 0;
        uint32_t wcnt = 0;  // unused: tasks of different requests are counted per lane
        hot_loop<false, APUSH, false>(c, l, r, v, sp, cnt, wcnt, nd, live, stride,
                                      live == __ballot(1) ? 0u : SERVE_IDLE_TRIPS);
        tsum += cnt;
        cnt = 0;
    }
}

__global__ void zero_u64_kernel(unsigned long long *p, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = 0;
}

__global__ void reset_state_kernel(DevState *st) {
    if (threadIdx.x == 0) {
        st->cursor = 0;
        st->total = 0;
        st->nodes = 0;
        st->error = 0;
    }
}

__global__ void fold_result_kernel(const DevState *st, uint64_t *result, const unsigned long long *xflags) {
    if (threadIdx.x == 0) {
        result[0] += st->total;
        result[1] += st->nodes;
        result[2] |= st->error | (xflags ? *xflags : 0ull);
    }
}

constexpr uint32_t EXPAND_BLOCK = 256;
constexpr uint32_t EXPAND_ITEMS = 8;  // tasks per thread per tile

// Expand every task by one row: out gets (cur|p, (left|p)<<1, (right|p)>>1) for every legal
// queen p of the next row, and a full board is carried over as it is; the counts of a task's
// outputs sum to its own count.  The order is not the input's: only totals use this.  A block
// works on tiles of EXPAND_BLOCK * EXPAND_ITEMS tasks and reserves a tile's outputs with one
// atomic on *out_n (one per task wave made that counter the bottleneck, ~1e8 atomics/s on one
// address); waves then write their outputs round by round, every lane with a j-th output
// writing it to the next consecutive slot, so one store instruction covers a contiguous range.
__global__ void __launch_bounds__(EXPAND_BLOCK)
expand_kernel(const uint32_t *__restrict__ in, uint64_t n, const unsigned long long *__restrict__ n_dev,
              uint32_t N, uint32_t *__restrict__ out, uint64_t cap, unsigned long long *__restrict__ out_n,
              unsigned long long *__restrict__ flags) {
    if (n_dev) n = min(n, static_cast<uint64_t>(*n_dev));
    __shared__ unsigned long long wave_out[EXPAND_BLOCK / 64];
    __shared__ unsigned long long tile_base;
    const uint32_t last = (1u << N) - 1u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t tile = static_cast<uint64_t>(EXPAND_BLOCK) * EXPAND_ITEMS;
    for (uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * tile; t0 < n; t0 += static_cast<uint64_t>(gridDim.x) * tile) {
        // This thread's tasks: t0 + j * EXPAND_BLOCK + threadIdx.x (coalesced loads).
        uint32_t c[EXPAND_ITEMS], l[EXPAND_ITEMS], r[EXPAND_ITEMS], v[EXPAND_ITEMS];
        uint32_t k = 0;
        for (uint32_t j = 0; j < EXPAND_ITEMS; ++j) {
            const uint64_t i = t0 + static_cast<uint64_t>(j) * EXPAND_BLOCK + threadIdx.x;
            c[j] = l[j] = r[j] = v[j] = 0;
            if (i < n) {
                c[j] = in[3 * i];
                l[j] = in[3 * i + 1];
                r[j] = in[3 * i + 2];
                // A full board has no children but counts 1: carried over as its own output,
                // marked by v == ~0 (a real candidate set never has bits above N).
                v[j] = c[j] == last ? ~0u : (last & ~(c[j] | l[j] | r[j]));
            }
            k += v[j] == ~0u ? 1u : __popc(v[j]);
        }
        const uint64_t wk = wave_sum_u64(k);
        if ((threadIdx.x & 63u) == 0) wave_out[wave] = wk;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long total = 0;
            for (uint32_t w = 0; w < EXPAND_BLOCK / 64; ++w) total += wave_out[w];
            tile_base = total ? atomicAdd(out_n, total) : 0;
            if (tile_base + total > cap) {
                // More outputs than the buffer holds (the caller's size was wrong): write
                // nothing past it and report; the consumers clamp their counts to `cap`.
                atomicOr(flags, 2ull);
                tile_base = ~0ull;
            }
        }
        __syncthreads();
        const unsigned long long tb = tile_base;
        unsigned long long o = tb;
        for (uint32_t w = 0; w < wave; ++w) o += wave_out[w];
        __syncthreads();  // wave_out and tile_base are rewritten by the next tile
        if (tb == ~0ull) continue;
        for (uint32_t j = 0; j < EXPAND_ITEMS; ++j) {
            uint32_t vj = v[j];
            for (;;) {
                const uint64_t m = __ballot(vj != 0);
                if (m == 0) break;
                if (vj != 0) {
                    const uint64_t slot = o + lane_rank(m);
                    if (vj == ~0u) {
                        out[3 * slot] = c[j];
                        out[3 * slot + 1] = l[j];
                        out[3 * slot + 2] = r[j];
                        vj = 0;
                    } else {
                        const uint32_t p = vj & (0u - vj);
                        vj ^= p;
                        out[3 * slot] = c[j] | p;
                        out[3 * slot + 1] = (l[j] | p) << 1;
                        out[3 * slot + 2] = (r[j] | p) >> 1;
                    }
                }
                o += __popcll(m);
            }
        }
    }
}

// *out += the number of outputs expand_kernel would write for these tasks (children of every
// task, 1 for a full board).  *out must be zero on entry (launch_child_count zeroes it).
__global__ void child_count_kernel(const uint32_t *__restrict__ in, uint64_t n,
                                   const unsigned long long *__restrict__ n_dev, uint32_t N,
                                   unsigned long long *__restrict__ out) {
    if (n_dev) n = min(n, static_cast<uint64_t>(*n_dev));
    const uint32_t last = (1u << N) - 1u;
    uint64_t k = 0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t c = in[3 * i];
        k += c == last ? 1u : __popc(last & ~(c | in[3 * i + 1] | in[3 * i + 2]));
    }
    k = wave_sum_u64(k);
    if ((threadIdx.x & 63u) == 0 && k) atomicAdd(out, static_cast<unsigned long long>(k));
}

const void *select_kernel(bool nodes, bool per_task, bool apush) {
    const void *k[2][2][2] = {
        {{reinterpret_cast<const void *>(&dfs_kernel<false, false, false>),
          reinterpret_cast<const void *>(&dfs_kernel<false, false, true>)},
         {reinterpret_cast<const void *>(&dfs_kernel<false, true, false>),
          reinterpret_cast<const void *>(&dfs_kernel<false, true, true>)}},
        {{reinterpret_cast<const void *>(&dfs_kernel<true, false, false>),
          reinterpret_cast<const void *>(&dfs_kernel<true, false, true>)},
         {reinterpret_cast<const void *>(&dfs_kernel<true, true, false>),
          reinterpret_cast<const void *>(&dfs_kernel<true, true, true>)}}};
    return k[nodes][per_task][apush];
}

// split_kernel for S = split rows (1..MAX_SPLIT) and a counting mode (0 totals, 1 per-task, 2 nodes).
template <int MODE>
const void *split_kernel_for(int split, bool apush) {
    const void *k[MAX_SPLIT][2] = {
        {reinterpret_cast<const void *>(&split_kernel<false, 1, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 1, MODE>)},
        {reinterpret_cast<const void *>(&split_kernel<false, 2, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 2, MODE>)},
        {reinterpret_cast<const void *>(&split_kernel<false, 3, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 3, MODE>)},
        {reinterpret_cast<const void *>(&split_kernel<false, 4, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 4, MODE>)},
        {reinterpret_cast<const void *>(&split_kernel<false, 5, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 5, MODE>)},
        {reinterpret_cast<const void *>(&split_kernel<false, 6, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 6, MODE>)},
        {reinterpret_cast<const void *>(&split_kernel<false, 7, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 7, MODE>)},
        {reinterpret_cast<const void *>(&split_kernel<false, 8, MODE>), reinterpret_cast<const void *>(&split_kernel<true, 8, MODE>)}};
    return k[split - 1][apush ? 1 : 0];
}

const void *select_split_kernel(int split, bool apush, bool per_task, bool nodes) {
    if (nodes) return split_kernel_for<2>(split, apush);
    return per_task ? split_kernel_for<1>(split, apush) : split_kernel_for<0>(split, apush);
}

// The DFS kernel of a launch: split_kernel with a device split (totals, per-task counts or node
// counts), else dfs_kernel (also per-task and node counts together).
const void *select_count_kernel(bool nodes, bool per_task, bool apush, int split) {
    if (split > 0 && !(nodes && per_task)) return select_split_kernel(split, apush, per_task, nodes);
    return select_kernel(nodes, per_task, apush);
}

const void *select_batch_kernel(bool apush) {
    return apush ? reinterpret_cast<const void *>(&batch_kernel<true>)
                 : reinterpret_cast<const void *>(&batch_kernel<false>);
}

// Dynamic LDS above 64 KiB must be allowed per kernel (gfx950 has 160 KiB per CU).
hipError_t allow_lds(const void *k, size_t lds) {
    if (lds <= 65536) return hipSuccess;
    return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
}

}  // namespace

hipError_t batch_occupancy(int block, int stack_depth, bool apush, int *blocks_per_cu) {
    const void *k = select_batch_kernel(apush);
    const size_t lds = dfs_lds_bytes(block, stack_depth + (apush ? 1 : 0));
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, block, lds);
}

hipError_t batch_fit(int block, int stack_depth, int waves_per_cu, int push, bool *apush, int *blocks_per_cu) {
    int occ[2] = {0, 0};
    hipError_t e = batch_occupancy(block, stack_depth, false, &occ[0]);
    if (e == hipSuccess) e = batch_occupancy(block, stack_depth, true, &occ[1]);
    if (e != hipSuccess) return e;
    if (waves_per_cu > 0)
        for (int &o : occ) o = std::min(o, std::max(1, waves_per_cu / (block / 64)));
    *apush = push == 2 || (push == 0 && prefer_always_push(occ[0], occ[1]));
    *blocks_per_cu = occ[*apush ? 1 : 0];
    return hipSuccess;
}

size_t serve_lds_bytes(int block, int levels) {
    // The counting layout plus SERVE_BUF ring entries per wave.
    return dfs_lds_bytes(block, levels) + static_cast<size_t>(block / 64) * SERVE_BUF * sizeof(uint4);
}

hipError_t serve_fit(int block, int stack_depth, int push, bool *apush, int *blocks_per_cu) {
    int occ[2] = {0, 0};
    for (int i = 0; i < 2; ++i) {
        const void *k = i ? reinterpret_cast<const void *>(&serve_kernel<true>)
                          : reinterpret_cast<const void *>(&serve_kernel<false>);
        const size_t lds = serve_lds_bytes(block, stack_depth + i);
        hipError_t e = allow_lds(k, lds);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[i], k, block, lds);
        if (e != hipSuccess) return e;
    }
    *apush = push == 2 || (push == 0 && prefer_always_push(occ[0], occ[1]));
    *blocks_per_cu = occ[*apush ? 1 : 0];
    return hipSuccess;
}

hipError_t launch_serve(const ServeArgs &a, hipStream_t stream) {
    const size_t lds = serve_lds_bytes(a.block, a.stack_depth + (a.apush ? 1 : 0));
    if (a.apush)
        serve_kernel<true><<<a.grid, a.block, lds, stream>>>(a);
    else
        serve_kernel<false><<<a.grid, a.block, lds, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &a, hipStream_t stream) {
    zero_u64_kernel<<<64, 256, 0, stream>>>(reinterpret_cast<unsigned long long *>(a.counts), a.ninst);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    reset_state_kernel<<<1, 64, 0, stream>>>(a.state);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = dfs_lds_bytes(a.block, a.stack_depth + (a.apush ? 1 : 0));
    e = allow_lds(select_batch_kernel(a.apush), lds);
    if (e != hipSuccess) return e;
    const uint32_t chunk = static_cast<uint32_t>(a.chunk), sd = static_cast<uint32_t>(a.stack_depth);
    auto *counts = reinterpret_cast<unsigned long long *>(a.counts);
    if (a.apush)
        batch_kernel<true><<<a.grid, a.block, lds, stream>>>(a.tasks, a.ntasks, chunk, sd, a.state, counts);
    else
        batch_kernel<false><<<a.grid, a.block, lds, stream>>>(a.tasks, a.ntasks, chunk, sd, a.state, counts);
    return hipGetLastError();
}

size_t dfs_lds_bytes(int block, int levels) {
    // Dead frame + stack + per-wave claim range and reduction slot (lds_layout()).
    return sizeof(uint4) * (1 + static_cast<size_t>(block) * static_cast<size_t>(levels)) +
           2 * MAX_WAVES_PER_BLOCK * 2 * sizeof(unsigned long long);
}

hipError_t dfs_occupancy(int block, int stack_depth, bool count_nodes, bool per_task, bool apush, int split,
                         int *blocks_per_cu) {
    const void *k = select_count_kernel(count_nodes, per_task, apush, split);
    const size_t lds = dfs_lds_bytes(block, stack_depth + (apush ? 1 : 0));
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, block, lds);
}

hipError_t launch_zero(unsigned long long *p, uint64_t n, hipStream_t stream) {
    zero_u64_kernel<<<static_cast<unsigned>(std::min<uint64_t>(64, (n + 255) / 256 + 1)), 256, 0, stream>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_expand(const uint32_t *in, uint64_t n_max, const unsigned long long *n_dev, int N, uint32_t *out,
                         uint64_t cap, unsigned long long *out_n, unsigned long long *flags, hipStream_t stream) {
    zero_u64_kernel<<<1, 64, 0, stream>>>(out_n, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t tile = static_cast<uint64_t>(EXPAND_BLOCK) * EXPAND_ITEMS;
    const uint64_t blocks = std::min<uint64_t>((n_max + tile - 1) / tile, 8192);
    expand_kernel<<<static_cast<unsigned>(std::max<uint64_t>(1, blocks)), EXPAND_BLOCK, 0, stream>>>(
        in, n_max, n_dev, static_cast<uint32_t>(N), out, cap, out_n, flags);
    return hipGetLastError();
}

hipError_t launch_child_count(const uint32_t *in, uint64_t n_max, const unsigned long long *n_dev, int N,
                              unsigned long long *out, hipStream_t stream) {
    zero_u64_kernel<<<1, 64, 0, stream>>>(out, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t blocks = std::min<uint64_t>((n_max + 255) / 256, 4096);
    child_count_kernel<<<static_cast<unsigned>(std::max<uint64_t>(1, blocks)), 256, 0, stream>>>(
        in, n_max, n_dev, static_cast<uint32_t>(N), out);
    return hipGetLastError();
}

hipError_t launch_count(const LaunchArgs &a, hipStream_t stream) {
    reset_state_kernel<<<1, 64, 0, stream>>>(a.state);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = dfs_lds_bytes(a.block, a.stack_depth + (a.apush ? 1 : 0));
    if (a.split < 0 || a.split > MAX_SPLIT) return hipErrorInvalidValue;
    const bool split = a.split > 0 && !(a.count_nodes && a.per_task);
    const void *k = select_count_kernel(a.count_nodes, a.per_task != nullptr, a.apush, a.split);
    e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    if (split && a.per_task && a.ntasks) {
        // split_kernel adds to per_task[] (dfs_kernel writes each entry once)
        zero_u64_kernel<<<static_cast<unsigned>(std::min<uint64_t>(2048, (a.ntasks + 255) / 256)), 256, 0, stream>>>(
            reinterpret_cast<unsigned long long *>(a.per_task), a.ntasks);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    uint32_t N = static_cast<uint32_t>(a.N);
    uint32_t chunk = static_cast<uint32_t>(a.chunk);
    uint32_t sd = static_cast<uint32_t>(a.stack_depth);
    const uint32_t *tasks = a.tasks;
    uint64_t ntasks = a.ntasks;
    DevState *state = a.state;
    uint64_t *per_task = a.per_task;
    uint64_t *trace = a.trace;
    const unsigned long long *ntasks_dev = a.ntasks_dev;
    void *args[] = {&tasks, &ntasks, &N, &chunk, &sd, &state, &per_task, &trace, &ntasks_dev};
    e = hipLaunchKernel(k, dim3(a.grid), dim3(a.block), args, lds, stream);
    if (e != hipSuccess) return e;
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.result) {
        fold_result_kernel<<<1, 64, 0, stream>>>(a.state, a.result, a.xflags);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace nq
