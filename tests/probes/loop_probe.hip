// Timing probe for gfx950: the plain DFS hot loop's instruction mix (13 VALU, the exec switch of
// the pop, the per-wave board count, one LDS write and one LDS read per iteration) at fixed LDS
// addresses and a fixed trip count, with the LDS accesses changed one at a time, at 4, 8, 12 and
// 16 resident waves per CU.  It locates the resource that bounds the DFS iteration: which LDS
// access, and at what width, the loop can afford.  The last mode is the product's iteration with
// the forced move (20 VALU per iteration for 1.37 placements).
//
// Every lane keeps one slot per level (lane-interleaved 16-B slots, no bank conflicts, the DFS
// stack's pattern) and never moves its stack pointer: the pointer arithmetic of the loop is kept
// (same instructions and dependencies) with a zero stride.  The data evolve as in the DFS (the
// popped frame is the frame written), so exec masks vary per lane as they do there; counts are
// not checked, only time.  The grid is exactly the resident waves, so the kernel time divided by
// the iterations of one wave is the time of one wave-iteration.
//
//   loop_probe [trips] [first mode]      one line per (mode, waves per CU), best of 3 launches
//
// Every LDS address is below 16 + 2 * 256 * 16 + 4096 bytes, inside every allocation used.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#define CHECK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                               \
            exit(1);                                                                              \
        }                                                                                         \
    } while (0)

// Writes of the frame v[20:23] (or part of it) to the lane's slot, reads back into it.
#define W128 "ds_write_b128 %[d], v[20:23]\n\t"
#define W96 "ds_write_b96 %[d], v[20:22]\n\t"
#define W64 "ds_write_b64 %[d], v[20:21]\n\t"
#define W32 "ds_write_b32 %[d], v23\n\t"
#define R128 "ds_read_b128 v[20:23], %[d]\n\t"
// Exec-masked b128 writes: a quarter of the lanes, or all of them, behind the same exec switch.
#define WM25 "v_cmpx_eq_u32_e64 %[m1], 0, %[sel]\n\t" W128 "s_mov_b64 exec, %[live]\n\t"
#define WM100 "v_cmpx_ne_u32_e64 %[m1], -1, %[sel]\n\t" W128 "s_mov_b64 exec, %[live]\n\t"
// The b128 write followed by one more VALU that nothing waits for (one VALU above the DFS's 13).
#define W128V W128 "v_xor_b32 v28, v28, v21\n\t"

// The product's iteration (nq_dfs.hip, unconditional push, per-wave count) with WR at the push
// and RD at the pop.
#define ITER(WR, RD)                                                                              \
    "v_sub_u32 %[t], 0, v23\n\t"                                                                  \
    "v_and_b32 %[p], v23, %[t]\n\t"                                                               \
    "v_xor_b32 v23, v23, %[p]\n\t" WR                                                             \
    "v_min_u32 %[t], 1, v23\n\t"                                                                  \
    "v_mad_u32_u24 %[d], %[t], 0, %[d]\n\t"                                                       \
    "v_or_b32 v22, v22, %[p]\n\t"                                                                 \
    "v_or_b32 v20, v20, %[p]\n\t"                                                                 \
    "v_lshrrev_b32 v22, 1, v22\n\t"                                                               \
    "v_add_lshl_u32 v21, v21, %[p], 1\n\t"                                                        \
    "v_cmp_eq_u32 vcc, -1, v20\n\t"                                                               \
    "v_bitop3_b32 v23, v21, v22, v20 bitop3:1\n\t"                                                \
    "v_cmpx_eq_u32_e64 %[m1], 0, v23\n\t"                                                         \
    "v_sub_u32_e64 %[d], %[d], 0 clamp\n\t" RD                                                    \
    "s_mov_b64 exec, %[live]\n\t"                                                                 \
    "s_bcnt1_i32_b64 %[bc], vcc\n\t"                                                              \
    "s_add_u32 %[wcnt], %[wcnt], %[bc]\n\t"                                                       \
    "s_waitcnt lgkmcnt(0)\n\t"

// The same iteration with the pop's read issued before the push's write: the parent frame is kept
// in v[24:27] (a second register set) and written after the read, one level up.
#define ITER_READ_FIRST                                                                           \
    "v_sub_u32 %[t], 0, v23\n\t"                                                                  \
    "v_and_b32 %[p], v23, %[t]\n\t"                                                               \
    "v_xor_b32 v27, v23, %[p]\n\t"                                                                \
    "v_min_u32 %[t], 1, v27\n\t"                                                                  \
    "v_mad_u32_u24 %[d], %[t], 0, %[d]\n\t"                                                       \
    "v_or_b32 v22, v22, %[p]\n\t"                                                                 \
    "v_or_b32 v20, v20, %[p]\n\t"                                                                 \
    "v_lshrrev_b32 v22, 1, v22\n\t"                                                               \
    "v_add_lshl_u32 v21, v21, %[p], 1\n\t"                                                        \
    "v_cmp_eq_u32 vcc, -1, v20\n\t"                                                               \
    "v_bitop3_b32 v23, v21, v22, v20 bitop3:1\n\t"                                                \
    "v_cmpx_eq_u32_e64 %[m1], 0, v23\n\t"                                                         \
    "v_sub_u32_e64 %[d], %[d], 0 clamp\n\t"                                                       \
    "ds_read_b128 v[20:23], %[d]\n\t"                                                             \
    "s_mov_b64 exec, %[live]\n\t"                                                                 \
    "ds_write_b128 %[d], v[24:27] offset:4096\n\t"                                                \
    "s_bcnt1_i32_b64 %[bc], vcc\n\t"                                                              \
    "s_add_u32 %[wcnt], %[wcnt], %[bc]\n\t"                                                       \
    "s_waitcnt lgkmcnt(0)\n\t"

// The product's iteration with the forced move (nq_dfs.hip: NQ_EXTRACT, NQ_PUSH_ALWAYS, NQ_CHILD,
// NQ_FORCED, NQ_COUNT_POP with the per-wave count): 20 VALU, one push and one pop.
#define ITER_FORCED                                                                               \
    "v_sub_u32 %[t], 0, v23\n\t"                                                                  \
    "v_and_b32 %[p], v23, %[t]\n\t"                                                               \
    "v_xor_b32 v23, v23, %[p]\n\t" W128                                                          \
    "v_min_u32 %[t], 1, v23\n\t"                                                                  \
    "v_mad_u32_u24 %[d], %[t], 0, %[d]\n\t"                                                       \
    "v_or_b32 v22, v22, %[p]\n\t"                                                                 \
    "v_or_b32 v20, v20, %[p]\n\t"                                                                 \
    "v_lshrrev_b32 v22, 1, v22\n\t"                                                               \
    "v_add_lshl_u32 v21, v21, %[p], 1\n\t"                                                        \
    "v_bitop3_b32 v23, v21, v22, v20 bitop3:1\n\t"                                                \
    "v_bcnt_u32_b32 %[t], v23, 0\n\t"                                                             \
    "v_cmpx_eq_u32_e64 %[m1], 1, %[t]\n\t"                                                        \
    "v_or_b32 v22, v22, v23\n\t"                                                                  \
    "v_or_b32 v20, v20, v23\n\t"                                                                  \
    "v_lshrrev_b32 v22, 1, v22\n\t"                                                               \
    "v_add_lshl_u32 v21, v21, v23, 1\n\t"                                                         \
    "v_bitop3_b32 v23, v21, v22, v20 bitop3:1\n\t"                                                \
    "s_mov_b64 exec, %[live]\n\t"                                                                 \
    "v_cmp_eq_u32 vcc, -1, v20\n\t"                                                               \
    "v_cmpx_eq_u32_e64 %[m1], 0, v23\n\t"                                                         \
    "v_sub_u32_e64 %[d], %[d], 0 clamp\n\t" R128                                                  \
    "s_mov_b64 exec, %[live]\n\t"                                                                 \
    "s_bcnt1_i32_b64 %[bc], vcc\n\t"                                                              \
    "s_add_u32 %[wcnt], %[wcnt], %[bc]\n\t"                                                       \
    "s_waitcnt lgkmcnt(0)\n\t"

#define X4(B) B B B B
#define X16(B) X4(B) X4(B) X4(B) X4(B)

#define LOOP(BODY)                                                                                \
    __asm__ volatile("s_mov_b64 %[live], exec\n"                                                  \
                     "1:\n\t" X16(BODY)                                                            \
                     "s_sub_u32 %[n], %[n], 1\n\t"                                                \
                     "s_cmp_lg_u32 %[n], 0\n\t"                                                   \
                     "s_cbranch_scc1 1b\n\t"                                                      \
                     : "+{v20}"(c), "+{v21}"(l), "+{v22}"(r), "+{v23}"(v), [d] "+v"(d), [n] "+s"(n), \
                       [wcnt] "+s"(wcnt), [t] "=&v"(t), [p] "=&v"(p), [m1] "=&s"(m1), [bc] "=&s"(bc), \
                       [live] "=&s"(live)                                                         \
                     : [sel] "v"(sel)                                                             \
                     : "memory", "scc", "vcc", "v24", "v25", "v26", "v27", "v28")

constexpr int NMODES = 12;
const char *const MODE_NAMES[NMODES] = {
    "b128 write + b128 read (the DFS)", "b128 read only",         "b128 write only",
    "no LDS access",                    "b96 write + b128 read",  "b64 write + b128 read",
    "b32 write + b128 read",            "b128 write, 25% lanes",  "b128 write, all lanes, same exec switch",
    "b128 read issued before the b128 write",
    "the DFS plus one VALU (14 VALU)",
    "the DFS with the forced move (20 VALU)"};

template <int MODE>
__global__ void __launch_bounds__(256) loop_probe(uint32_t trips, uint32_t *sink) {
    extern __shared__ uint4 lds[];
    lds[1 + threadIdx.x] = make_uint4(0xFFFF0000u | threadIdx.x, 0, 0, 0x0000FFFFu);
    lds[1 + 256 + threadIdx.x] = make_uint4(0xFFFF0000u | threadIdx.x, 0, 0, 0x0000FFFFu);
    __syncthreads();
    uint32_t d = 16u + threadIdx.x * 16u;  // this lane's level-0 slot
    uint32_t c = 0xFFFF0000u | threadIdx.x, l = threadIdx.x * 7u, r = threadIdx.x * 13u, v = 0x0000FFFFu;
    uint32_t n = trips, wcnt = 0, t, p, bc;
    const uint32_t sel = threadIdx.x & 3u;
    uint64_t m1, live;
    if (MODE == 0) LOOP(ITER(W128, R128));
    if (MODE == 1) LOOP(ITER("", R128));
    if (MODE == 2) LOOP(ITER(W128, ""));
    if (MODE == 3) LOOP(ITER("", ""));
    if (MODE == 4) LOOP(ITER(W96, R128));
    if (MODE == 5) LOOP(ITER(W64, R128));
    if (MODE == 6) LOOP(ITER(W32, R128));
    if (MODE == 7) LOOP(ITER(WM25, R128));
    if (MODE == 8) LOOP(ITER(WM100, R128));
    if (MODE == 9) LOOP(ITER_READ_FIRST);
    if (MODE == 10) LOOP(ITER(W128V, R128));
    if (MODE == 11) LOOP(ITER_FORCED);
    sink[blockIdx.x * 256 + threadIdx.x] = c ^ l ^ r ^ v ^ wcnt;
}

typedef void (*kfn)(uint32_t, uint32_t *);

int main(int argc, char **argv) {
    const uint32_t trips = argc > 1 ? static_cast<uint32_t>(atoi(argv[1])) : 20000;
    const int m0 = argc > 2 ? atoi(argv[2]) : 0;
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    const kfn ks[NMODES] = {loop_probe<0>, loop_probe<1>, loop_probe<2>, loop_probe<3>, loop_probe<4>,
                            loop_probe<5>, loop_probe<6>, loop_probe<7>, loop_probe<8>, loop_probe<9>,
                            loop_probe<10>, loop_probe<11>};
    const int wpcs[4] = {4, 8, 12, 16};
    // LDS per 256-lane workgroup such that exactly waves/4 workgroups fit in 160 KiB.
    const size_t ldsz[4] = {96 * 1024, 60 * 1024, 44 * 1024, 34 * 1024};
    const size_t used = 16 + 2 * 256 * 16 + 4096;
    uint32_t *sink;
    CHECK(hipMalloc(&sink, sizeof(uint32_t) * cus * 4 * 256));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    printf("device %s, %d CUs, %u trips x 16 iterations per wave; CU cycles at 2.4 GHz\n", prop.gcnArchName, cus, trips);
    for (int m = m0; m < NMODES; ++m) {
        for (int w = 0; w < 4; ++w) {
            if (ldsz[w] < used || ldsz[w] > prop.sharedMemPerBlock) {
                fprintf(stderr, "LDS size %zu out of range\n", ldsz[w]);
                return 1;
            }
            const int blocks = cus * (wpcs[w] / 4);
            CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(ks[m]), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(ldsz[w])));
            float ms = 1e30f;  // best of three launches (the clock is not pinned)
            for (int rep = 0; rep < 3; ++rep) {
                CHECK(hipEventRecord(e0, 0));
                hipLaunchKernelGGL(ks[m], dim3(blocks), dim3(256), ldsz[w], 0, trips, sink);
                CHECK(hipGetLastError());
                CHECK(hipEventRecord(e1, 0));
                CHECK(hipDeviceSynchronize());
                float t = 0;
                CHECK(hipEventElapsedTime(&t, e0, e1));
                if (t < ms) ms = t;
            }
            const double ns = ms * 1e6 / (static_cast<double>(trips) * 16.0);  // per wave-iteration
            printf("mode %d  %-40s waves/CU %2d  %7.2f ns per wave-iteration  %6.1f wave cycles  %5.2f CU cycles\n", m,
                   MODE_NAMES[m], wpcs[w], ns, ns * 2.4, ns * 2.4 / wpcs[w]);
            fflush(stdout);
        }
    }
    return 0;
}
