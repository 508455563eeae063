// LDS pipe probe for gfx950: cycles per wave-wide ds_read / ds_write of 4, 8, 12 and 16 bytes
// per lane, at several resident waves per CU, with the DFS stack's access pattern (every lane
// its own slot in lane-interleaved levels, so no bank conflicts).  It prices the LDS side of
// the DFS iteration (one ds_write_b128 push and one ds_read_b128 pop per node).
//
// Each wave issues ITERS groups of 8 independent accesses to 8 levels of its lanes' slots and
// waits for them at the end of each group (inside the same asm statement); its shader-clock
// cycles (s_memtime) over the whole loop are recorded.  Reported: CU cycles per wave
// instruction = wave cycles / (instructions per wave x resident waves per CU), i.e. the LDS
// pipe's cost per instruction when all resident waves keep it busy.
//
//   lds_probe [iters]      prints one line per (op, bytes, waves/CU)
//
// Every address is checked on the host against the workgroup's LDS allocation before launch.
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

constexpr int LEVELS = 8;

typedef uint32_t v3u __attribute__((ext_vector_type(3)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Slot size per lane for an access of B bytes: 4 and 8 B slots are packed, 12 and 16 B accesses
// use 16-B slots (their alignment), as the DFS frames do.
__host__ __device__ constexpr uint32_t slot_bytes(int B) { return B <= 8 ? B : 16; }

#define RD8(OP, R)                                                                                \
    OP " %0, %8\n\t" OP " %1, %8 offset:%c9\n\t" OP " %2, %8 offset:%c10\n\t" OP                 \
       " %3, %8 offset:%c11\n\t" OP " %4, %8 offset:%c12\n\t" OP " %5, %8 offset:%c13\n\t" OP      \
       " %6, %8 offset:%c14\n\t" OP " %7, %8 offset:%c15\n\t"                                    \
       "s_waitcnt lgkmcnt(0)\n\t"
#define WR8(OP)                                                                                   \
    OP " %0, %1\n\t" OP " %0, %1 offset:%c2\n\t" OP " %0, %1 offset:%c3\n\t" OP                   \
       " %0, %1 offset:%c4\n\t" OP " %0, %1 offset:%c5\n\t" OP " %0, %1 offset:%c6\n\t" OP         \
       " %0, %1 offset:%c7\n\t" OP " %0, %1 offset:%c8\n\t"                                      \
       "s_waitcnt lgkmcnt(0)\n\t"

// Level stride in bytes for a 256-lane workgroup: 256 * slot.  Offsets are immediates (< 64 KiB).
template <int B>
struct Lvl {
    static constexpr uint32_t s = 256u * slot_bytes(B);
};

// Operations timed: 8 reads, 8 writes, 4 writes + 4 reads, or a chain of dependent reads.
enum Op { READ = 0, WRITE = 1, MIX = 2, CHAIN = 3 };

// A chain of 8 dependent ds_read_b32: each address is the value the previous read returned
// (every slot holds its own address, written before the loop), so each read waits for the last.
#define CHAIN8                                                                                    \
    "ds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\tds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\t"     \
    "ds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\tds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\t"     \
    "ds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\tds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\t"     \
    "ds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\tds_read_b32 %0, %0\n\ts_waitcnt lgkmcnt(0)\n\t"

// Four b128 writes to levels 0..3 and four b128 reads of levels 4..7 (the DFS's mix of one push
// and one pop), waited for together.
#define MIX8                                                                                      \
    "ds_write_b128 %4, %5\n\tds_read_b128 %0, %4 offset:%c9\n\tds_write_b128 %4, %5 offset:%c6\n\t"     \
    "ds_read_b128 %1, %4 offset:%c10\n\tds_write_b128 %4, %5 offset:%c7\n\tds_read_b128 %2, %4 offset:%c11\n\t" \
    "ds_write_b128 %4, %5 offset:%c8\n\tds_read_b128 %3, %4 offset:%c12\n\ts_waitcnt lgkmcnt(0)\n\t"

template <int OP, int B, int ACTIVE>
__global__ void __launch_bounds__(256) probe(uint32_t iters, uint64_t *cycles, uint32_t *sink) {
    extern __shared__ uint32_t lds[];
    const uint32_t addr = threadIdx.x * slot_bytes(B);  // level 0 of this lane
    constexpr uint32_t S = Lvl<B>::s;
    uint32_t acc = threadIdx.x;
    if (OP == CHAIN) {
        lds[threadIdx.x] = threadIdx.x * 4u;  // slot i holds its own byte address
        __syncthreads();
    }
    // Lanes taking part: ACTIVE of every 64, spread evenly (exec-masked instructions).
    const bool active = ((threadIdx.x & 63u) % (64 / ACTIVE)) == 0;
    const uint64_t t0 = __builtin_amdgcn_s_memtime();
    for (uint32_t it = 0; it < iters; ++it) {
        if (!active) continue;
        if (OP == CHAIN) {
            uint32_t a = addr;
            __asm__ volatile(CHAIN8 : "+v"(a) : : "memory");
            acc += a;
        } else if (OP == MIX) {
            const v4u d = {acc, acc, acc, acc};
            v4u x0, x1, x2, x3;
            __asm__ volatile(MIX8
                             : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3)
                             : "v"(addr), "v"(d), "i"(S), "i"(2 * S), "i"(3 * S), "i"(4 * S), "i"(5 * S), "i"(6 * S),
                               "i"(7 * S)
                             : "memory");
            acc += x0.x ^ x1.y ^ x2.z ^ x3.w;
        } else if (OP == WRITE) {
            if constexpr (B == 4) {
                __asm__ volatile(WR8("ds_write_b32") : : "v"(addr), "v"(acc), "i"(S), "i"(2 * S), "i"(3 * S),
                                 "i"(4 * S), "i"(5 * S), "i"(6 * S), "i"(7 * S) : "memory");
            } else if constexpr (B == 8) {
                const uint64_t d = acc;
                __asm__ volatile(WR8("ds_write_b64") : : "v"(addr), "v"(d), "i"(S), "i"(2 * S), "i"(3 * S),
                                 "i"(4 * S), "i"(5 * S), "i"(6 * S), "i"(7 * S) : "memory");
            } else if constexpr (B == 12) {
                const v3u d = {acc, acc, acc};
                __asm__ volatile(WR8("ds_write_b96") : : "v"(addr), "v"(d), "i"(S), "i"(2 * S), "i"(3 * S),
                                 "i"(4 * S), "i"(5 * S), "i"(6 * S), "i"(7 * S) : "memory");
            } else {
                const v4u d = {acc, acc, acc, acc};
                __asm__ volatile(WR8("ds_write_b128") : : "v"(addr), "v"(d), "i"(S), "i"(2 * S), "i"(3 * S),
                                 "i"(4 * S), "i"(5 * S), "i"(6 * S), "i"(7 * S) : "memory");
            }
            acc += 1;
        } else {
            if constexpr (B == 4) {
                uint32_t x0, x1, x2, x3, x4, x5, x6, x7;
                __asm__ volatile(RD8("ds_read_b32", 0)
                                 : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3), "=&v"(x4), "=&v"(x5), "=&v"(x6), "=&v"(x7)
                                 : "v"(addr), "i"(S), "i"(2 * S), "i"(3 * S), "i"(4 * S), "i"(5 * S), "i"(6 * S),
                                   "i"(7 * S)
                                 : "memory");
                acc += x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7;
            } else if constexpr (B == 8) {
                uint64_t x0, x1, x2, x3, x4, x5, x6, x7;
                __asm__ volatile(RD8("ds_read_b64", 0)
                                 : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3), "=&v"(x4), "=&v"(x5), "=&v"(x6), "=&v"(x7)
                                 : "v"(addr), "i"(S), "i"(2 * S), "i"(3 * S), "i"(4 * S), "i"(5 * S), "i"(6 * S),
                                   "i"(7 * S)
                                 : "memory");
                acc += static_cast<uint32_t>(x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7);
            } else if constexpr (B == 12) {
                v3u x0, x1, x2, x3, x4, x5, x6, x7;
                __asm__ volatile(RD8("ds_read_b96", 0)
                                 : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3), "=&v"(x4), "=&v"(x5), "=&v"(x6), "=&v"(x7)
                                 : "v"(addr), "i"(S), "i"(2 * S), "i"(3 * S), "i"(4 * S), "i"(5 * S), "i"(6 * S),
                                   "i"(7 * S)
                                 : "memory");
                acc += (x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7).x;
            } else {
                v4u x0, x1, x2, x3, x4, x5, x6, x7;
                __asm__ volatile(RD8("ds_read_b128", 0)
                                 : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3), "=&v"(x4), "=&v"(x5), "=&v"(x6), "=&v"(x7)
                                 : "v"(addr), "i"(S), "i"(2 * S), "i"(3 * S), "i"(4 * S), "i"(5 * S), "i"(6 * S),
                                   "i"(7 * S)
                                 : "memory");
                acc += x0.x ^ x1.y ^ x2.z ^ x3.w ^ x4.x ^ x5.y ^ x6.z ^ x7.w;
            }
        }
    }
    const uint64_t t1 = __builtin_amdgcn_s_memtime();
    if ((threadIdx.x & 63u) == 0) cycles[blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64] = t1 - t0;
    sink[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <int OP, int B, int ACTIVE = 64>
void run(int cus, uint32_t iters, int wgs_per_cu) {
    const int block = 256;
    // LDS per workgroup: the 8 levels, padded so that exactly wgs_per_cu workgroups fit a CU.
    const size_t used = static_cast<size_t>(LEVELS) * Lvl<B>::s;
    const size_t lds = std::max<size_t>(used, (160 * 1024) / wgs_per_cu - 256);
    // Host-side bounds check: the highest byte any lane touches.
    const size_t top = (LEVELS - 1) * static_cast<size_t>(Lvl<B>::s) + (block - 1) * slot_bytes(B) + B;
    if (top > lds || (LEVELS - 1) * Lvl<B>::s > 65535) {
        fprintf(stderr, "probe B=%d out of bounds: %zu > %zu\n", B, top, lds);
        exit(1);
    }
    auto k = probe<OP, B, ACTIVE>;
    CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds)));
    int occ = 0;
    CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void *>(k), block, lds));
    const int grid = cus * occ;
    const int waves = grid * (block / 64);
    uint64_t *d_cyc = nullptr;
    uint32_t *d_sink = nullptr;
    CHECK(hipMalloc(&d_cyc, waves * sizeof(uint64_t)));
    CHECK(hipMalloc(&d_sink, static_cast<size_t>(grid) * block * sizeof(uint32_t)));
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, 0, iters / 8 + 1, d_cyc, d_sink);  // warm
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, 0, iters, d_cyc, d_sink);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipDeviceSynchronize());
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::vector<uint64_t> cyc(waves);
    CHECK(hipMemcpy(cyc.data(), d_cyc, waves * sizeof(uint64_t), hipMemcpyDeviceToHost));
    double mean = 0;
    for (uint64_t c : cyc) mean += static_cast<double>(c);
    mean /= waves;
    const int waves_per_cu = occ * (block / 64);
    const double insts = 8.0 * iters;  // per wave
    static const char *names[] = {"read ", "write", "mix  ", "chain"};
    if (OP == CHAIN) {
        // Dependent reads: wave cycles per read is the latency (with waves_per_cu waves doing it).
        printf("%s b32  waves/CU %2d  %6.1f wave cycles per dependent read (latency)  kernel %.2f ms\n", names[OP],
               waves_per_cu, mean / insts, ms);
    } else {
        const double cyc_per_inst = mean / (insts * waves_per_cu);
        const double bytes_per_cyc = 64.0 * B * ACTIVE / 64.0 / cyc_per_inst;
        printf("%s b%-3d waves/CU %2d lanes %2d  %6.2f CU cycles per wave instruction  %6.1f B/cycle/CU  kernel "
               "%.2f ms\n",
               names[OP], 8 * B, waves_per_cu, ACTIVE, cyc_per_inst, bytes_per_cyc, ms);
    }
    fflush(stdout);
    CHECK(hipFree(d_cyc));
    CHECK(hipFree(d_sink));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
}

int main(int argc, char **argv) {
    const uint32_t iters = argc > 1 ? static_cast<uint32_t>(atoi(argv[1])) : 20000;
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int wgs : {1, 2, 3, 4}) {
        run<READ, 4>(cus, iters, wgs);
        run<READ, 8>(cus, iters, wgs);
        run<READ, 12>(cus, iters, wgs);
        run<READ, 16>(cus, iters, wgs);
        run<WRITE, 4>(cus, iters, wgs);
        run<WRITE, 8>(cus, iters, wgs);
        run<WRITE, 12>(cus, iters, wgs);
        run<WRITE, 16>(cus, iters, wgs);
        run<MIX, 16>(cus, iters, wgs);
        run<CHAIN, 4>(cus, iters, wgs);
    }
    // Exec-masked b128 accesses: does the cost follow the active lanes?
    for (int wgs : {3, 4}) {
        run<WRITE, 16, 32>(cus, iters, wgs);
        run<WRITE, 16, 16>(cus, iters, wgs);
        run<WRITE, 16, 1>(cus, iters, wgs);
        run<READ, 16, 32>(cus, iters, wgs);
        run<READ, 16, 16>(cus, iters, wgs);
        run<READ, 16, 1>(cus, iters, wgs);
    }
    return 0;
}
