// Write-only rate of this device for the task generator's largest row: 453,688,251 tasks of 12
// bytes (the N=27, 7-row list, 5.44 GB) written contiguously, once with one 12-byte store per lane
// (the generator's store) and once with 16-byte stores of the same bytes, grid-stride, at 8,192,
// 16,384 and 32,768 workgroups of 256.  The generator's scatter is measured against this.
//
//   write_probe
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define CHECK(x)                                                                \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
            exit(1);                                                            \
        }                                                                       \
    } while (0)

struct alignas(4) Task3 {
    uint32_t c, l, r;
};

__global__ void write12(Task3 *__restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        out[i] = Task3{static_cast<uint32_t>(i), static_cast<uint32_t>(i) + 1u, static_cast<uint32_t>(i) + 2u};
}

__global__ void write16(uint4 *__restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        out[i] = make_uint4(static_cast<uint32_t>(i), 1u, 2u, 3u);
}

int main() {
    const uint64_t n12 = 453688251ull, bytes = n12 * 12ull, n16 = bytes / 16;  // n16 * 16 <= bytes
    void *buf = nullptr;
    CHECK(hipMalloc(&buf, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int grid : {8192, 16384, 32768}) {
        for (int rep = 0; rep < 3; ++rep) {
            float ms12 = 0, ms16 = 0;
            CHECK(hipEventRecord(e0));
            write12<<<grid, 256>>>(static_cast<Task3 *>(buf), n12);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipDeviceSynchronize());
            CHECK(hipEventElapsedTime(&ms12, e0, e1));
            CHECK(hipEventRecord(e0));
            write16<<<grid, 256>>>(static_cast<uint4 *>(buf), n16);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipDeviceSynchronize());
            CHECK(hipEventElapsedTime(&ms16, e0, e1));
            printf("grid %d: 12-byte stores %.3f ms (%.2f TB/s), 16-byte stores %.3f ms (%.2f TB/s)\n", grid, ms12,
                   bytes / (ms12 * 1e9), ms16, (n16 * 16.0) / (ms16 * 1e9));
        }
    }
    CHECK(hipFree(buf));
    return 0;
}
