// Copy kernel of known size, the calibration for rocprofv3's FETCH_SIZE / WRITE_SIZE counters:
// it reads and writes exactly `bytes` (default 256 MiB) once, 16 B per lane per access, so the
// counters' reading for it gives the factor between the counter and the bytes moved.
//
//   copy_probe [MiB]
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

__global__ void copy_kernel(const uint4 *__restrict__ in, uint4 *__restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        out[i] = in[i];
}

int main(int argc, char **argv) {
    const uint64_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 256;
    const uint64_t bytes = mib << 20, n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    CHECK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    CHECK(hipEventRecord(e0));
    copy_kernel<<<4096, 256>>>(a, b, n);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipDeviceSynchronize());
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    printf("copy %llu MiB (%llu KiB read, %llu KiB written) in %.3f ms: %.1f GB/s\n",
           static_cast<unsigned long long>(mib), static_cast<unsigned long long>(bytes >> 10),
           static_cast<unsigned long long>(bytes >> 10), ms, 2.0 * bytes / (ms * 1e6));
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
