// Order-preserving expansion of batched task lists on the device (gfx950).
//
// A batched task is {cur, left, right, tag} with N in the low 5 bits of the tag.  One level
// replaces every task that still has more than `target` free rows (N - popcount(cur)) by its
// children, in column order, and keeps every other task; the output keeps the input order (an
// exclusive scan of the per-task output counts gives each task its output offset).  Expanding
// level by level this way yields the tasks in the order a depth-first expansion of each input
// task would (children in column order), i.e. the order the host builds them in.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "nq_kernels.h"

namespace nq {

namespace {

__device__ __forceinline__ bool needs_expansion(const uint4 &t, uint32_t target) {
    const uint32_t N = t.w & 31u;
    const uint32_t d = __popc(t.x);
    return d < N && N - d > target;
}

__global__ void batch_child_count_kernel(const uint4 *__restrict__ in, uint64_t n, uint32_t target,
                                         unsigned long long *__restrict__ counts) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint4 t = in[i];
        const uint32_t last = (1u << (t.w & 31u)) - 1u;
        counts[i] = needs_expansion(t, target) ? __popc(last & ~(t.x | t.y | t.z)) : 1u;
    }
}

__global__ void batch_expand_kernel(const uint4 *__restrict__ in, uint64_t n, uint32_t target,
                                    const unsigned long long *__restrict__ offsets, uint4 *__restrict__ out) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint4 t = in[i];
        uint64_t o = offsets[i];
        if (!needs_expansion(t, target)) {
            out[o] = t;
            continue;
        }
        const uint32_t last = (1u << (t.w & 31u)) - 1u;
        uint32_t v = last & ~(t.x | t.y | t.z);
        while (v) {
            const uint32_t p = v & (0u - v);
            v ^= p;
            out[o++] = make_uint4(t.x | p, (t.y | p) << 1, (t.z | p) >> 1, t.w);
        }
    }
}

unsigned grid_for(uint64_t n) {
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)));
}

}  // namespace

hipError_t batch_expand(uint4 **tasks, uint64_t *n, uint32_t target, int levels, hipStream_t stream) {
    unsigned long long *counts = nullptr, *offsets = nullptr;
    void *temp = nullptr;
    size_t cap = 0, temp_bytes = 0;
    hipError_t e = hipSuccess;
    auto release = [&]() {
        if (counts) (void)hipFree(counts);
        if (offsets) (void)hipFree(offsets);
        if (temp) (void)hipFree(temp);
    };
    for (int lev = 0; lev < levels && *n > 0; ++lev) {
        const uint64_t m = *n;
        if (m > cap) {
            release();
            counts = offsets = nullptr;
            temp = nullptr;
            cap = m;
            if ((e = hipMalloc(&counts, cap * sizeof(unsigned long long))) != hipSuccess) break;
            if ((e = hipMalloc(&offsets, cap * sizeof(unsigned long long))) != hipSuccess) break;
            temp_bytes = 0;
            if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, counts, offsets, m, stream)) != hipSuccess)
                break;
            if ((e = hipMalloc(&temp, std::max<size_t>(temp_bytes, 1))) != hipSuccess) break;
        }
        batch_child_count_kernel<<<grid_for(m), 256, 0, stream>>>(*tasks, m, target, counts);
        if ((e = hipGetLastError()) != hipSuccess) break;
        size_t bytes = temp_bytes;
        if ((e = hipcub::DeviceScan::ExclusiveSum(temp, bytes, counts, offsets, m, stream)) != hipSuccess) break;
        unsigned long long tail[2] = {0, 0};
        if ((e = hipMemcpyAsync(&tail[0], offsets + m - 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                stream)) != hipSuccess)
            break;
        if ((e = hipMemcpyAsync(&tail[1], counts + m - 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                stream)) != hipSuccess)
            break;
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) break;
        const uint64_t total = tail[0] + tail[1];
        uint4 *out = nullptr;
        if ((e = hipMalloc(&out, std::max<uint64_t>(total, 1) * sizeof(uint4))) != hipSuccess) break;
        batch_expand_kernel<<<grid_for(m), 256, 0, stream>>>(*tasks, m, target, offsets, out);
        if ((e = hipGetLastError()) != hipSuccess) {
            (void)hipFree(out);
            break;
        }
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) {
            (void)hipFree(out);
            break;
        }
        (void)hipFree(*tasks);
        *tasks = out;
        *n = total;
    }
    release();
    return e;
}

}  // namespace nq
