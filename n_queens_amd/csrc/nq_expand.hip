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

// A wave-uniform copy of the first active lane's 64-bit value.
__device__ __forceinline__ uint64_t first_lane_u64(uint64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

unsigned grid_for(uint64_t n) {
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)));
}

// A 12-byte task of the reference's list, moved as one 12-byte access.
struct alignas(4) Task3 {
    uint32_t c, l, r;
};

// The reference's 12-byte tasks {cur, left, right}, one row deeper, unrestricted (rows >= 2 of
// src/n_queens.cpp:74-113 use the full mask): children per task, then every child at its
// scanned offset, lowest column first.
__global__ void gen_child_count_kernel(const uint32_t *__restrict__ in, uint64_t n, uint32_t last,
                                       unsigned long long *__restrict__ counts) {
    // n + 1 entries: counts[n] = 0 closes the scan (its offset is the row's total)
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i <= n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        if (i == n) {
            counts[i] = 0;
            continue;
        }
        const Task3 t = reinterpret_cast<const Task3 *>(in)[i];
        counts[i] = __popc(last & ~(t.c | t.l | t.r));
    }
}

// One wave per 64 consecutive parents.  Every lane writes its parent's children, lowest column
// first, into the wave's staging buffer in LDS at their place in the wave's run (the scan's
// offsets relative to the wave's first), then the wave copies the buffer out with lane k of each
// pass storing child k: contiguous 12-byte stores, as if one lane made each child, without the
// cross-lane search for the parent.  A run longer than the buffer goes through it in windows.
// Launched with 256 lanes: four waves, one stage each.
constexpr uint32_t GEN_STAGE = 1024;  // children staged per wave: 12 KiB of LDS, 12 waves per CU

__global__ void __launch_bounds__(256) gen_expand_kernel(const uint32_t *__restrict__ in, uint64_t n, uint32_t last,
                                  const unsigned long long *__restrict__ offsets, uint32_t *__restrict__ out) {
    __shared__ Task3 stage[4][GEN_STAGE];
    const uint32_t lane = threadIdx.x & 63u;
    Task3 *buf = stage[threadIdx.x >> 6];
    const uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
    for (uint64_t base = ((blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6) * 64; base < n;
         base += nwaves * 64) {
        const uint64_t i = base + lane;
        const bool ok = i < n;
        uint32_t c = 0, l = 0, r = 0, v = 0;
        uint64_t o = 0;
        if (ok) {
            const Task3 t = reinterpret_cast<const Task3 *>(in)[i];
            c = t.c;
            l = t.l;
            r = t.r;
            v = last & ~(c | l | r);
            o = offsets[i];
        }
        const uint64_t o0 = first_lane_u64(o);  // lane 0's parent exists (base < n)
        const uint32_t lastlane = static_cast<uint32_t>(n - 1 - base < 63 ? n - 1 - base : 63);
        const uint32_t rel = ok ? static_cast<uint32_t>(o - o0) : 0u;  // at most 64 * 31 children per wave
        const uint32_t total = __shfl(rel + static_cast<uint32_t>(__popc(v)), lastlane);
        for (uint32_t w = 0; w < total; w += GEN_STAGE) {
            // this lane's children with places in [w, w + GEN_STAGE)
            uint32_t x = v, pos = rel;
            while (x && pos < w + GEN_STAGE) {
                const uint32_t q = x & (0u - x);
                x ^= q;
                if (pos >= w) buf[pos - w] = Task3{c | q, (l | q) << 1, (r | q) >> 1};
                ++pos;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t len = total - w < GEN_STAGE ? total - w : GEN_STAGE;
            Task3 *dst = reinterpret_cast<Task3 *>(out) + o0 + w;
            for (uint32_t j = lane; j < len; j += 64) dst[j] = buf[j];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
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

hipError_t generate_expand(int N, const uint32_t *seeds, uint64_t nseeds, int levels, uint32_t *d_out, uint64_t cap,
                           uint64_t *n_out, bool *too_large, hipStream_t stream) {
    const uint32_t last = N >= 32 ? 0xFFFFFFFFu : ((1u << N) - 1u);
    *n_out = nseeds;
    *too_large = false;
    if (levels <= 0 || nseeds == 0) {
        hipError_t e = hipSuccess;
        if (levels <= 0 && d_out && cap >= nseeds && nseeds) {
            e = hipMemcpyAsync(d_out, seeds, nseeds * 3 * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
        }
        if (levels > 0) *n_out = 0;
        return e;
    }
    // One device block per row: the row's tasks, then the scan of its child counts (m + 1
    // entries: the last count is 0, so the last offset is the next row's size), then the scan's
    // scratch.  A row's block is allocated once its size is known, before the row is written.
    struct Row {
        char *block = nullptr;
        uint32_t *tasks = nullptr;
        unsigned long long *counts = nullptr, *offsets = nullptr;
        void *temp = nullptr;
        size_t temp_bytes = 0;
    };
    auto round_up = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    auto alloc_row = [&](uint64_t m, Row *row) -> hipError_t {
        if (m + 1 > 0x7FFFFFFFull) {  // one scan per row takes at most 2^31 - 1 items
            *too_large = true;
            return hipSuccess;
        }
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, row->temp_bytes, row->counts, row->offsets,
                                                        static_cast<int>(m + 1), stream);
        if (e != hipSuccess) return e;
        const size_t t = round_up(m * 3 * sizeof(uint32_t)), c = round_up((m + 1) * sizeof(unsigned long long));
        if ((e = hipMalloc(&row->block, t + 2 * c + std::max<size_t>(row->temp_bytes, 1))) != hipSuccess) return e;
        row->tasks = reinterpret_cast<uint32_t *>(row->block);
        row->counts = reinterpret_cast<unsigned long long *>(row->block + t);
        row->offsets = reinterpret_cast<unsigned long long *>(row->block + t + c);
        row->temp = row->block + t + 2 * c;
        return hipSuccess;
    };
    Row cur;
    hipError_t e = alloc_row(nseeds, &cur);
    if (e == hipSuccess && !*too_large)
        e = hipMemcpyAsync(cur.tasks, seeds, nseeds * 3 * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    uint64_t m = nseeds;
    for (int lev = 0; e == hipSuccess && !*too_large && lev < levels; ++lev) {
        gen_child_count_kernel<<<grid_for(m + 1), 256, 0, stream>>>(cur.tasks, m, last, cur.counts);
        if ((e = hipGetLastError()) != hipSuccess) break;
        size_t bytes = cur.temp_bytes;
        if ((e = hipcub::DeviceScan::ExclusiveSum(cur.temp, bytes, cur.counts, cur.offsets, static_cast<int>(m + 1),
                                                  stream)) != hipSuccess)
            break;
        unsigned long long total = 0;
        if ((e = hipMemcpyAsync(&total, cur.offsets + m, sizeof total, hipMemcpyDeviceToHost, stream)) != hipSuccess)
            break;
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) break;
        if (lev == levels - 1) {
            *n_out = total;
            if (d_out && cap >= total && total) {  // else the count only
                gen_expand_kernel<<<grid_for(m), 256, 0, stream>>>(cur.tasks, m, last, cur.offsets, d_out);
                if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(stream);
            }
            break;
        }
        if (total == 0) {
            *n_out = 0;
            break;
        }
        Row next;
        if ((e = alloc_row(total, &next)) != hipSuccess || *too_large) break;
        gen_expand_kernel<<<grid_for(m), 256, 0, stream>>>(cur.tasks, m, last, cur.offsets, next.tasks);
        if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(stream);
        (void)hipFree(cur.block);
        cur = next;
        m = total;
    }
    if (cur.block) (void)hipFree(cur.block);
    return e;
}

}  // namespace nq
