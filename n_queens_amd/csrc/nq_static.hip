// nq_count_static: nq_count with no device allocation, for callers that cannot allocate device
// memory (tight serving loops, harnesses that forbid hipMalloc).  No reference counterpart.
//
// Everything nq_count keeps in library-allocated device memory lives in one file-scope device
// array, `scratch`, allocated with the code object when the library is loaded on a device and
// held until the process ends (NQ_STATIC_SCRATCH_BYTES, about 20 MiB, on every device the
// library runs on):
//
//   [0, 32)         DevState: task cursor, total, nodes, error flags (reset by a kernel per call)
//   [32, 56)        result words [sum, nodes, flags] (zeroed by a kernel per call)
//   [256, ...)      task staging: NQ_STATIC_MAX_TASKS x 12 B (the caller's list is copied in)
//   after that      per-task counts: NQ_STATIC_MAX_TASKS x 8 B (copied back when asked for)
//
// In host code `scratch` names a host-side stand-in, not a device address: the device address
// comes from hipGetSymbolAddress, once per device with that device current, and only that
// address is ever handed to kernels and copies.  Calls on one device share its scratch and are
// serialised here by a per-device lock; the scratch keeps the last call's contents, and every
// call resets what must start at zero.  Calls from different processes each have their own.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>

#include "nq.h"
#include "nq_internal.h"
#include "nq_kernels.h"

namespace nq {
namespace {

constexpr size_t STATE_OFF = 0;
constexpr size_t RESULT_OFF = 32;
constexpr size_t TASKS_OFF = 256;
constexpr size_t PER_TASK_OFF = TASKS_OFF + 12ull * NQ_STATIC_MAX_TASKS;
static_assert(PER_TASK_OFF % 256 == 0, "per-task counts stay 256-B aligned");
static_assert(PER_TASK_OFF + 8ull * NQ_STATIC_MAX_TASKS == NQ_STATIC_SCRATCH_BYTES, "scratch layout");

__device__ __align__(256) unsigned char scratch[NQ_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu[MAX_DEVICES];
unsigned char *g_scratch[MAX_DEVICES];  // device address of `scratch` per device (under g_mu[dev])
int g_cus[MAX_DEVICES];

#define NQ_HIP(call)                                      \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return fail_hip(e_, #call); \
    } while (0)

// The device address of this device's scratch (looked up on the first call on the device) and
// its compute-unit count.  Called with g_mu[dev] held.
int device_scratch(int dev, unsigned char **ptr, int *cus) {
    if (!g_scratch[dev]) {
        void *p = nullptr;
        NQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(NQ_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        NQ_HIP(hipDeviceGetAttribute(&g_cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *ptr = g_scratch[dev];
    *cus = g_cus[dev];
    return NQ_OK;
}

}  // namespace
}  // namespace nq

using namespace nq;

extern "C" int nq_count_static(int N, const uint32_t *tasks, uint64_t ntasks, const nq_config *cfg, uint64_t *sum,
                               uint64_t *per_task, uint64_t *nodes) {
    if (N < 1 || N > MAX_N) return fail(NQ_ERR_ARG, "N must be 1..28 for the GPU kernel");
    if (!sum) return fail(NQ_ERR_ARG, "sum is NULL");
    if (ntasks && !tasks) return fail(NQ_ERR_ARG, "tasks is NULL");
    if (ntasks > NQ_STATIC_MAX_TASKS) {
        char buf[160];
        snprintf(buf, sizeof buf, "%llu tasks exceed nq_count_static's scratch of %llu tasks",
                 static_cast<unsigned long long>(ntasks), static_cast<unsigned long long>(NQ_STATIC_MAX_TASKS));
        return fail(NQ_ERR_ARG, buf);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NQ_ERR_NO_DEVICE, "Failed to find any gpu!");
    int dev = 0;
    NQ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(NQ_ERR_ARG, "device index beyond the static scratch table");

    std::lock_guard<std::mutex> g(g_mu[dev]);
    unsigned char *base = nullptr;
    int cus = 0;
    int rc = device_scratch(dev, &base, &cus);
    if (rc) return rc;
    nq_config c;
    rc = plan_count(cus, N, task_min_depth(N, tasks, ntasks), ntasks, cfg, nodes != nullptr, per_task != nullptr,
                    &c);
    if (rc) return rc;

    auto *d_state = reinterpret_cast<DevState *>(base + STATE_OFF);
    auto *d_result = reinterpret_cast<uint64_t *>(base + RESULT_OFF);
    auto *d_tasks = reinterpret_cast<uint32_t *>(base + TASKS_OFF);
    auto *d_per_task = reinterpret_cast<uint64_t *>(base + PER_TASK_OFF);
    hipStream_t s = nullptr;  // the null stream: nothing to create
    if (ntasks) NQ_HIP(hipMemcpyAsync(d_tasks, tasks, ntasks * 12, hipMemcpyHostToDevice, s));
    hipError_t e = launch_zero(reinterpret_cast<unsigned long long *>(d_result), 3, s);
    if (e != hipSuccess) return fail_hip(e, "launch_zero");
    LaunchArgs a{d_tasks, ntasks, N, c.block, c.grid, c.stack_depth, c.chunk, d_state,
                 per_task ? d_per_task : nullptr, d_result, nodes != nullptr};
    a.apush = c.push == NQ_PUSH_ALWAYS;
    a.split = c.split;
    e = launch_count(a, s);
    if (e != hipSuccess) return fail_hip(e, "launch_count");
    uint64_t res[3] = {0, 0, 0};
    NQ_HIP(hipMemcpyAsync(res, d_result, sizeof res, hipMemcpyDeviceToHost, s));
    if (per_task && ntasks) NQ_HIP(hipMemcpyAsync(per_task, d_per_task, ntasks * 8, hipMemcpyDeviceToHost, s));
    NQ_HIP(hipStreamSynchronize(s));
    if (res[2] & DEV_ERR_LDS) return fail(NQ_ERR_STATE, "DFS kernel LDS does not start at address 0");
    if (res[2]) return fail(NQ_ERR_DEPTH, "a task needed more stack frames than configured");
    *sum = res[0];
    if (nodes) *nodes = res[1];
    return NQ_OK;
}
