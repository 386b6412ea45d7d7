// mpm_noalloc.hip -- mpm_step for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The substep's scratch -- the accumulator and velocity
// grids that mpm_step otherwise takes as a caller workspace -- lives in the
// file-scope device array `scratch` below, sized for grids up to
// kNoAllocMaxGrid^3 nodes.  Every GPU the library runs on holds its own
// copy of the array until the process ends.  Calls that overlap on one GPU
// share that copy and race; the caller serialises them (one stream, or
// events).
#include <hip/hip_runtime.h>
#include <mutex>

#include "mpm_common.h"
#include "mpm_hip.h"

using namespace mpm;

namespace {
constexpr int kNoAllocMaxGrid = 64;
constexpr int64_t kScratchNodes = (int64_t)kNoAllocMaxGrid * kNoAllocMaxGrid * kNoAllocMaxGrid;
constexpr int kMaxDevices = 64;

// accumulator grid then velocity grid, one float4 per node each
static __device__ __align__(256) float4 scratch[2 * kScratchNodes];

// the device address of `scratch` on each device, fetched once per device
// with that device current (the host-side name of the array is not a device
// address)
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

int scratch_of_current_device(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return set_err(-60, "mpm_step_noalloc: no current device (or device index >= 64)");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return set_err(-61, "mpm_step_noalloc: hipGetSymbolAddress(scratch) failed");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}
}  // namespace

extern "C" {

int mpm_noalloc_max_grid(void) { return kNoAllocMaxGrid; }

int mpm_step_noalloc(const MpmParams* prm, const MpmState* st, const MpmCollider* colliders, int n_colliders,
                     float time, float dt, void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if (prm->n_grid > kNoAllocMaxGrid)
        return set_err(-62, "mpm_step_noalloc: n_grid exceeds the static scratch (at most 64^3 nodes); use mpm_step");
    void* ws = nullptr;
    if ((rc = scratch_of_current_device(&ws))) return rc;
    // mpm_step clears the accumulator with a kernel before its P2G and
    // overwrites every velocity node, so nothing of an earlier call leaks in
    return mpm_step(prm, st, colliders, n_colliders, ws, (int64_t)sizeof(float4) * 2 * kScratchNodes, time, dt,
                    stream);
}

}  // extern "C"
