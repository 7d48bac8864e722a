// noalloc.cpp -- bepu_get_motion_states_ranges_noalloc: the readback of
// body ranges for callers that cannot allocate (code under HIP graph capture,
// tight serving loops).  Its device scratch is a fixed array of this file,
// resolved to a device address once per device.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "../../include/bepu_gpu.h"
#include "world_host.h"

using namespace bepu;

// ---------------------------------------------------------------------------
// Scratch of bepu_get_motion_states_ranges_noalloc: a fixed device array, so
// that the call needs no allocation.  Each device holds its own copy; host
// code reaches it only through the per-device address from
// hipGetSymbolAddress (the array's name is a host-side stand-in).
namespace {
constexpr uint32_t kNoAllocMaxRanges = BEPU_NOALLOC_MAX_RANGES;
constexpr uint32_t kNoAllocMaxBodies = BEPU_NOALLOC_MAX_BODIES;
struct NoAllocScratch {
    uint32_t first[kNoAllocMaxRanges];
    uint32_t count[kNoAllocMaxRanges];
    uint32_t offset[kNoAllocMaxRanges];
    bepu_motion_state states[kNoAllocMaxBodies];
};
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;                                // serialises the address lookups
std::atomic<NoAllocScratch*> g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device
std::mutex g_noalloc_mutex[kMaxDevices];                   // one no-allocation call at a time per device
}  // namespace

static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

// Resolves the scratch array's address on `device` (the current device) once.
static bepu_status resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return BEPU_E_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device].load()) return BEPU_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BEPU_E_DEVICE;
    g_scratch[device].store(static_cast<NoAllocScratch*>(p));
    return BEPU_OK;
}

// Exclusive scan of the range counts into the range offsets (one workgroup).
__global__ void __launch_bounds__(1024) k_noalloc_offsets(NoAllocScratch* sc, uint32_t n) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t t = threadIdx.x;
    const uint32_t b = t * per;
    uint32_t sum = 0;
    for (uint32_t i = b; i < b + per && i < n; ++i) sum += sc->count[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        uint32_t v = t >= d ? part[t - d] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t i = b; i < b + per && i < n; ++i) {
        sc->offset[i] = run;
        run += sc->count[i];
    }
}

// One workgroup per range: the range's bodies to their MotionState records.
__global__ void __launch_bounds__(256) k_noalloc_gather(DeviceWorld w, NoAllocScratch* sc) {
    const uint32_t r = blockIdx.x;
    const uint32_t first = sc->first[r], count = sc->count[r], off = sc->offset[r];
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
        const uint32_t b = first + i;
        const float4 p = w.pos[b], q = w.quat[b], v = w.linv[b], a = w.angv[b];
        bepu_motion_state& m = sc->states[off + i];
        m.position[0] = p.x; m.position[1] = p.y; m.position[2] = p.z;
        m.orientation[0] = q.x; m.orientation[1] = q.y; m.orientation[2] = q.z; m.orientation[3] = q.w;
        m.linear_velocity[0] = v.x; m.linear_velocity[1] = v.y; m.linear_velocity[2] = v.z;
        m.angular_velocity[0] = a.x; m.angular_velocity[1] = a.y; m.angular_velocity[2] = a.z;
    }
}

extern "C" {

bepu_status bepu_get_motion_states_ranges_noalloc(bepu_world w, const uint32_t* firsts, const uint32_t* counts,
                                                  uint32_t nranges, bepu_motion_state* out) {
    if (!w || (nranges && (!firsts || !counts || !out))) return BEPU_E_INVALID_ARG;
    if (nranges > kNoAllocMaxRanges) return BEPU_E_CAPACITY;
    // Bodies added since the last step would need an upload, which may allocate.
    if (!w->staged.empty()) return BEPU_E_INVALID_ARG;
    uint64_t total = 0;
    for (uint32_t r = 0; r < nranges; ++r) {
        if ((uint64_t)firsts[r] + counts[r] > w->d.nbodies) return BEPU_E_INVALID_ARG;
        total += counts[r];
    }
    if (total > kNoAllocMaxBodies) return BEPU_E_CAPACITY;
    if (w->device < 0 || w->device >= kMaxDevices) return BEPU_E_DEVICE;
    // First call on this device: resolve the scratch array's address there.
    if (!g_scratch[w->device].load()) {
        int prev = 0;
        if (hipGetDevice(&prev) != hipSuccess) return BEPU_E_DEVICE;
        if (prev != w->device && hipSetDevice(w->device) != hipSuccess) return BEPU_E_DEVICE;
        bepu_status rs = resolve_scratch(w->device);
        if (prev != w->device) (void)hipSetDevice(prev);
        if (rs != BEPU_OK) return rs;
    }
    NoAllocScratch* sc = g_scratch[w->device].load();
    // The array is the device's, shared by every world on it: one call at a
    // time (each call waits for its copies before it returns).
    std::lock_guard<std::mutex> lock(g_noalloc_mutex[w->device]);
    if (nranges) {
        if (hipMemcpyAsync(sc->first, firsts, nranges * sizeof(uint32_t), hipMemcpyHostToDevice, w->stream) !=
                hipSuccess ||
            hipMemcpyAsync(sc->count, counts, nranges * sizeof(uint32_t), hipMemcpyHostToDevice, w->stream) !=
                hipSuccess)
            return BEPU_E_DEVICE;
        hipLaunchKernelGGL(k_noalloc_offsets, dim3(1), dim3(1024), 0, w->stream, sc, nranges);
        hipLaunchKernelGGL(k_noalloc_gather, dim3(nranges), dim3(256), 0, w->stream, w->d, sc);
        if (hipGetLastError() != hipSuccess) return BEPU_E_DEVICE;
        if (total && hipMemcpyAsync(out, sc->states, (size_t)total * sizeof(bepu_motion_state),
                                    hipMemcpyDeviceToHost, w->stream) != hipSuccess)
            return BEPU_E_DEVICE;
    }
    return poll_errors(w);
}

}  // extern "C"
