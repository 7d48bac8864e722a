// pairbuf.h -- broadphase helpers shared by the tree and the grid: the
// closed-interval box test and a per-workgroup LDS buffer of found pairs that
// is appended to the global overlap list with one atomic per flush.
#pragma once
#include "world.h"

namespace bepu {

// BoundingBox.Intersects (BoundingBox.cs:58-69): closed intervals on all axes.
__device__ __forceinline__ bool overlap(float4 amin, float4 amax, float4 bmin, float4 bmax) {
    if (bmin.x > amax.x || bmin.y > amax.y || bmin.z > amax.z) return false;
    if (amin.x > bmax.x || amin.y > bmax.y || amin.z > bmax.z) return false;
    return true;
}

// Appends one pair straight to the global overlap list.
__device__ __forceinline__ void pair_out(DeviceWorld& w, uint64_t key) {
    uint32_t idx = atomicAdd(&w.counters[ctr(CTR_PAIRS)], 1u);
    if (idx < w.pair_cap) w.pair_keys[idx] = key;
    else w.counters[ctr(CTR_OVERFLOW)] = 1;
}

// Appends the pairs of the lanes with `found` set to the LDS buffer with one
// LDS atomic per wavefront; entries past `cap` go straight to the global list.
// Safe in divergent code (only live lanes take part).
__device__ __forceinline__ void buf_push(DeviceWorld& w, uint64_t* buf, uint32_t* buf_n, uint32_t cap, bool found,
                                         uint64_t key) {
    uint64_t mask = __ballot(found);
    if (!mask) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(buf_n, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (found) {
        uint32_t at = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (at < cap) buf[at] = key;
        else pair_out(w, key);
    }
}

// Writes the buffered pairs out with one global atomic; every thread of the
// workgroup calls it (it synchronises) and the buffer is empty afterwards.
template <int BLOCK>
__device__ __forceinline__ void buf_flush(DeviceWorld& w, uint64_t* buf, uint32_t* buf_n, uint32_t cap,
                                          uint32_t* base_slot) {
    __syncthreads();
    uint32_t n = min(*buf_n, cap);
    if (threadIdx.x == 0) *base_slot = n ? atomicAdd(&w.counters[ctr(CTR_PAIRS)], n) : 0;
    __syncthreads();
    uint32_t base = *base_slot;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        uint32_t idx = base + i;
        if (idx < w.pair_cap) w.pair_keys[idx] = buf[i];
        else w.counters[ctr(CTR_OVERFLOW)] = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) *buf_n = 0;
    __syncthreads();
}

}  // namespace bepu
