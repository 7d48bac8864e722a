// wave.h -- wavefront-level helpers (64-lane ballots and shuffles) and the
// lock-free union-find used to build simulation islands.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bepu {

// Appends `value` for every lane with `take` set: one atomic per wavefront.
// Call from converged control flow (every live lane of the wavefront).
__device__ __forceinline__ void wave_push(uint32_t* list, uint32_t* counter, bool take, uint32_t value) {
    uint64_t active = __ballot(1);
    uint64_t mask = __ballot(take);
    if (!mask) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)active) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (take) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// Workgroup-wide append: one global atomic per workgroup.  Every thread of
// the workgroup calls it (it synchronises); BLOCK is a multiple of 64.
// s_tmp holds BLOCK / 64 + 1 words of LDS.
template <int BLOCK>
__device__ __forceinline__ void block_push(uint32_t* list, uint32_t* counter, bool take, uint32_t value,
                                           uint32_t* s_tmp) {
    constexpr int NW = BLOCK / 64;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    uint64_t mask = __ballot(take);
    if (lane == 0) s_tmp[wv] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int k = 0; k < NW; ++k) {
            uint32_t c = s_tmp[k];
            s_tmp[k] = tot;
            tot += c;
        }
        s_tmp[NW] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    if (take) list[s_tmp[NW] + s_tmp[wv] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = value;
    __syncthreads();
}

// Adds `inc` to counter[key] for every valid lane and returns the value each
// lane would have seen from its own atomicAdd.  Lanes sharing the first valid
// lane's key are combined into a single atomic (a large island's bodies and
// constraints all hit one counter), so every valid lane must pass the same
// `inc`.  Call from converged control flow.
template <typename T>
__device__ __forceinline__ T wave_atomic_add(T* counter, uint32_t key, bool valid, T inc) {
    uint64_t vmask = __ballot(valid);
    if (!vmask) return 0;
    int lane = __lane_id();
    int lead = __ffsll((unsigned long long)vmask) - 1;
    uint32_t k0 = __shfl(key, lead);
    bool same = valid && key == k0;
    uint64_t smask = __ballot(same);
    T old = 0;
    if (lane == lead) old = atomicAdd(&counter[k0], inc * (T)__popcll(smask));
    old = __shfl(old, lead);
    if (same) return old + inc * (T)__popcll(smask & ((1ull << lane) - 1ull));
    if (valid) return atomicAdd(&counter[key], inc);
    return 0;
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Union-find with hooking of the larger root under the smaller one, so every
// island's representative is its smallest dynamic body id.  Finds halve the
// path as they go; concurrent unions and finds are safe because a parent
// pointer only ever moves to a smaller id of the same set.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    while (true) {
        uint32_t p = ld_agent(&parent[x]);
        if (p == x) return x;
        uint32_t g = ld_agent(&parent[p]);
        if (g != p) __hip_atomic_store(&parent[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
    }
}

__device__ __forceinline__ void uf_union(uint32_t* parent, uint32_t a, uint32_t b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        uint32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(&parent[a], &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

}  // namespace bepu
