// k_broadphase.hip -- broadphase overlap finding: Morton codes of AABB
// centres, radix sort, Karras LBVH build, bottom-up refit and a stack
// traversal per leaf whose hits are compacted with wavefront ballots.
//
// Replaces DynamicHierarchy.UpdateSingleThreaded (DynamicHierarchy.cs:194-214).
// The pair SET is what the reference produces: leaf boxes are the bodies'
// AABBs (DynamicHierarchyNode.cs:766-769), leaf-leaf overlap is the closed
// interval test of BoundingBox.Intersects (BoundingBox.cs:58-69) and pairs
// pass the default collision rule (kinematic-kinematic is NoBroadPhase,
// CollisionRules.cs:175-178; dynamic bodies are always active here).
#include <hipcub/hipcub.hpp>

#include "world.h"

namespace bepu {

__device__ __forceinline__ float ord2f(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return __uint_as_float(u);
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void __launch_bounds__(256) k_morton(DeviceWorld w) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nbodies) return;
    float lo[3] = {ord2f(w.bounds[0]), ord2f(w.bounds[1]), ord2f(w.bounds[2])};
    float hi[3] = {ord2f(w.bounds[3]), ord2f(w.bounds[4]), ord2f(w.bounds[5])};
    float4 mn = w.bmin[i], mx = w.bmax[i];
    float c[3] = {0.5f * (mn.x + mx.x), 0.5f * (mn.y + mx.y), 0.5f * (mn.z + mx.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float ext = hi[k] - lo[k];
        float t = ext > 0 ? (c[k] - lo[k]) / ext : 0.5f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        q[k] = min((uint32_t)(t * 1024.0f), 1023u);
    }
    w.morton[i] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    w.leaf_ids_tmp[i] = i;
}

// Karras 2012 ("Maximizing parallelism in the construction of BVHs, octrees
// and k-d trees"): internal node i covers a contiguous range of sorted leaves.
__device__ __forceinline__ int delta(const uint32_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ void __launch_bounds__(256) k_lbvh_build(DeviceWorld w) {
    int n = (int)w.nbodies;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const uint32_t* k = w.morton_sorted;
    int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(k, n, i, j);
    int s = 0;
    for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    int left = (lo == gamma) ? ~gamma : gamma;             // leaf when the range is a single leaf
    int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    w.node_child[i] = make_int2(left, right);
    w.node_parent[left >= 0 ? left : (n - 1) + ~left] = i;
    w.node_parent[right >= 0 ? right : (n - 1) + ~right] = i;
    if (i == 0) w.node_parent[0] = -1;
}

// Bottom-up refit: the second child to arrive at a node writes its box.
__global__ void __launch_bounds__(256) k_lbvh_refit(DeviceWorld w) {
    int n = (int)w.nbodies;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int node = w.node_parent[(n - 1) + j];
    while (node >= 0) {
        __threadfence();
        if (atomicAdd(&w.node_visit[node], 1u) == 0) return;
        __threadfence();
        int2 c = w.node_child[node];
        float4 amin, amax, bmin, bmax;
        if (c.x >= 0) { amin = w.node_min[c.x]; amax = w.node_max[c.x]; }
        else { uint32_t b = w.leaf_ids[~c.x]; amin = w.bmin[b]; amax = w.bmax[b]; }
        if (c.y >= 0) { bmin = w.node_min[c.y]; bmax = w.node_max[c.y]; }
        else { uint32_t b = w.leaf_ids[~c.y]; bmin = w.bmin[b]; bmax = w.bmax[b]; }
        w.node_min[node] = make_float4(fminf(amin.x, bmin.x), fminf(amin.y, bmin.y), fminf(amin.z, bmin.z), 0.0f);
        w.node_max[node] = make_float4(fmaxf(amax.x, bmax.x), fmaxf(amax.y, bmax.y), fmaxf(amax.z, bmax.z), 0.0f);
        node = w.node_parent[node];
    }
}

__device__ __forceinline__ bool overlap(float4 amin, float4 amax, float4 bmin, float4 bmax) {
    if (bmin.x > amax.x || bmin.y > amax.y || bmin.z > amax.z) return false;
    if (amin.x > bmax.x || amin.y > bmax.y || amin.z > bmax.z) return false;
    return true;
}

// Appends `key` for every lane with `found` set; one atomic per wavefront.
__device__ __forceinline__ void wave_append(DeviceWorld& w, bool found, uint64_t key) {
    uint64_t active = __ballot(1);
    uint64_t mask = __ballot(found);
    if (mask == 0) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)active) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&w.counters[CTR_PAIRS], (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (found) {
        uint32_t idx = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (idx < w.pair_cap) w.pair_keys[idx] = key;
        else w.counters[CTR_OVERFLOW] = 1;
    }
}

constexpr int kStack = 64;

// One thread per sorted leaf j; reports leaves j' > j whose boxes overlap.
__global__ void __launch_bounds__(256) k_traverse(DeviceWorld w) {
    __shared__ int stack_mem[kStack][256];
    int n = (int)w.nbodies;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = j < n;
    uint32_t body = alive ? w.leaf_ids[j] : 0;
    float4 qmin = alive ? w.bmin[body] : make_float4(0, 0, 0, 0);
    float4 qmax = alive ? w.bmax[body] : make_float4(0, 0, 0, 0);
    bool qdyn = alive ? ((w.bflags[body] & 1u) != 0) : false;
    int sp = 0;
    if (alive && n > 1) stack_mem[sp++][threadIdx.x] = 0;
    // All lanes of the wave iterate together so that ballots see every lane.
    while (__ballot(sp > 0) != 0) {
        int node = -1;
        if (sp > 0) node = stack_mem[--sp][threadIdx.x];
        int2 c = node >= 0 ? w.node_child[node] : make_int2(0, 0);
        for (int side = 0; side < 2; ++side) {
            bool found = false;
            uint64_t key = 0;
            if (node >= 0) {
                int ch = side == 0 ? c.x : c.y;
                if (ch >= 0) {
                    if (overlap(qmin, qmax, w.node_min[ch], w.node_max[ch])) {
                        if (sp < kStack) stack_mem[sp++][threadIdx.x] = ch;
                        else w.counters[CTR_OVERFLOW] = 2;
                    }
                } else {
                    int leaf = ~ch;
                    if (leaf > j) {
                        uint32_t other = w.leaf_ids[leaf];
                        bool odyn = (w.bflags[other] & 1u) != 0;
                        if ((qdyn || odyn) && overlap(qmin, qmax, w.bmin[other], w.bmax[other])) {
                            found = true;
                            key = pair_key(body, other);
                        }
                    }
                }
            }
            wave_append(w, found, key);
        }
    }
}

size_t broadphase_sort_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 30);
    return bytes;
}

void launch_broadphase(const DeviceWorld& w, void* temp, size_t temp_bytes, hipStream_t st) {
    uint32_t n = w.nbodies;
    if (n < 2) return;
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, st, w);
    size_t bytes = temp_bytes;
    hipcub::DeviceRadixSort::SortPairs(temp, bytes, w.morton, w.morton_sorted, w.leaf_ids_tmp, w.leaf_ids, (int)n, 0,
                                       30, st);
    hipLaunchKernelGGL(k_lbvh_build, dim3((n - 1 + 255) / 256), dim3(256), 0, st, w);
    launch_fill_u32(w.node_visit, 0u, n - 1, st);
    hipLaunchKernelGGL(k_lbvh_refit, dim3((n + 255) / 256), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_traverse, dim3((n + 255) / 256), dim3(256), 0, st, w);
}

}  // namespace bepu
