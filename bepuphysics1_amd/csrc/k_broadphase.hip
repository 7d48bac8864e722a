// k_broadphase.hip -- broadphase overlap finding: Morton codes of AABB
// centres, radix sort, an implicit balanced BVH over the Morton order and a
// stack traversal per leaf whose hits are compacted with wavefront ballots.
//
// Replaces DynamicHierarchy.UpdateSingleThreaded (DynamicHierarchy.cs:194-214).
// The pair SET is what the reference produces: leaf boxes are the bodies'
// AABBs (DynamicHierarchyNode.cs:766-769), leaf-leaf overlap is the closed
// interval test of BoundingBox.Intersects (BoundingBox.cs:58-69) and pairs
// pass the default collision rule (kinematic-kinematic is NoBroadPhase,
// CollisionRules.cs:175-178; dynamic bodies are always active here).
//
// Tree: level 0 holds the leaves in Morton order; node i of level L covers
// leaves [i * 2^L, (i + 1) * 2^L).  Levels 1..8 are reduced inside the
// workgroup that loaded the leaves (LDS), the few upper levels by one
// workgroup, so the build needs no cross-workgroup synchronisation.  Each
// node carries the largest leaf size below it and whether any leaf below is
// dynamic.  A pair is reported by the leaf with the smaller (size, index);
// subtrees that cannot hold a leaf ranked above the query are pruned, so a
// large static body (a ground) does not walk the tree.
#include <hipcub/hipcub.hpp>

#include "world.h"

namespace bepu {

constexpr int kTreeBlock = 256;   // leaves per low-level build block (levels 1..8)
constexpr int kTreeLow = 8;
constexpr int kStack = 32;

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

// Node encoding: min.xyz + w = largest leaf size below; max.xyz + w = 1 if a
// dynamic leaf is below.  Leaf size = sum of the AABB's edge lengths.
__device__ __forceinline__ void merge(float4& amin, float4& amax, float4 bmin, float4 bmax) {
    amin = make_float4(fminf(amin.x, bmin.x), fminf(amin.y, bmin.y), fminf(amin.z, bmin.z), fmaxf(amin.w, bmin.w));
    amax = make_float4(fmaxf(amax.x, bmax.x), fmaxf(amax.y, bmax.y), fmaxf(amax.z, bmax.z), fmaxf(amax.w, bmax.w));
}

__global__ void __launch_bounds__(kTreeBlock) k_tree_low(DeviceWorld w) {
    __shared__ float4 smin[kTreeBlock], smax[kTreeBlock];
    uint32_t n = w.nbodies;
    uint32_t t = threadIdx.x;
    uint32_t j = blockIdx.x * kTreeBlock + t;
    if (j < n) {
        uint32_t b = w.leaf_ids[j];
        float4 mn = w.bmin[b], mx = w.bmax[b];
        float ext = (mx.x - mn.x) + (mx.y - mn.y) + (mx.z - mn.z);
        float dyn = (w.bflags[b] & 1u) ? 1.0f : 0.0f;
        mn.w = ext;
        mx.w = dyn;
        w.node_min[j] = mn;
        w.node_max[j] = mx;
        smin[t] = mn;
        smax[t] = mx;
    }
    __syncthreads();
    for (int L = 1; L <= kTreeLow && L < (int)w.tree_levels; ++L) {
        uint32_t cnt_prev = w.tree_cnt[L - 1], cnt = w.tree_cnt[L];
        uint32_t per = kTreeBlock >> L;              // nodes of level L owned by this block
        uint32_t gi = blockIdx.x * per + t;          // global index at level L
        float4 mn, mx;
        bool ok = t < per && gi < cnt;
        if (ok) {
            uint32_t c0 = 2 * gi;
            mn = smin[2 * t];
            mx = smax[2 * t];
            if (c0 + 1 < cnt_prev) merge(mn, mx, smin[2 * t + 1], smax[2 * t + 1]);
        }
        __syncthreads();
        if (ok) {
            smin[t] = mn;
            smax[t] = mx;
            w.node_min[w.tree_off[L] + gi] = mn;
            w.node_max[w.tree_off[L] + gi] = mx;
        }
        __syncthreads();
    }
}

// Levels above kTreeLow: one workgroup, level by level.
__global__ void __launch_bounds__(1024) k_tree_high(DeviceWorld w) {
    for (int L = kTreeLow + 1; L < (int)w.tree_levels; ++L) {
        uint32_t cnt_prev = w.tree_cnt[L - 1], cnt = w.tree_cnt[L];
        const float4* cmin = w.node_min + w.tree_off[L - 1];
        const float4* cmax = w.node_max + w.tree_off[L - 1];
        for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
            float4 mn = cmin[2 * i], mx = cmax[2 * i];
            if (2 * i + 1 < cnt_prev) merge(mn, mx, cmin[2 * i + 1], cmax[2 * i + 1]);
            w.node_min[w.tree_off[L] + i] = mn;
            w.node_max[w.tree_off[L] + i] = mx;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ bool overlap(float4 amin, float4 amax, float4 bmin, float4 bmax) {
    if (bmin.x > amax.x || bmin.y > amax.y || bmin.z > amax.z) return false;
    if (amin.x > bmax.x || amin.y > bmax.y || amin.z > bmax.z) return false;
    return true;
}

constexpr int kPairBuf = 1024;  // pairs buffered in LDS per workgroup before one global append

// Flushes the workgroup's buffered pairs with a single global atomic.
__device__ __forceinline__ void flush_pairs(DeviceWorld& w, uint64_t* buf, uint32_t n, uint32_t* base_slot) {
    if (threadIdx.x == 0) *base_slot = n ? atomicAdd(&w.counters[CTR_PAIRS], n) : 0;
    __syncthreads();
    uint32_t base = *base_slot;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t idx = base + i;
        if (idx < w.pair_cap) w.pair_keys[idx] = buf[i];
        else w.counters[CTR_OVERFLOW] = 1;
    }
}

// One thread per sorted leaf j; reports overlapping leaves ranked above it.
// Hits are appended to an LDS buffer (wave ballot + one LDS atomic per wave)
// and written out with one global atomic per workgroup.
__global__ void __launch_bounds__(256) k_traverse(DeviceWorld w) {
    __shared__ uint32_t stack_mem[kStack][256];
    __shared__ uint64_t buf[kPairBuf];
    __shared__ uint32_t buf_n, base_slot;
    uint32_t n = w.nbodies;
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) buf_n = 0;
    __syncthreads();
    bool alive = j < n;
    float4 qmin = alive ? w.node_min[j] : make_float4(0, 0, 0, 0);
    float4 qmax = alive ? w.node_max[j] : make_float4(0, 0, 0, 0);
    uint32_t body = alive ? w.leaf_ids[j] : 0;
    float qext = qmin.w;
    bool qdyn = qmax.w != 0.0f;
    int top = (int)w.tree_levels - 1;
    int sp = 0;
    if (alive && n > 1) stack_mem[sp++][threadIdx.x] = ((uint32_t)top << 27) | 0u;
    int lane = __lane_id();
    // The loop runs until every wave of the block is done so that buffer
    // flushes (block barriers) are reached by all threads.
    while (true) {
        bool busy = sp > 0;
        int L = -1;
        uint32_t idx = 0;
        if (busy) {
            uint32_t e = stack_mem[--sp][threadIdx.x];
            L = (int)(e >> 27);
            idx = e & 0x7ffffffu;
        }
        for (int side = 0; side < 2; ++side) {
            bool found = false;
            uint64_t key = 0;
            if (L > 0) {
                int cl = L - 1;
                uint32_t ci = 2 * idx + side;
                if (ci < w.tree_cnt[cl]) {
                    float4 cmin = w.node_min[w.tree_off[cl] + ci];
                    float4 cmax = w.node_max[w.tree_off[cl] + ci];
                    if (cl == 0) {
                        float cext = cmin.w;
                        bool mine = cext > qext || (cext == qext && ci > j);
                        bool cdyn = cmax.w != 0.0f;
                        if (mine && (qdyn || cdyn) && overlap(qmin, qmax, cmin, cmax)) {
                            found = true;
                            key = pair_key(body, w.leaf_ids[ci]);
                        }
                    } else {
                        uint32_t hi_leaf = min(n, (ci + 1) << cl) - 1;
                        bool ranked = cmin.w > qext || (cmin.w == qext && hi_leaf > j);
                        if (ranked && (qdyn || cmax.w != 0.0f) && overlap(qmin, qmax, cmin, cmax)) {
                            if (sp < kStack) stack_mem[sp++][threadIdx.x] = ((uint32_t)cl << 27) | ci;
                            else w.counters[CTR_OVERFLOW] = 2;
                        }
                    }
                }
            }
            uint64_t mask = __ballot(found);
            if (mask) {
                uint32_t wbase = 0;
                int leader = __ffsll((unsigned long long)mask) - 1;
                if (lane == leader) wbase = atomicAdd(&buf_n, (uint32_t)__popcll(mask));
                wbase = __shfl(wbase, leader);
                if (found) {
                    uint32_t at = wbase + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                    if (at < kPairBuf) buf[at] = key;
                    else {  // buffer full: this pair goes straight out
                        uint32_t idx2 = atomicAdd(&w.counters[CTR_PAIRS], 1u);
                        if (idx2 < w.pair_cap) w.pair_keys[idx2] = key;
                        else w.counters[CTR_OVERFLOW] = 1;
                    }
                }
            }
        }
        // Block-level continuation test and buffer flush when nearly full.
        if (__syncthreads_or(sp > 0) == 0) break;
        if (buf_n >= kPairBuf - 2 * 256) {
            uint32_t cnt = min(buf_n, (uint32_t)kPairBuf);
            flush_pairs(w, buf, cnt, &base_slot);
            __syncthreads();
            if (threadIdx.x == 0) buf_n = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    flush_pairs(w, buf, min(buf_n, (uint32_t)kPairBuf), &base_slot);
}

size_t broadphase_sort_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 30);
    return bytes;
}

void tree_layout(uint32_t n, uint32_t* levels, uint32_t* off, uint32_t* cnt) {
    uint32_t L = 0, o = 0, c = n;
    while (true) {
        off[L] = o;
        cnt[L] = c;
        o += c;
        ++L;
        if (c <= 1 || L >= kMaxTreeLevels) break;
        c = (c + 1) / 2;
    }
    *levels = L;
}

void launch_broadphase(const DeviceWorld& w, void* temp, size_t temp_bytes, hipStream_t st) {
    uint32_t n = w.nbodies;
    if (n < 2) return;
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, st, w);
    size_t bytes = temp_bytes;
    (void)hipcub::DeviceRadixSort::SortPairs(temp, bytes, w.morton, w.morton_sorted, w.leaf_ids_tmp, w.leaf_ids,
                                             (int)n, 0, 30, st);
    hipLaunchKernelGGL(k_tree_low, dim3((n + kTreeBlock - 1) / kTreeBlock), dim3(kTreeBlock), 0, st, w);
    if (w.tree_levels > kTreeLow + 1) hipLaunchKernelGGL(k_tree_high, dim3(1), dim3(1024), 0, st, w);
    hipLaunchKernelGGL(k_traverse, dim3((n + 255) / 256), dim3(256), 0, st, w);
}

}  // namespace bepu
