// k_bvh.hip -- tree broadphase: Morton codes of AABB centres, radix sort, an
// implicit balanced tree over the Morton order and a bottom-up walk per leaf
// whose hits are compacted with wavefront ballots.
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
#include "pairbuf.h"

namespace bepu {

constexpr int kTreeBlock = 256;   // leaves per low-level build block (levels 1..8)
constexpr int kTreeLow = 8;

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
        bool dyn = (w.bflags[b] & 1u) != 0;
        mn.w = ext;
        // Leaf encoding: max.w holds the body id with the dynamic flag in bit 31
        // (raw bits, never used as a float); internal nodes hold the flag as 0/1.
        mx.w = __uint_as_float(b | (dyn ? 0x80000000u : 0u));
        w.node_min[j] = mn;
        w.node_max[j] = mx;
        mx.w = dyn ? 1.0f : 0.0f;
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

constexpr int kTravBlock = 64;     // one wavefront per workgroup: no cross-wave barriers
constexpr int kDescend = 3;        // levels skipped per pop: 8 descendants tested with independent loads
constexpr int kFanout = 1 << kDescend;
constexpr int kPairBuf = 2 * kFanout * kTravBlock;
// A pop pushes at most kFanout - 1 entries net per kDescend levels: a search
// rooted at level k <= 30 needs at most 7 * 10 + 1 entries.
constexpr int kStack = 72;

// Leaf test: ranked above the query, collision rule, box overlap, same sub-world.
__device__ __forceinline__ bool leaf_hit(const DeviceWorld& w, float4 qmin, float4 qmax, bool qdyn, uint32_t qworld,
                                         uint32_t j, uint32_t ci, float4 cmin, float4 cmax) {
    uint32_t cb = __float_as_uint(cmax.w);
    bool mine = cmin.w > qmin.w || (cmin.w == qmin.w && ci > j);
    return mine && (qdyn || (cb >> 31)) && overlap(qmin, qmax, cmin, cmax) && w.wid[cb & 0x7fffffffu] == qworld;
}

__device__ __forceinline__ bool node_hit(float4 qmin, float4 qmax, bool qdyn, uint32_t j, uint32_t n, int level,
                                         uint32_t ci, float4 cmin, float4 cmax) {
    uint32_t hi_leaf = min(n, (ci + 1) << level) - 1;
    bool ranked = cmin.w > qmin.w || (cmin.w == qmin.w && hi_leaf > j);
    return ranked && (qdyn || cmax.w != 0.0f) && overlap(qmin, qmax, cmin, cmax);
}

// One thread per sorted leaf j; reports the overlapping leaves ranked above
// it (larger size, or equal size and larger index).  The walk is bottom-up:
// the sibling of every ancestor of j is tested first, then each qualifying
// sibling subtree is searched depth-first, a pop testing the (up to) 8
// descendants three levels down with independent loads.
__global__ void __launch_bounds__(kTravBlock) k_traverse(DeviceWorld w) {
    __shared__ uint32_t stack_mem[kStack][kTravBlock];
    __shared__ uint64_t buf[kPairBuf];
    __shared__ uint32_t buf_n, base_slot;
    __shared__ uint32_t s_off[kMaxTreeLevels], s_cnt[kMaxTreeLevels];
    const uint32_t n = w.nbodies;
    const int tid = threadIdx.x;
    if (tid < (int)kMaxTreeLevels) {
        s_off[tid] = w.tree_off[tid];
        s_cnt[tid] = w.tree_cnt[tid];
    }
    if (tid == 0) buf_n = 0;
    __syncthreads();
    const uint32_t j = blockIdx.x * kTravBlock + tid;
    const bool alive = j < n;
    const int top = (int)w.tree_levels - 1;
    const float4 qmin = alive ? w.node_min[j] : make_float4(0, 0, 0, 0);
    const float4 qmax = alive ? w.node_max[j] : make_float4(0, 0, 0, 0);
    const uint32_t qbits = __float_as_uint(qmax.w);
    const uint32_t body = qbits & 0x7fffffffu;
    const bool qdyn = (qbits >> 31) != 0;
    const uint32_t qworld = alive ? w.wid[body] : 0;

    // Ancestor siblings: bit k of `todo` = the sibling subtree at level k qualifies.
    uint32_t todo = 0;
    bool found = false;
    uint64_t key = 0;
    if (alive) {
        for (int k = 0; k < top; ++k) {
            uint32_t sib = (j >> k) ^ 1u;
            if (sib >= s_cnt[k]) continue;
            float4 cmin = w.node_min[s_off[k] + sib], cmax = w.node_max[s_off[k] + sib];
            if (k == 0) {
                if (leaf_hit(w, qmin, qmax, qdyn, qworld, j, sib, cmin, cmax)) {
                    found = true;
                    key = pair_key(body, __float_as_uint(cmax.w) & 0x7fffffffu);
                }
            } else if (node_hit(qmin, qmax, qdyn, j, n, k, sib, cmin, cmax)) {
                todo |= 1u << k;
            }
        }
    }
    buf_push(w, buf, &buf_n, kPairBuf, found, key);

    int sp = 0;
    while (true) {
        if (sp == 0 && todo) {
            int k = __ffs(todo) - 1;
            todo &= todo - 1;
            stack_mem[sp++][tid] = ((uint32_t)k << 27) | ((j >> k) ^ 1u);
        }
        bool busy = sp > 0;
        if (!__syncthreads_or(busy)) break;
        int L = 0;
        uint32_t idx = 0;
        if (busy) {
            uint32_t e = stack_mem[--sp][tid];
            L = (int)(e >> 27);
            idx = e & 0x7ffffffu;
        }
        const int d = min(kDescend, L);
        const int cl = L - d;
        const uint32_t cfirst = idx << d;
        // cfirst < s_cnt[cl] holds for every node of level L.
        const uint32_t lim = busy ? min(1u << d, s_cnt[cl] - cfirst) : 0u;
        const uint32_t base = s_off[cl] + cfirst;
        uint32_t fmask = 0;
        uint64_t keys[kFanout];
#pragma unroll
        for (int c = 0; c < kFanout; ++c) {
            keys[c] = 0;
            if ((uint32_t)c < lim) {
                const uint32_t ci = cfirst + (uint32_t)c;
                const float4 cmin = w.node_min[base + c], cmax = w.node_max[base + c];
                if (cl == 0) {
                    if (leaf_hit(w, qmin, qmax, qdyn, qworld, j, ci, cmin, cmax)) {
                        fmask |= 1u << c;
                        keys[c] = pair_key(body, __float_as_uint(cmax.w) & 0x7fffffffu);
                    }
                } else if (node_hit(qmin, qmax, qdyn, j, n, cl, ci, cmin, cmax)) {
                    if (sp < kStack) stack_mem[sp++][tid] = ((uint32_t)cl << 27) | ci;
                    else w.counters[ctr(CTR_OVERFLOW)] = 2;
                }
            }
        }
        if (__ballot(fmask != 0)) {
#pragma unroll
            for (int c = 0; c < kFanout; ++c) buf_push(w, buf, &buf_n, kPairBuf, (fmask >> c) & 1u, keys[c]);
        }
        // At most kFanout * 64 pairs join per iteration: flush before the buffer can overflow.
        __syncthreads();
        if (buf_n > (uint32_t)(kPairBuf - kFanout * kTravBlock)) buf_flush<kTravBlock>(w, buf, &buf_n, kPairBuf, &base_slot);
    }
    buf_flush<kTravBlock>(w, buf, &buf_n, kPairBuf, &base_slot);
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

// Tree build and walk over leaves already sorted by Morton code.
void launch_bvh_after_sort(const DeviceWorld& w, hipStream_t st) {
    uint32_t n = w.nbodies;
    hipLaunchKernelGGL(k_tree_low, dim3((n + kTreeBlock - 1) / kTreeBlock), dim3(kTreeBlock), 0, st, w);
    if (w.tree_levels > kTreeLow + 1) hipLaunchKernelGGL(k_tree_high, dim3(1), dim3(1024), 0, st, w);
    hipLaunchKernelGGL(k_traverse, dim3((n + kTravBlock - 1) / kTravBlock), dim3(kTravBlock), 0, st, w);
}

void launch_morton(const DeviceWorld& w, hipStream_t st) {
    uint32_t n = w.nbodies;
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, st, w);
}

}  // namespace bepu
