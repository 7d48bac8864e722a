// k_convex.hip -- narrowphase of general convex pairs (any pair with a hull
// or a capsule): GeneralConvexContactManifold.Update
// (CollisionTests/Manifolds/GeneralConvexContactManifold.cs:71-126) on the
// pairs the main narrowphase kernel queued.  One thread per pair; the pair
// tester's iterative GJK / MPR (convex.h) diverges per pair, so it runs in a
// kernel of its own rather than inflating the closed-form pairs' kernel.
#include "body_shape.h"
#include "convex.h"
#include "manifold.h"
#include "wave.h"

namespace bepu {

namespace {

// ContactReducer.ReduceContacts for four contacts and one candidate
// (CollisionTests/ContactReducer.cs:197-318): the list index to remove (or -1).
__device__ __forceinline__ int reduce_contacts(const PairState& s, f3 cpos, float cdepth, bool* add) {
    f3 P[4];
    float D[4];
    for (int i = 0; i < 4; ++i) {
        float4 c = s.cpos[s.clist[i]];
        P[i] = f3_of(c);
        D[i] = c.w;
    }
    float maxDepth = -FLT_MAX;
    int deepest = -1;
    for (int i = 0; i < 4; i++)
        if (D[i] > maxDepth) {
            deepest = i;
            maxDepth = D[i];
        }
    if (cdepth > maxDepth) deepest = 4;
    f3 deepPos = deepest < 4 ? P[deepest] : cpos;
    float furthestDist = 0;
    int furthest = -1;
    for (int i = 0; i < 4; i++) {
        float d = distsq3(P[i], deepPos);
        if (d > furthestDist) {
            furthestDist = d;
            furthest = i;
        }
    }
    if (distsq3(cpos, deepPos) > furthestDist) furthest = 4;
    f3 farPos = furthest < 4 ? P[furthest] : cpos;
    f3 yAxis = cross3(deepPos - farPos, f3_of(s.cnrm[s.clist[0]]));
    float minY = FLT_MAX, maxY = -FLT_MAX;
    int iminY = -1, imaxY = -1;
    for (int i = 0; i < 4; i++) {
        float d = dot3(P[i], yAxis);
        if (d < minY) {
            iminY = i;
            minY = d;
        }
        if (d > maxY) {
            imaxY = i;
            maxY = d;
        }
    }
    float d = dot3(cpos, yAxis);
    if (d < minY) iminY = 4;
    if (d > maxY) imaxY = 4;
    if (deepest == 4 || furthest == 4 || iminY == 4 || imaxY == 4) {
        *add = true;
        for (int i = 0; i < 4; i++)
            if (!(i == deepest || i == furthest || i == iminY || i == imaxY)) return i;
        return -1;
    }
    *add = false;
    return -1;
}

// GeneralConvexContactManifold.Add: the contact and its supplement data
// (GeneralConvexContactManifold.cs:128-137).
__device__ __forceinline__ void add_general(PairState& s, ConvexState& e, f3 pos, f3 nrm, float depth, int id, const cvx::RT& ta,
                            const cvx::RT& tb) {
    int slot = 0;
    while (s.flags & (16u << slot)) ++slot;
    e.offa[slot] = f4_of(cvx::rt_transform_by_inverse(pos, ta), depth);
    e.offb[slot] = f4_of(cvx::rt_transform_by_inverse(pos, tb), 0.0f);
    add_contact_inplace(s, pos, nrm, depth, id);
}

// GeneralConvexContactManifold.Update (GeneralConvexContactManifold.cs:71-126) with
// ContactRefresher.ContactRefresh (CollisionTests/ContactRefresher.cs:18-69) and
// IsContactUnique (GeneralConvexContactManifold.cs:145-166).
__device__ __forceinline__ void update_general(const DeviceWorld& w, const cvx::Shape& sa, const cvx::Shape& sb,
                                               PairState& s, ConvexState& e) {
    float4 PA = w.pos[s.a], QA = w.quat[s.a], PB = w.pos[s.b], QB = w.quat[s.b];
    cvx::RT ta{f3_of(PA), q4{QA.x, QA.y, QA.z, QA.w}};
    cvx::RT tb{f3_of(PB), q4{QB.x, QB.y, QB.z, QB.w}};
    int rem[4], nrem = 0;
    for (int k = 0; k < s.ccount; k++) {
        int slot = s.clist[k];
        f3 na = cvx::rt_transform(f3_of(e.offa[slot]), ta);
        f3 nb = cvx::rt_transform(f3_of(e.offb[slot]), tb);
        f3 ab = nb - na;
        f3 n = f3_of(s.cnrm[slot]);
        float d = dot3(ab, n);
        f3 t = ab - n * d;
        d = lensq3(t);
        if (d > w.s.contact_invalidation_length_sq) {
            rem[nrem++] = k;
        } else {
            d = dot3(ab, n);
            float depth = e.offa[slot].w - d;
            s.cpos[slot].w = depth;
            if (depth < -w.s.max_contact_distance) rem[nrem++] = k;
            else s.cpos[slot] = f4_of((nb + na) * .5f, depth);
        }
    }
    for (int i = nrem - 1; i >= 0; i--) remove_contact_inplace(s, rem[i]);
    cvx::TesterState st{__float_as_int(e.sep.w), f3_of(e.sep), f3_of(e.dir)};
    // Keep the state opaque to the optimiser: a state known on one path (a new
    // pair starts Separated) would otherwise get its own clone of the whole
    // pair test, doubling the code and its register pressure.
    asm volatile("" : "+v"(st.state));
    f3 pos, nrm;
    float depth;
    int id;
    if (cvx::candidate(sa, sb, ta, tb, f3_of(w.linv[s.a]), f3_of(w.linv[s.b]), st, &pos, &nrm, &depth, &id)) {
        for (int i = 0; i < s.ccount; ++i)
            if (dot3(f3_of(s.cnrm[s.clist[i]]), nrm) < 0) {
                remove_contact_inplace(s, i);
                break;
            }
        bool unique = true;
        for (int i = 0; i < s.ccount; i++) {
            int slot = s.clist[i];
            if (distsq3(f3_of(s.cpos[slot]), pos) < w.s.contact_min_separation_sq) {
                s.cnrm[slot] = f4_of(nrm, s.cnrm[slot].w);
                s.cpos[slot] = f4_of(pos, depth);
                e.offa[slot] = f4_of(cvx::rt_transform_by_inverse(pos, ta), depth);
                e.offb[slot] = f4_of(cvx::rt_transform_by_inverse(pos, tb), 0.0f);
                unique = false;
                break;
            }
        }
        if (unique) {
            if (s.ccount == 4) {
                bool add;
                int r = reduce_contacts(s, pos, depth, &add);
                if (r >= 0) remove_contact_inplace(s, r);
                if (add) add_general(s, e, pos, nrm, depth, id, ta, tb);
            } else {
                add_general(s, e, pos, nrm, depth, id, ta, tb);
            }
        }
    } else {
        for (int i = s.ccount - 1; i >= 0; i--) remove_contact_inplace(s, i);
    }
    e.sep = f4_of(st.separating_axis, __int_as_float(st.state));
    e.dir = f4_of(st.local_direction, 0.0f);
    s.tester = st.state;
}

constexpr int kConvexBlock = 64;
constexpr uint32_t kConvexMaxGrid = 8192;  // one pair per thread up to 524,288 queued pairs

// Copies body i's hull vertices (if it is a hull) to dst[k * kBlock]: a
// thread's vertices lie one lane apart, so a wavefront's LDS reads of vertex
// k are consecutive 16-byte words (no bank conflicts).
template <int kBlock>
__device__ __forceinline__ void stage_hull(const DeviceWorld& w, uint32_t i, float4* dst) {
    float4 sh = w.shape[i];
    if (sh.w != (float)BEPU_SHAPE_CONVEX_HULL) return;
    const float4* src = w.hull_verts + __float_as_uint(sh.y);
    const int n = (int)__float_as_uint(sh.z);
    constexpr int kChunk = 8;
    for (int k0 = 0; k0 < n; k0 += kChunk) {
        float4 v[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) v[k] = src[min(k0 + k, n - 1)];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (k0 + k < n) dst[(k0 + k) * kBlock] = v[k];
    }
}

}  // namespace

// kStage: each pair's hull vertices are first copied to this thread's columns
// of an LDS block, so that every support scan of its GJK / MPR iterations
// reads LDS instead of making two dependent round trips to L2 (at most
// kConvexStageVerts vertices per hull; the host checks).
//
// `choose` 0: run; 1 (the staging variant) or 2 (the global one): run only
// when this step's pairs favour it.  Copying the hulls pays off when few of
// the pairs are Shallow (resting contacts, a short closest-point test each):
// in a settling batch the Separated and Deep pairs of the falling phase gain
// from it, the settled piles' Shallow pairs lose.  Both variants compute the
// same bits; both launches read the same counters, so exactly one runs.
template <bool kStage, int kBlock>
__global__ void __launch_bounds__(kBlock) k_convex(DeviceWorld w, uint32_t choose) {
    __shared__ float4 stage[kStage ? 2 * kConvexStageVerts * kBlock : 1];
    // The state regions of the work list, costliest first: Deep pairs (MPR
    // with surface casts and refinement), then Shallow (closest-point GJK),
    // then Separated (boolean GJK).  Workgroups start in index order, so the
    // long ones start first and the short ones fill in behind them.
    // Six lists: per state, the pairs of two hulls (from the region's front)
    // then the others (from its back).
    uint32_t cnt[6], beg[7];
    cnt[0] = w.counters[ctr(CTR_CONVEX_DEEP)];
    cnt[1] = w.counters[ctr(CTR_CONVEX_OTHER0 + 2)];
    cnt[2] = w.counters[ctr(CTR_CONVEX_SHALLOW)];
    cnt[3] = w.counters[ctr(CTR_CONVEX_OTHER0 + 1)];
    cnt[4] = w.counters[ctr(CTR_CONVEX)];
    cnt[5] = w.counters[ctr(CTR_CONVEX_OTHER0)];
    beg[0] = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) beg[q + 1] = beg[q] + cnt[q];
    const uint32_t n = beg[6];
    if (blockIdx.x * kBlock >= n) return;  // uniform per workgroup
    if (choose) {
        const uint32_t shallow = cnt[2] + cnt[3];
        const bool staging_pays = shallow < n - shallow;
        if ((choose == 1) != staging_pays) return;  // uniform per launch
    }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        int q = 0;
#pragma unroll
        for (int k = 1; k < 6; ++k) q += i >= beg[k];
        const uint32_t state = 2 - (uint32_t)(q >> 1), k = i - beg[q];
        const uint32_t at = (q & 1) ? w.pair_cap - 1 - k : k;
        uint2 wk = w.conv_work[(size_t)state * w.pair_cap + at];
        // The pair's state is updated in place in global memory: its contact
        // slots are indexed at run time, which would put a private copy in
        // scratch memory.
        PairState& s = w.pairs_cur[wk.x];
        ConvexState& e = w.conv_cur[wk.x];
        // k_pair_prep carried last step's state over; a new pair starts from
        // zero.
        if (wk.y == ~0u) {
            float4 z = make_float4(0, 0, 0, 0);
            e.sep = z;
            e.dir = z;
            e.offa[0] = z; e.offa[1] = z; e.offa[2] = z; e.offa[3] = z;
            e.offb[0] = z; e.offb[1] = z; e.offb[2] = z; e.offb[3] = z;
        }
        cvx::Shape sa, sb;
        if (kStage) {
            float4* ca = stage + threadIdx.x;
            float4* cb = stage + kConvexStageVerts * kBlock + threadIdx.x;
            stage_hull<kBlock>(w, s.a, ca);
            stage_hull<kBlock>(w, s.b, cb);
            sa = shape_of(w, s.a, ca, kBlock);
            sb = shape_of(w, s.b, cb, kBlock);
        } else {
            float4 ha = w.shape[s.a], hb = w.shape[s.b];
            sa = shape_of(w, s.a, w.hull_verts + __float_as_uint(ha.y), 1);
            sb = shape_of(w, s.b, w.hull_verts + __float_as_uint(hb.y), 1);
        }
        update_general(w, sa, sb, s, e);
        w.hot_cur[wk.x] = hot_of(s);
        // A constraint joins the islands of its two bodies when both are dynamic.
        if (s.ccount > 0 && (w.bflags[s.a] & 1u) && (w.bflags[s.b] & 1u)) uf_union(w.uf_parent, s.a, s.b);
    }
}

void launch_convex(const DeviceWorld& w, uint32_t lds_mode, hipStream_t st) {
    if (!w.conv_work || w.pair_cap == 0) return;
    uint32_t grid = std::min<uint32_t>(kConvexMaxGrid, (w.pair_cap + kConvexBlock - 1) / kConvexBlock);
    const bool can_stage = w.hull_max_verts <= kConvexStageVerts;
    if (can_stage && lds_mode == 2) {
        hipLaunchKernelGGL((k_convex<true, kConvexBlock>), dim3(grid), dim3(kConvexBlock), 0, st, w, 1u);
        hipLaunchKernelGGL((k_convex<false, kConvexBlock>), dim3(grid), dim3(kConvexBlock), 0, st, w, 2u);
    } else if (can_stage && lds_mode == 1) {
        hipLaunchKernelGGL((k_convex<true, kConvexBlock>), dim3(grid), dim3(kConvexBlock), 0, st, w, 0u);
    } else {
        hipLaunchKernelGGL((k_convex<false, kConvexBlock>), dim3(grid), dim3(kConvexBlock), 0, st, w, 0u);
    }
}

}  // namespace bepu
