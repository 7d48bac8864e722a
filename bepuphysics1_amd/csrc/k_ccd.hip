// k_ccd.hip -- continuous position update (ContinuousPositionUpdater,
// PositionUpdating/ContinuousPositionUpdater.cs:134-154) after the
// integration kernel turned every active body and moved the discrete ones:
//
//   k_toi       one thread per pair slot: a pair with a continuous body and
//               an active one recomputes its time of impact
//               (ConvexPairHandler.UpdateTimeOfImpact, ConvexPairHandler.cs:
//               37-111) with the sphere casts of cast.h; every pair of a
//               moving passive or continuous body takes part in that body's
//               smallest time of impact (an atomic minimum over the float
//               bits, which order as the values for times in (0, 1]);
//   k_ccd_move  one thread per body: Entity.UpdatePositionContinuously
//               (EntityBase.cs:1093-1121), the linear motion scaled by that
//               smallest time of impact, which then resets to 1.
//
// Launched only in worlds where a body was made passive or continuous.
#include "body_shape.h"
#include "cast.h"
#include "mesh.h"

namespace bepu {

namespace {

constexpr uint32_t kOneBits = 0x3f800000u;  // 1.0f

// The body moves this step: the integration kernel's test (an active body).
__device__ __forceinline__ bool moves(const DeviceWorld& w, uint32_t i, uint32_t f) {
    if (!(f & BF_PRESENT) || (f & BF_ASLEEP)) return false;
    return (f & BF_DYNAMIC) || w.stab[i].y <= 0.0f;
}

// The collidable's world transform while the times of impact are computed: a
// moving passive or continuous body has the pose of the step's start (its
// orientation was turned, its transform not refreshed), the others their
// current pose.
__device__ __forceinline__ cvx::RT ccd_pose(const DeviceWorld& w, uint32_t i, uint32_t f, bool moving) {
    const float4 q = ((f & BF_CCD_MODE) && moving) ? w.qold[i] : w.quat[i];
    cvx::RT t;
    t.pos = f3_of(w.pos[i]);
    t.q = q4{q.x, q.y, q.z, q.w};
    return t;
}

__device__ __forceinline__ cvx::Shape pose_shape(const DeviceWorld& w, uint32_t i) {
    return shape_of(w, i, w.hull_verts + __float_as_uint(w.shape[i].y), 1);
}

// ConvexPairHandler.UpdateTimeOfImpact for the pair (a, b) in the handler's
// order: the discrete bodies already moved, so the sweep is the relative
// motion of the others over the step; each body's core sphere (its minimum
// radius times the core shape scaling) is cast at the other shape when the
// sweep is longer than that radius.
__device__ __forceinline__ float pair_toi(const DeviceWorld& w, uint32_t a, uint32_t b, uint32_t fa, uint32_t fb,
                                          bool ma, bool mb) {
    const f3 va = f3_of(w.linv[a]), vb = f3_of(w.linv[b]);
    f3 vel;
    if (!(fa & BF_CCD_MODE)) vel = vb;
    else if (!(fb & BF_CCD_MODE)) vel = -va;
    else vel = vb - va;
    vel = vel * w.s.dt;
    const float vsq = lensq3(vel);
    const float scaling = w.s.core_shape_scaling;
    const cvx::RT ta = ccd_pose(w, a, fa, ma), tb = ccd_pose(w, b, fb, mb);
    float toi = 1, t;
    const float ra = w.minrad[a] * scaling;
    if (ra * ra < vsq && cvx::ccd_sphere_cast(ta.pos, -vel, ra, pose_shape(w, b), tb, toi, scaling, t)) toi = t;
    const float rb = w.minrad[b] * scaling;
    if (rb * rb < vsq && cvx::ccd_sphere_cast(tb.pos, vel, rb, pose_shape(w, a), ta, toi, scaling, t)) toi = t;
    // Embedded from the start: no clamping (the contacts handle it).
    if (toi == 0) toi = 1;
    return toi;
}

// StaticMeshPairHandler.UpdateTimeOfImpact (StaticMeshPairHandler.cs:95-159)
// and TerrainPairHandler.UpdateTimeOfImpact (TerrainPairHandler.cs:117-176)
// for an active continuous convex a (the mesh or terrain m never moves): its
// core sphere swept along its motion against each triangle of the manifold's
// overlapped triangles (a mesh: those its box overlaps; a terrain: those of
// the query box of its manifold's last update), each triangle moved by the
// convex's position; a hit further than Toolbox.BigEpsilon (on a one-sided
// mesh only from the side the mesh keeps; on a terrain only against its up
// side) shortens the time of impact.
__device__ float mesh_toi(const DeviceWorld& w, uint32_t a, uint32_t m, uint32_t slot, float toi) {
    const f3 velocity = f3_of(w.linv[a]) * w.s.dt;
    const float vsq = lensq3(velocity);
    const float minimumRadius = w.minrad[a] * w.s.core_shape_scaling;
    toi = 1;
    if (!(minimumRadius * minimumRadius < vsq)) return toi;
    const float4 msh = w.shape[m];
    const int sided = (int)msh.y;
    const f3 pa = f3_of(w.pos[a]);
    const cvx::RT identity{f3{0, 0, 0}, q4{0, 0, 0, 1}};
    // One triangle in world space, moved by the convex's position; `up` the
    // terrain's up direction (a mesh: unused).
    auto cast = [&](cvx::TriShape& t, bool terrain, f3 up) {
        t.v0 = t.v0 - pa;
        t.v1 = t.v1 - pa;
        t.v2 = t.v2 - pa;
        t.margin = 0;
        float hitT;
        f3 hitN;
        if (cvx::ccd_sphere_cast(f3{0, 0, 0}, velocity, minimumRadius, t, identity, toi, w.s.core_shape_scaling, hitT,
                                 &hitN) &&
            hitT > cvx::kBigEpsilon) {
            if (terrain) {
                // TerrainPairHandler.UpdateTimeOfImpact (TerrainPairHandler.cs:148-170):
                // the normal AC x AB turned to the terrain's up side.
                const f3 n = cross3(t.v2 - t.v0, t.v1 - t.v0);
                float d = dot3(n, up);
                if (d < 0) d = dot3(n, hitN);
                else d = -dot3(n, hitN);
                if (d < 0) toi = hitT;
            } else if (sided != SIDED_DOUBLE) {
                const float d = dot3(cross3(t.v1 - t.v0, t.v2 - t.v0), hitN);
                if ((sided == SIDED_COUNTERCLOCKWISE && d < 0) || (sided == SIDED_CLOCKWISE && d > 0)) toi = hitT;
            } else {
                toi = hitT;
            }
        }
    };
    if (msh.w == (float)BEPU_SHAPE_TERRAIN) {
        // The triangles the manifold's last update found: its query box.
        const float4 lo = w.conv_cur[slot].sep, hi = w.conv_cur[slot].dir;
        if (lo.w == 0.0f) return toi;
        const TerrainInfo ti = terrain_info(w, __float_as_uint(msh.x));
        const f3 up{ti.m.m21, ti.m.m22, ti.m.m23};
        terrain_overlaps(w, ti, f3_of(lo), f3_of(hi), [&](int tri) {
            f3 v[3];
            terrain_world_triangle(w, ti, tri, v);
            cvx::TriShape t;
            t.v0 = v[0];
            t.v1 = v[1];
            t.v2 = v[2];
            cast(t, true, up);
        });
        return toi;
    }
    const MeshInfo mi = mesh_info(w, __float_as_uint(msh.x));
    const f3 qmin = f3_of(w.bmin[a]), qmax = f3_of(w.bmax[a]);
    const float4* nodes = w.mesh_nodes + 2 * (size_t)mi.nbase;
    auto overlaps = [&](int node) {
        const float4 lo = nodes[2 * node], hi = nodes[2 * node + 1];
        if (qmin.x > hi.x || qmin.y > hi.y || qmin.z > hi.z) return false;
        if (lo.x > qmax.x || lo.y > qmax.y || lo.z > qmax.z) return false;
        return true;
    };
    int stack[kMeshStack];
    int top = 0;
    if (mi.ntri > 0 && overlaps(0)) stack[top++] = 0;
    while (top > 0) {
        const int node = stack[--top];
        const int ca = __float_as_int(nodes[2 * node].w), cb = __float_as_int(nodes[2 * node + 1].w);
        if (ca >= 0) {
            if (overlaps(cb) && top < kMeshStack) stack[top++] = cb;
            if (overlaps(ca) && top < kMeshStack) stack[top++] = ca;
            continue;
        }
        cvx::TriShape t;
        const int32_t* ix = w.mesh_idx + mi.ibase + cb;
        t.v0 = affine_transform(f3_of(w.mesh_verts[mi.vbase + ix[0]]), mi.m, mi.t);
        t.v1 = affine_transform(f3_of(w.mesh_verts[mi.vbase + ix[1]]), mi.m, mi.t);
        t.v2 = affine_transform(f3_of(w.mesh_verts[mi.vbase + ix[2]]), mi.m, mi.t);
        cast(t, false, f3{0, 0, 0});
    }
    return toi;
}

__global__ void __launch_bounds__(256) k_toi(DeviceWorld w) {
    const uint32_t n = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const PairHot h = w.hot_cur[k];
        if (h.flags & (PF_DEAD | PF_VOID)) continue;
        const uint32_t fa = w.bflags[h.a], fb = w.bflags[h.b];
        if (!((fa | fb) & BF_CCD_MODE)) continue;  // no passive or continuous body reads it
        const bool ma = moves(w, h.a, fa), mb = moves(w, h.b, fb);
        float toi;
        if (fb & BF_MESH) {
            // A mesh pair (the convex is A): requested by an active continuous convex.
            toi = w.pairs_cur[k].toi;
            if ((fa & BF_CONTINUOUS) && ma) {
                toi = mesh_toi(w, h.a, h.b, k, toi);
                w.pairs_cur[k].toi = toi;
            }
            const uint32_t bits = __float_as_uint(toi);
            if ((fa & BF_CCD_MODE) && ma && bits < kOneBits) atomicMin(&w.toi[h.a], bits);
            continue;
        }
        // Requested by the pair's continuous bodies (EntityBase.cs:1071-1082)
        // whenever one of the two is active; otherwise the pair keeps the
        // time of impact it had.
        if (((fa | fb) & BF_CONTINUOUS) && (ma || mb)) {
            toi = pair_toi(w, h.a, h.b, fa, fb, ma, mb);
            w.pairs_cur[k].toi = toi;
        } else {
            toi = w.pairs_cur[k].toi;
        }
        const uint32_t bits = __float_as_uint(toi);
        if ((fa & BF_CCD_MODE) && ma && bits < kOneBits) atomicMin(&w.toi[h.a], bits);
        if ((fb & BF_CCD_MODE) && mb && bits < kOneBits) atomicMin(&w.toi[h.b], bits);
    }
}

__global__ void __launch_bounds__(256) k_ccd_move(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nbodies) return;
    const uint32_t f = w.bflags[i];
    if (!(f & BF_CCD_MODE) || !moves(w, i, f)) return;
    const float least = __uint_as_float(w.toi[i]);
    w.toi[i] = kOneBits;
    const f3 v = f3_of(w.linv[i]);
    const f3 inc = v * (w.s.dt * least);
    const f3 p = f3_of(w.pos[i]) + inc;
    w.pos[i] = f4_of(p, 0.0f);
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) {
        w.wfault[w.wid[i]] = 1u;
        atomicOr(&w.counters[ctr(CTR_STICKY)], STICKY_NONFINITE);
    }
}

}  // namespace

void launch_ccd(const DeviceWorld& w, hipStream_t st) {
    if (w.pair_cap) {
        const uint32_t grid = std::min<uint32_t>((w.pair_cap + 255) / 256, 2048);
        hipLaunchKernelGGL(k_toi, dim3(grid), dim3(256), 0, st, w);
    }
    if (w.nbodies) hipLaunchKernelGGL(k_ccd_move, dim3((w.nbodies + 255) / 256), dim3(256), 0, st, w);
}

}  // namespace bepu
