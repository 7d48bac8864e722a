// body_shape.h -- the convex shape view of a body, for the device passes
// that run the general convex routines (k_convex, k_ccd).
#pragma once
#include "convex.h"
#include "world.h"

namespace bepu {

// The shape of body i.  A hull's vertices are read from `verts` (the device
// pool, or this pair's copy in LDS) at stride `vstride`.
__device__ __forceinline__ cvx::Shape shape_of(const DeviceWorld& w, uint32_t i, const float4* verts, int vstride) {
    float4 sh = w.shape[i];
    cvx::Shape s;
    s.kind = (int)sh.w;
    s.half = f3{0, 0, 0};
    s.half_length = 0;
    s.verts = nullptr;
    s.nverts = 0;
    s.vstride = vstride;
    if (s.kind == BEPU_SHAPE_BOX) {
        s.half = f3_of(sh);
        s.margin = w.s.default_margin;  // BoxShape's margin: CollisionDetectionSettings.DefaultMargin
    } else if (s.kind == BEPU_SHAPE_CAPSULE) {
        s.margin = sh.x;
        s.half_length = sh.y;
    } else if (s.kind == BEPU_SHAPE_CONVEX_HULL) {
        s.margin = sh.x;
        s.verts = verts;
        s.nverts = (int)__float_as_uint(sh.z);
    } else {
        s.margin = sh.x;  // sphere: the radius
    }
    return s;
}

// Shape.GetBoundingBox of a body's shape at pose (P, q), without the
// velocity expansion: BoxShape.cs:149-174, SphereShape.cs:75-86,
// CapsuleShape.cs:95-137, ConvexHullShape.cs:206-286; a static mesh's box is
// its tree's root box (StaticMesh.UpdateBoundingBox, StaticMesh.cs:189-192), a
// terrain's the box its shape was made with.
__device__ __forceinline__ void shape_aabb(const DeviceWorld& w, float4 sh, f3 P, q4 q, f3* mn, f3* mx) {
    if (sh.w == (float)BEPU_SHAPE_BOX) {
        m3 o = mat_from_quat(q);
        f3 right{sign_ref(o.m11) * sh.x, sign_ref(o.m21) * sh.y, sign_ref(o.m31) * sh.z};
        f3 up{sign_ref(o.m12) * sh.x, sign_ref(o.m22) * sh.y, sign_ref(o.m32) * sh.z};
        f3 back{sign_ref(o.m13) * sh.x, sign_ref(o.m23) * sh.y, sign_ref(o.m33) * sh.z};
        f3 off;
        off.x = right.x * o.m11 + right.y * o.m21 + right.z * o.m31;
        off.y = up.x * o.m12 + up.y * o.m22 + up.z * o.m32;
        off.z = back.x * o.m13 + back.y * o.m23 + back.z * o.m33;
        f3 p = P;
        *mx = p + off;
        *mn = p - off;
    } else if (sh.w == (float)BEPU_SHAPE_CAPSULE) {
        // CapsuleShape.GetBoundingBox (CapsuleShape.cs:95-137) with Quaternion.TransformY
        // (Quaternion.cs:586-608): the segment's upper end, grown by the radius.
        float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
        float xx2 = q.x * x2, xy2 = q.x * y2, yz2 = q.y * z2, zz2 = q.z * z2;
        float wx2 = q.w * x2, wz2 = q.w * z2;
        float hl = sh.y, r = sh.x;
        f3 up{hl * (xy2 - wz2), hl * (1.f - xx2 - zz2), hl * (yz2 + wx2)};
        f3 lo, hi;
        if (up.x > 0) { hi.x = up.x + r; lo.x = -up.x - r; } else { hi.x = -up.x + r; lo.x = up.x - r; }
        if (up.y > 0) { hi.y = up.y + r; lo.y = -up.y - r; } else { hi.y = -up.y + r; lo.y = up.y - r; }
        if (up.z > 0) { hi.z = up.z + r; lo.z = -up.z - r; } else { hi.z = -up.z + r; lo.z = up.z - r; }
        f3 p = P;
        *mn = p + lo;
        *mx = p + hi;
    } else if (sh.w == (float)BEPU_SHAPE_CONVEX_HULL) {
        // ConvexHullShape.GetBoundingBox (ConvexHullShape.cs:206-286): extreme
        // vertices along the rotated axes (first extreme kept on ties), then
        // EntityShape.TransformLocalExtremePoints (EntityShape.cs:58-63).
        m3 o = mat_from_quat(q);
        const float4* hv = w.hull_verts + __float_as_uint(sh.y);
        uint32_t nv = __float_as_uint(sh.z);
        f3 right{o.m11, o.m21, o.m31}, upv{o.m12, o.m22, o.m32}, back{o.m13, o.m23, o.m33};
        f3 v0 = f3_of(hv[0]);
        float maxX = dot3(v0, right), minX = maxX;
        float maxY = dot3(v0, upv), minY = maxY;
        float maxZ = dot3(v0, back), minZ = maxZ;
        uint32_t iminX = 0, imaxX = 0, iminY = 0, imaxY = 0, iminZ = 0, imaxZ = 0;
        for (uint32_t k = 1; k < nv; ++k) {
            f3 v = f3_of(hv[k]);
            float d = dot3(v, right);
            if (d < minX) { minX = d; iminX = k; } else if (d > maxX) { maxX = d; imaxX = k; }
            d = dot3(v, upv);
            if (d < minY) { minY = d; iminY = k; } else if (d > maxY) { maxY = d; imaxY = k; }
            d = dot3(v, back);
            if (d < minZ) { minZ = d; iminZ = k; } else if (d > maxZ) { maxZ = d; imaxZ = k; }
        }
        f3 a = f3_of(hv[imaxX]), b = f3_of(hv[imaxY]), c = f3_of(hv[imaxZ]);
        f3 pos{a.x * o.m11 + a.y * o.m21 + a.z * o.m31, b.x * o.m12 + b.y * o.m22 + b.z * o.m32,
               c.x * o.m13 + c.y * o.m23 + c.z * o.m33};
        a = f3_of(hv[iminX]);
        b = f3_of(hv[iminY]);
        c = f3_of(hv[iminZ]);
        f3 neg{a.x * o.m11 + a.y * o.m21 + a.z * o.m31, b.x * o.m12 + b.y * o.m22 + b.z * o.m32,
               c.x * o.m13 + c.y * o.m23 + c.z * o.m33};
        float m = sh.x;
        *mx = f3{P.x + pos.x + m, P.y + pos.y + m, P.z + pos.z + m};
        *mn = f3{P.x + neg.x - m, P.y + neg.y - m, P.z + neg.z - m};
    } else if (sh.w == (float)BEPU_SHAPE_STATIC_MESH) {
        const uint32_t nbase = __float_as_uint(w.mesh_info[4 * __float_as_uint(sh.x) + 2].w);
        *mn = f3_of(w.mesh_nodes[2 * (size_t)nbase]);
        *mx = f3_of(w.mesh_nodes[2 * (size_t)nbase + 1]);
    } else if (sh.w == (float)BEPU_SHAPE_TERRAIN) {
        // Terrain.UpdateBoundingBox (Terrain.cs:195-212): fixed, made with the shape.
        const float4* r = w.terrain_info + 6 * (size_t)__float_as_uint(sh.x);
        *mn = f3_of(r[4]);
        *mx = f3_of(r[5]);
    } else {
        float r = sh.x;
        *mn = f3{P.x - r, P.y - r, P.z - r};
        *mx = f3{P.x + r, P.y + r, P.z + r};
    }
}

}  // namespace bepu
