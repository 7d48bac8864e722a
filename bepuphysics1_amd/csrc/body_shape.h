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

}  // namespace bepu
