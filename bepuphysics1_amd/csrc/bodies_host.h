// bodies_host.h -- host-side entity construction (one-time per body):
// shape volume distributions, mass properties and the inverse inertia, as
// Entity/ConvexShape construction computes them in the reference.
#pragma once
#include <cmath>
#include <cstring>

#include "../../include/bepu_gpu.h"
#include "dmath.h"

namespace bepu {

// BoxShape.ComputeDescription (BoxShape.cs:118-138): volume distribution of a box.
inline m3 box_volume_distribution(float width, float height, float length) {
    float w2 = width * width, h2 = height * height, l2 = length * length;
    const float inv12 = 1 / 12.0f;
    return m3{(h2 + l2) * inv12, 0, 0, 0, (w2 + l2) * inv12, 0, 0, 0, (w2 + h2) * inv12};
}

// SphereShape.ComputeDescription (SphereShape.cs:52-67)
inline m3 sphere_volume_distribution(float radius) {
    float d = ((2.0f / 5.0f) * radius * radius);
    return m3{d, 0, 0, 0, d, 0, 0, 0, d};
}

// Matrix3x3.AdaptiveDeterminant / AdaptiveInvert (Matrix3x3.cs:550-664, 1244-1290).
// The corner minor is M11*M33 - M13*M12 in the reference (Matrix3x3.cs:1269); kept.
inline m3 adaptive_invert(const m3& m) {
    float det = m.m11 * m.m22 * m.m33 + m.m12 * m.m23 * m.m31 + m.m13 * m.m21 * m.m32 -
                m.m31 * m.m22 * m.m13 - m.m32 * m.m23 * m.m11 - m.m33 * m.m21 * m.m12;
    int sub;
    if (det != 0) sub = 0;
    else if ((det = m.m11 * m.m22 - m.m12 * m.m21) != 0) sub = 1;
    else if ((det = m.m22 * m.m33 - m.m23 * m.m32) != 0) sub = 2;
    else if ((det = m.m11 * m.m33 - m.m13 * m.m12) != 0) sub = 3;
    else if (m.m11 != 0) { det = m.m11; sub = 4; }
    else if (m.m22 != 0) { det = m.m22; sub = 5; }
    else if (m.m33 != 0) { det = m.m33; sub = 6; }
    else { det = 0; sub = -1; }
    float di = 1 / det;
    m3 r{0, 0, 0, 0, 0, 0, 0, 0, 0};
    switch (sub) {
    case 0:
        r.m11 = (m.m22 * m.m33 - m.m23 * m.m32) * di;
        r.m12 = (m.m13 * m.m32 - m.m33 * m.m12) * di;
        r.m13 = (m.m12 * m.m23 - m.m22 * m.m13) * di;
        r.m21 = (m.m23 * m.m31 - m.m21 * m.m33) * di;
        r.m22 = (m.m11 * m.m33 - m.m13 * m.m31) * di;
        r.m23 = (m.m13 * m.m21 - m.m11 * m.m23) * di;
        r.m31 = (m.m21 * m.m32 - m.m22 * m.m31) * di;
        r.m32 = (m.m12 * m.m31 - m.m11 * m.m32) * di;
        r.m33 = (m.m11 * m.m22 - m.m12 * m.m21) * di;
        break;
    case 1: r.m11 = m.m22 * di; r.m12 = -m.m12 * di; r.m21 = -m.m21 * di; r.m22 = m.m11 * di; break;
    case 2: r.m22 = m.m33 * di; r.m23 = -m.m23 * di; r.m32 = -m.m32 * di; r.m33 = m.m22 * di; break;
    case 3: r.m11 = m.m33 * di; r.m13 = -m.m13 * di; r.m31 = -m.m31 * di; r.m33 = m.m11 * di; break;
    case 4: r.m11 = 1 / m.m11; break;
    case 5: r.m22 = 1 / m.m22; break;
    case 6: r.m33 = 1 / m.m33; break;
    default: break;
    }
    return r;
}

// CapsuleShape.ComputeDescription (CapsuleShape.cs:76-92): a cylindrical
// approximation of the capsule's volume distribution.
inline m3 capsule_volume_distribution(float length, float radius) {
    float effective = length + radius / 2;
    float diag = (.0833333333f * effective * effective + .25f * radius * radius);
    return m3{diag, 0, 0, 0, .5f * radius * radius, 0, 0, 0, diag};
}

// Device shape records: capsule (radius, half length, 0, kind); hull (margin,
// first vertex, vertex count as uint bits, kind).
inline float4 capsule_shape(float length, float radius) {
    return make_float4(radius, length * 0.5f, 0, (float)BEPU_SHAPE_CAPSULE);
}
inline float4 hull_shape(float margin, uint32_t first, uint32_t count) {
    float f, c;
    std::memcpy(&f, &first, 4);
    std::memcpy(&c, &count, 4);
    return make_float4(margin, f, c, (float)BEPU_SHAPE_CONVEX_HULL);
}

inline m3 scale3(const m3& m, float s) {
    return m3{m.m11 * s, m.m12 * s, m.m13 * s, m.m21 * s, m.m22 * s, m.m23 * s, m.m31 * s, m.m32 * s, m.m33 * s};
}

}  // namespace bepu
