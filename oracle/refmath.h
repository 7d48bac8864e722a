// refmath.h -- FP32 math of BEPUutilities restated with the reference's exact
// operation order (build with -ffp-contract=off; no fast-math).
//
// TEST INFRASTRUCTURE ONLY.  This file belongs to the oracle: it is compiled
// into liboracle_refcpu.so, which only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load.  The product library never includes it.
//
// Every function cites the reference member it restates
// (paths relative to the reference repository root).
#pragma once
#include <cmath>
#include <cstdint>

namespace ref {

struct V3 { float x, y, z; };
struct Q4 { float x, y, z, w; };
// Row-major 3x3 matching BEPUutilities.Matrix3x3 field names M11..M33.
struct M3 { float m11, m12, m13, m21, m22, m23, m31, m32, m33; };

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }

// Vector3.Add / Subtract / Negate / Multiply (BEPUutilities/Vector3.cs)
inline V3 add(const V3& a, const V3& b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(const V3& a, const V3& b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 neg(const V3& a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 mul(const V3& a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
// Vector3.Dot (Vector3.cs:104-118)
inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Vector3.Cross (Vector3.cs:548-556)
inline V3 cross(const V3& a, const V3& b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Vector3.LengthSquared / Length (Vector3.cs:64-76).  (float)Math.Sqrt(double(x))
// equals the correctly rounded float square root, so sqrtf is exact here.
inline float length_sq(const V3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float length(const V3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector3.DistanceSquared (Vector3.cs:350-356)
inline float distance_sq(const V3& a, const V3& b) {
    float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
    return x * x + y * y + z * z;
}
// Vector3.Normalize (Vector3.cs:81-87): the reciprocal square root is taken in double.
inline V3 normalize(const V3& v) {
    float inv = (float)(1.0 / std::sqrt((double)(v.x * v.x + v.y * v.y + v.z * v.z)));
    return V3{v.x * inv, v.y * inv, v.z * inv};
}
// Vector3.Divide (Vector3.cs:175-181): multiply by a float reciprocal.
inline V3 divide(const V3& v, float d) {
    float inv = 1 / d;
    return V3{v.x * inv, v.y * inv, v.z * inv};
}

// Quaternion.Multiply (BEPUutilities/Quaternion.cs:65-79)
inline Q4 qmul(const Q4& a, const Q4& b) {
    float x = a.x, y = a.y, z = a.z, w = a.w;
    float bX = b.x, bY = b.y, bZ = b.z, bW = b.w;
    Q4 r;
    r.x = x * bW + bX * w + y * bZ - z * bY;
    r.y = y * bW + bY * w + z * bX - x * bZ;
    r.z = z * bW + bZ * w + x * bY - y * bX;
    r.w = w * bW - x * bX - y * bY - z * bZ;
    return r;
}
inline Q4 qadd(const Q4& a, const Q4& b) { return Q4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
// Quaternion.Normalize (Quaternion.cs:250-259, 262-269): FP64 reciprocal square root.
inline Q4 qnormalize(const Q4& q) {
    float inv = (float)(1.0 / std::sqrt((double)(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w)));
    return Q4{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
inline Q4 qconj(const Q4& q) { return Q4{-q.x, -q.y, -q.z, q.w}; }
// Quaternion.Transform (Quaternion.cs:509-535)
inline V3 qtransform(const V3& v, const Q4& r) {
    float x2 = r.x + r.x, y2 = r.y + r.y, z2 = r.z + r.z;
    float xx2 = r.x * x2, xy2 = r.x * y2, xz2 = r.x * z2;
    float yy2 = r.y * y2, yz2 = r.y * z2, zz2 = r.z * z2;
    float wx2 = r.w * x2, wy2 = r.w * y2, wz2 = r.w * z2;
    float tx = v.x * (1.0f - yy2 - zz2) + v.y * (xy2 - wz2) + v.z * (xz2 + wy2);
    float ty = v.x * (xy2 + wz2) + v.y * (1.0f - xx2 - zz2) + v.z * (yz2 - wx2);
    float tz = v.x * (xz2 - wy2) + v.y * (yz2 + wx2) + v.z * (1.0f - xx2 - yy2);
    return V3{tx, ty, tz};
}

// Matrix3x3.CreateFromQuaternion (BEPUutilities/Matrix3x3.cs:1303-1327)
inline M3 m3_from_quat(const Q4& q) {
    float qX2 = q.x + q.x, qY2 = q.y + q.y, qZ2 = q.z + q.z;
    float XX = qX2 * q.x, YY = qY2 * q.y, ZZ = qZ2 * q.z;
    float XY = qX2 * q.y, XZ = qX2 * q.z, XW = qX2 * q.w;
    float YZ = qY2 * q.z, YW = qY2 * q.w, ZW = qZ2 * q.w;
    M3 r;
    r.m11 = 1 - YY - ZZ; r.m21 = XY - ZW; r.m31 = XZ + YW;
    r.m12 = XY + ZW; r.m22 = 1 - XX - ZZ; r.m32 = YZ - XW;
    r.m13 = XZ - YW; r.m23 = YZ + XW; r.m33 = 1 - XX - YY;
    return r;
}
// Matrix3x3.Multiply (Matrix3x3.cs:766-791)
inline M3 m3_mul(const M3& a, const M3& b) {
    M3 r;
    r.m11 = a.m11 * b.m11 + a.m12 * b.m21 + a.m13 * b.m31;
    r.m12 = a.m11 * b.m12 + a.m12 * b.m22 + a.m13 * b.m32;
    r.m13 = a.m11 * b.m13 + a.m12 * b.m23 + a.m13 * b.m33;
    r.m21 = a.m21 * b.m11 + a.m22 * b.m21 + a.m23 * b.m31;
    r.m22 = a.m21 * b.m12 + a.m22 * b.m22 + a.m23 * b.m32;
    r.m23 = a.m21 * b.m13 + a.m22 * b.m23 + a.m23 * b.m33;
    r.m31 = a.m31 * b.m11 + a.m32 * b.m21 + a.m33 * b.m31;
    r.m32 = a.m31 * b.m12 + a.m32 * b.m22 + a.m33 * b.m32;
    r.m33 = a.m31 * b.m13 + a.m32 * b.m23 + a.m33 * b.m33;
    return r;
}
// Matrix3x3.MultiplyTransposed: transpose(t) * m (Matrix3x3.cs:866-891)
inline M3 m3_mul_transposed(const M3& t, const M3& m) {
    M3 r;
    r.m11 = t.m11 * m.m11 + t.m21 * m.m21 + t.m31 * m.m31;
    r.m12 = t.m11 * m.m12 + t.m21 * m.m22 + t.m31 * m.m32;
    r.m13 = t.m11 * m.m13 + t.m21 * m.m23 + t.m31 * m.m33;
    r.m21 = t.m12 * m.m11 + t.m22 * m.m21 + t.m32 * m.m31;
    r.m22 = t.m12 * m.m12 + t.m22 * m.m22 + t.m32 * m.m32;
    r.m23 = t.m12 * m.m13 + t.m22 * m.m23 + t.m32 * m.m33;
    r.m31 = t.m13 * m.m11 + t.m23 * m.m21 + t.m33 * m.m31;
    r.m32 = t.m13 * m.m12 + t.m23 * m.m22 + t.m33 * m.m32;
    r.m33 = t.m13 * m.m13 + t.m23 * m.m23 + t.m33 * m.m33;
    return r;
}
// Matrix3x3.Multiply(matrix, scale) (Matrix3x3.cs:932-945)
inline M3 m3_scale(const M3& m, float s) {
    return M3{m.m11 * s, m.m12 * s, m.m13 * s, m.m21 * s, m.m22 * s, m.m23 * s,
              m.m31 * s, m.m32 * s, m.m33 * s};
}
// Matrix3x3.Transform: row vector times matrix (Matrix3x3.cs:1071-1085)
inline V3 m3_transform(const V3& v, const M3& m) {
    return V3{v.x * m.m11 + v.y * m.m21 + v.z * m.m31, v.x * m.m12 + v.y * m.m22 + v.z * m.m32,
              v.x * m.m13 + v.y * m.m23 + v.z * m.m33};
}
// Matrix3x3.TransformTranspose (Matrix3x3.cs:1113-1125)
inline V3 m3_transform_transpose(const V3& v, const M3& m) {
    return V3{v.x * m.m11 + v.y * m.m12 + v.z * m.m13, v.x * m.m21 + v.y * m.m22 + v.z * m.m23,
              v.x * m.m31 + v.y * m.m32 + v.z * m.m33};
}
// Matrix3x3.AdaptiveDeterminant + AdaptiveInvert (Matrix3x3.cs:550-664, 1244-1290)
inline M3 m3_adaptive_invert(const M3& m) {
    int sub;
    float det = m.m11 * m.m22 * m.m33 + m.m12 * m.m23 * m.m31 + m.m13 * m.m21 * m.m32 -
                m.m31 * m.m22 * m.m13 - m.m32 * m.m23 * m.m11 - m.m33 * m.m21 * m.m12;
    if (det != 0) sub = 0;
    else if ((det = m.m11 * m.m22 - m.m12 * m.m21) != 0) sub = 1;
    else if ((det = m.m22 * m.m33 - m.m23 * m.m32) != 0) sub = 2;
    // The reference computes this minor as M11*M33 - M13*M12 (Matrix3x3.cs:1269); kept as is.
    else if ((det = m.m11 * m.m33 - m.m13 * m.m12) != 0) sub = 3;
    else if ((det = m.m11) != 0) sub = 4;
    else if ((det = m.m22) != 0) sub = 5;
    else if ((det = m.m33) != 0) sub = 6;
    else { det = 0; sub = -1; }
    float di = 1 / det;
    M3 r{0, 0, 0, 0, 0, 0, 0, 0, 0};
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
    case 1:
        r.m11 = m.m22 * di; r.m12 = -m.m12 * di; r.m21 = -m.m21 * di; r.m22 = m.m11 * di;
        break;
    case 2:
        r.m22 = m.m33 * di; r.m23 = -m.m23 * di; r.m32 = -m.m32 * di; r.m33 = m.m22 * di;
        break;
    case 3:
        r.m11 = m.m33 * di; r.m13 = -m.m13 * di; r.m31 = -m.m31 * di; r.m33 = m.m11 * di;
        break;
    case 4: r.m11 = 1 / m.m11; break;
    case 5: r.m22 = 1 / m.m22; break;
    case 6: r.m33 = 1 / m.m33; break;
    default: break;
    }
    return r;
}

// MathHelper.Clamp / Min / Max (BEPUutilities/MathHelper.cs)
inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float maxf(float a, float b) { return a > b ? a : b; }
inline float minf(float a, float b) { return a < b ? a : b; }
// Math.Sign(float) as used by BoxShape.GetBoundingBox (an int multiplied into a float).
inline float signf_int(float v) { return v > 0 ? 1.0f : (v < 0 ? -1.0f : 0.0f); }

}  // namespace ref
