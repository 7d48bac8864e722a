// dmath.h -- FP32 vector math for the gfx950 kernels.
//
// Every helper evaluates its expression in the same association order as the
// BEPUutilities member it mirrors, so that with -ffp-contract=off and
// correctly rounded division/square root the device produces the same bits
// as the reference's C# code (cited per helper, paths relative to the
// reference repository root).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BEPU_HD __host__ __device__ __forceinline__
#define BEPU_D __device__ __forceinline__

struct f3 {
    float x, y, z;
};
struct q4 {
    float x, y, z, w;
};
// Row-major 3x3, fields M11..M33 as in BEPUutilities.Matrix3x3.
struct m3 {
    float m11, m12, m13, m21, m22, m23, m31, m32, m33;
};

BEPU_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
BEPU_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }  // Vector3.Add
BEPU_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }  // Vector3.Subtract
BEPU_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }                      // Vector3.Negate
BEPU_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }    // Vector3.Multiply
BEPU_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }     // Vector3.Dot (Vector3.cs:115-118)
BEPU_HD f3 cross3(f3 a, f3 b) {                                                  // Vector3.Cross (Vector3.cs:548-556)
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
BEPU_HD float lensq3(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }  // Vector3.LengthSquared
BEPU_HD float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }  // Vector3.Length
BEPU_HD float distsq3(f3 a, f3 b) {  // Vector3.DistanceSquared (Vector3.cs:350-356)
    float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
    return x * x + y * y + z * z;
}
// Vector3.Normalize (Vector3.cs:81-87): reciprocal square root in FP64.
BEPU_HD f3 normalize3(f3 v) {
    float inv = (float)(1.0 / sqrt((double)(v.x * v.x + v.y * v.y + v.z * v.z)));
    return f3{v.x * inv, v.y * inv, v.z * inv};
}
// Vector3.Divide (Vector3.cs:175-181)
BEPU_HD f3 divide3(f3 v, float d) {
    float inv = 1 / d;
    return f3{v.x * inv, v.y * inv, v.z * inv};
}

// Quaternion.Multiply (Quaternion.cs:65-79)
BEPU_HD q4 qmul(q4 a, q4 b) {
    q4 r;
    r.x = a.x * b.w + b.x * a.w + a.y * b.z - a.z * b.y;
    r.y = a.y * b.w + b.y * a.w + a.z * b.x - a.x * b.z;
    r.z = a.z * b.w + b.z * a.w + a.x * b.y - a.y * b.x;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
// Quaternion.Normalize (Quaternion.cs:250-269): FP64 reciprocal square root.
BEPU_HD q4 qnormalize(q4 q) {
    float inv = (float)(1.0 / sqrt((double)(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w)));
    return q4{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
BEPU_HD q4 qconj(q4 q) { return q4{-q.x, -q.y, -q.z, q.w}; }
// Quaternion.Transform (Quaternion.cs:509-535)
BEPU_HD f3 qrotate(f3 v, q4 r) {
    float x2 = r.x + r.x, y2 = r.y + r.y, z2 = r.z + r.z;
    float xx2 = r.x * x2, xy2 = r.x * y2, xz2 = r.x * z2;
    float yy2 = r.y * y2, yz2 = r.y * z2, zz2 = r.z * z2;
    float wx2 = r.w * x2, wy2 = r.w * y2, wz2 = r.w * z2;
    return f3{v.x * (1.0f - yy2 - zz2) + v.y * (xy2 - wz2) + v.z * (xz2 + wy2),
              v.x * (xy2 + wz2) + v.y * (1.0f - xx2 - zz2) + v.z * (yz2 - wx2),
              v.x * (xz2 - wy2) + v.y * (yz2 + wx2) + v.z * (1.0f - xx2 - yy2)};
}
// Matrix3x3.CreateFromQuaternion (Matrix3x3.cs:1303-1327)
BEPU_HD m3 mat_from_quat(q4 q) {
    float qX2 = q.x + q.x, qY2 = q.y + q.y, qZ2 = q.z + q.z;
    float XX = qX2 * q.x, YY = qY2 * q.y, ZZ = qZ2 * q.z;
    float XY = qX2 * q.y, XZ = qX2 * q.z, XW = qX2 * q.w;
    float YZ = qY2 * q.z, YW = qY2 * q.w, ZW = qZ2 * q.w;
    m3 r;
    r.m11 = 1 - YY - ZZ; r.m21 = XY - ZW; r.m31 = XZ + YW;
    r.m12 = XY + ZW; r.m22 = 1 - XX - ZZ; r.m32 = YZ - XW;
    r.m13 = XZ - YW; r.m23 = YZ + XW; r.m33 = 1 - XX - YY;
    return r;
}
// Matrix3x3.MultiplyTransposed then Matrix3x3.Multiply: R^T * L * R, the
// world inverse inertia of Entity.UpdateForForces (EntityBase.cs:960-961).
BEPU_HD m3 rotate_inertia(const m3& R, const m3& L) {
    m3 t;
    t.m11 = R.m11 * L.m11 + R.m21 * L.m21 + R.m31 * L.m31;
    t.m12 = R.m11 * L.m12 + R.m21 * L.m22 + R.m31 * L.m32;
    t.m13 = R.m11 * L.m13 + R.m21 * L.m23 + R.m31 * L.m33;
    t.m21 = R.m12 * L.m11 + R.m22 * L.m21 + R.m32 * L.m31;
    t.m22 = R.m12 * L.m12 + R.m22 * L.m22 + R.m32 * L.m32;
    t.m23 = R.m12 * L.m13 + R.m22 * L.m23 + R.m32 * L.m33;
    t.m31 = R.m13 * L.m11 + R.m23 * L.m21 + R.m33 * L.m31;
    t.m32 = R.m13 * L.m12 + R.m23 * L.m22 + R.m33 * L.m32;
    t.m33 = R.m13 * L.m13 + R.m23 * L.m23 + R.m33 * L.m33;
    m3 r;
    r.m11 = t.m11 * R.m11 + t.m12 * R.m21 + t.m13 * R.m31;
    r.m12 = t.m11 * R.m12 + t.m12 * R.m22 + t.m13 * R.m32;
    r.m13 = t.m11 * R.m13 + t.m12 * R.m23 + t.m13 * R.m33;
    r.m21 = t.m21 * R.m11 + t.m22 * R.m21 + t.m23 * R.m31;
    r.m22 = t.m21 * R.m12 + t.m22 * R.m22 + t.m23 * R.m32;
    r.m23 = t.m21 * R.m13 + t.m22 * R.m23 + t.m23 * R.m33;
    r.m31 = t.m31 * R.m11 + t.m32 * R.m21 + t.m33 * R.m31;
    r.m32 = t.m31 * R.m12 + t.m32 * R.m22 + t.m33 * R.m32;
    r.m33 = t.m31 * R.m13 + t.m32 * R.m23 + t.m33 * R.m33;
    return r;
}
// Matrix3x3.Transform: row vector times matrix (Matrix3x3.cs:1071-1085)
BEPU_HD f3 mat_transform(f3 v, const m3& m) {
    return f3{v.x * m.m11 + v.y * m.m21 + v.z * m.m31, v.x * m.m12 + v.y * m.m22 + v.z * m.m32,
              v.x * m.m13 + v.y * m.m23 + v.z * m.m33};
}
// Matrix3x3.Multiply (Matrix3x3.cs:766-791)
BEPU_HD m3 mat_mul(const m3& a, const m3& b) {
    m3 r;
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
// Matrix3x3.Invert (Matrix3x3.cs:505-530) with Determinant (:1231-1235)
BEPU_HD m3 mat_invert(const m3& m) {
    const float det = m.m11 * m.m22 * m.m33 + m.m12 * m.m23 * m.m31 + m.m13 * m.m21 * m.m32 -
                      m.m31 * m.m22 * m.m13 - m.m32 * m.m23 * m.m11 - m.m33 * m.m21 * m.m12;
    const float d = 1 / det;
    m3 r;
    r.m11 = (m.m22 * m.m33 - m.m23 * m.m32) * d;
    r.m12 = (m.m13 * m.m32 - m.m33 * m.m12) * d;
    r.m13 = (m.m12 * m.m23 - m.m22 * m.m13) * d;
    r.m21 = (m.m23 * m.m31 - m.m21 * m.m33) * d;
    r.m22 = (m.m11 * m.m33 - m.m13 * m.m31) * d;
    r.m23 = (m.m13 * m.m21 - m.m11 * m.m23) * d;
    r.m31 = (m.m21 * m.m32 - m.m22 * m.m31) * d;
    r.m32 = (m.m12 * m.m31 - m.m11 * m.m32) * d;
    r.m33 = (m.m11 * m.m22 - m.m12 * m.m21) * d;
    return r;
}
// Matrix3x3.TransformTranspose (Matrix3x3.cs:1113-1125)
BEPU_HD f3 mat_transform_t(f3 v, const m3& m) {
    return f3{v.x * m.m11 + v.y * m.m12 + v.z * m.m13, v.x * m.m21 + v.y * m.m22 + v.z * m.m23,
              v.x * m.m31 + v.y * m.m32 + v.z * m.m33};
}
BEPU_HD f3 mat_row(const m3& m, int i) {
    return i == 0 ? f3{m.m11, m.m12, m.m13} : (i == 1 ? f3{m.m21, m.m22, m.m23} : f3{m.m31, m.m32, m.m33});
}

// MathHelper.Clamp / Max / Min (MathHelper.cs:58-88)
BEPU_HD float clampf_ref(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
BEPU_HD float maxf_ref(float a, float b) { return a > b ? a : b; }
BEPU_HD float minf_ref(float a, float b) { return a < b ? a : b; }
// Math.Sign(float) multiplied into a float (BoxShape.GetBoundingBox, BoxShape.cs:158-163)
BEPU_HD float sign_ref(float v) { return v > 0 ? 1.0f : (v < 0 ? -1.0f : 0.0f); }

BEPU_HD f3 f3_of(float4 v) { return f3{v.x, v.y, v.z}; }
BEPU_HD float4 f4_of(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Pair key and the colouring priority: a bijective 64-bit mix (SplitMix64 finaliser).
BEPU_HD uint64_t pair_key(uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return ((uint64_t)lo << 32) | hi;
}
BEPU_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
