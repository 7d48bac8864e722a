// convex.h -- the general convex pair test on the device: support mappings
// of boxes, spheres, capsules and hulls, boolean GJK, closest-point GJK, and
// the MPR contact with its surface casts and depth refinement, driven by the
// GeneralConvexPairTester state machine.  One thread runs one pair; every
// expression keeps the reference's operand order (files cited per function,
// relative to the reference repository root) so the results are bit-identical
// under -ffp-contract=off with correctly rounded division and square root.
#pragma once
#include <float.h>

#include "../../include/bepu_gpu.h"
#include "dmath.h"

namespace bepu {
namespace cvx {

constexpr float kEpsilon = 1e-7f;     // Toolbox.Epsilon
constexpr float kBigEpsilon = 1e-5f;  // Toolbox.BigEpsilon
constexpr int kMaxGJKIterations = 15;           // GJKToolbox.cs:17
constexpr int kHighGJKIterations = 8;           // GJKToolbox.cs:22
constexpr float kProgressionEpsilon = 1e-8f;    // PairSimplex.cs:97
constexpr float kDistanceConvergenceEpsilon = 1e-7f;  // PairSimplex.cs:102
constexpr int kInnerIterationLimit = 15;        // MPRToolbox.cs:24-99
constexpr int kOuterIterationLimit = 15;
constexpr float kSurfaceEpsilon = 1e-7f;
constexpr float kDepthRefinementEpsilon = 1e-4f;
constexpr int kMaxDepthRefinementIterations = 3;

enum { S_EMPTY = 0, S_POINT, S_SEGMENT, S_TRIANGLE, S_TETRAHEDRON };  // SimplexState
enum { CT_SEPARATED = 0, CT_SHALLOW = 1, CT_DEEP = 2 };              // pair tester state

struct RT {
    f3 pos;
    q4 q;
};

// A posed-shape-independent view of a ConvexShape: BEPU_SHAPE_* kind, box
// half extents (margin included), capsule half length, hull vertices, margin.
struct Shape {
    int kind;
    f3 half;
    float half_length;
    const float4* verts;  // vertex k at verts[k * vstride]
    int nverts;
    int vstride;
    float margin;
};

BEPU_D f3 zero3() { return f3{0, 0, 0}; }

// GetLocalExtremePointWithoutMargin: BoxShape.cs:182-185, SphereShape.cs:95-98,
// CapsuleShape.cs:146-154, ConvexHullShape.cs:289-306.
BEPU_D f3 support_local(const Shape& s, f3 d) {
    if (s.kind == BEPU_SHAPE_BOX)
        return f3{sign_ref(d.x) * (s.half.x - s.margin), sign_ref(d.y) * (s.half.y - s.margin),
                  sign_ref(d.z) * (s.half.z - s.margin)};
    if (s.kind == BEPU_SHAPE_CAPSULE) {
        if (d.y > 0) return f3{0, s.half_length, 0};
        if (d.y < 0) return f3{0, -s.half_length, 0};
        return zero3();
    }
    if (s.kind == BEPU_SHAPE_CONVEX_HULL) {
        // The first vertex of largest dot product (ties keep the earlier one).
        // Vertices are loaded eight at a time so their latencies overlap.
        constexpr int kChunk = 8;
        float mx = 0;
        f3 best = zero3();
        for (int i0 = 0; i0 < s.nverts; i0 += kChunk) {
            float4 v[kChunk];
#pragma unroll
            for (int k = 0; k < kChunk; ++k) v[k] = s.verts[min(i0 + k, s.nverts - 1) * s.vstride];
#pragma unroll
            for (int k = 0; k < kChunk; ++k) {
                float dd = v[k].x * d.x + v[k].y * d.y + v[k].z * d.z;
                if (i0 + k < s.nverts && (i0 + k == 0 || dd > mx)) {
                    mx = dd;
                    best = f3_of(v[k]);
                }
            }
        }
        return best;
    }
    return zero3();
}

// A mesh triangle in the local space of the convex it is tested against
// (TriangleShape, CollisionShapes/ConvexShapes/TriangleShape.cs), as shape B of
// the GJK and MPR routines below (which take B's type as a template
// parameter, so the closed set of entity shapes above keeps its own code).
struct TriShape {
    f3 v0, v1, v2;
    float margin;
};

// TriangleShape.GetLocalExtremePointWithoutMargin (TriangleShape.cs:166-185)
BEPU_D f3 support_local(const TriShape& s, f3 d) {
    float dA = dot3(d, s.v0), dB = dot3(d, s.v1), dC = dot3(d, s.v2);
    if (dA > dB && dA > dC) return s.v0;
    if (dB > dC) return s.v1;
    return s.v2;
}

// Quaternion.Concatenate (BEPUutilities/Quaternion.cs:101-118)
BEPU_D q4 qconcat(q4 a, q4 b) {
    q4 r;
    r.x = a.w * b.x + a.x * b.w + a.z * b.y - a.y * b.z;
    r.y = a.w * b.y + a.y * b.w + a.x * b.z - a.z * b.x;
    r.z = a.w * b.z + a.z * b.w + a.y * b.x - a.x * b.y;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}

// RigidTransform.Transform / TransformByInverse (RigidTransform.cs:135-156)
BEPU_D f3 rt_transform(f3 p, const RT& t) { return qrotate(p, t.q) + t.pos; }
BEPU_D f3 rt_transform_by_inverse(f3 p, const RT& t) { return qrotate(p - t.pos, qconj(t.q)); }

// MinkowskiToolbox.GetLocalTransform (MinkowskiToolbox.cs:19-29)
BEPU_D RT local_transform(const RT& a, const RT& b) {
    q4 ca = qconj(a.q);
    RT r;
    r.q = qconcat(b.q, ca);
    r.pos = qrotate(b.pos - a.pos, ca);
    return r;
}

// ConvexShape.GetExtremePointWithoutMargin (ConvexShape.cs:71-80)
template <class SB>
BEPU_D f3 extreme_without_margin(const SB& s, f3 direction, const RT& t) {
    f3 d = qrotate(direction, qconj(t.q));
    f3 p = support_local(s, d);
    p = qrotate(p, t.q);
    return p + t.pos;
}

// GetLocalMinkowskiExtremePoint, first overload, with the contribution form of
// ExpandMinkowskiSum (MinkowskiToolbox.cs:38-50, 133-148).
template <class SB>
BEPU_D f3 mk_extreme(const Shape& a, const SB& b, f3 d, const RT& lb) {
    f3 e = support_local(a, d);
    f3 v = extreme_without_margin(b, -d, lb);
    e = e - v;
    float ls = lensq3(d);
    v = ls > kEpsilon ? d * ((a.margin + b.margin) / sqrtf(ls)) : zero3();
    return e + v;
}

// GetLocalMinkowskiExtremePoint with the A-side point and the per-shape
// ExpandMinkowskiSum (MinkowskiToolbox.cs:63-80, 160-176) when `per_shape`,
// else the first overload (as mk_extreme).  The two round differently; one
// function keeps a single copy of the support scans in the surface cast.
template <class SB>
BEPU_D f3 mk_extreme_sel(const Shape& a, const SB& b, f3 d, const RT& lb, bool per_shape, f3* ea) {
    f3 pa = support_local(a, d);
    f3 pb = extreme_without_margin(b, -d, lb);
    float ls = lensq3(d);
    if (!per_shape) {
        f3 e = pa - pb;
        f3 v = ls > kEpsilon ? d * ((a.margin + b.margin) / sqrtf(ls)) : zero3();
        return e + v;
    }
    if (ls > kEpsilon) {
        ls = 1 / sqrtf(ls);
        pa = pa + d * (a.margin * ls);
        pb = pb - d * (b.margin * ls);
    }
    *ea = pa;
    return pa - pb;
}

// ---------------------------------------------------------------------------
// SimpleSimplex (GJK/SimpleSimplex.cs): the boolean test.

struct SSimplex {
    f3 A, B, C, D;
    int state;
};

// SimpleSimplex.GetPointOnTriangleClosestToOrigin (SimpleSimplex.cs:108-199)
BEPU_D f3 ss_triangle(SSimplex& s) {
    f3 ab = s.B - s.A, ac = s.C - s.A;
    float d5 = -dot3(ab, s.C), d6 = -dot3(ac, s.C);
    if (d6 >= 0.f && d5 <= d6) {
        s.state = S_POINT;
        s.A = s.C;
        return s.A;
    }
    float d1 = -dot3(ab, s.A), d2 = -dot3(ac, s.A);
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 > 0.f && d6 < 0.f) {
        s.state = S_SEGMENT;
        s.B = s.C;
        float V = d2 / (d2 - d6);
        return ac * V + s.A;
    }
    float d3 = -dot3(ab, s.B), d4 = -dot3(ac, s.B);
    float va = d3 * d6 - d5 * d4;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = d4 - d3) > 0.f && (d6d5 = d5 - d6) > 0.f) {
        s.state = S_SEGMENT;
        f3 B = s.B, C = s.C;
        s.A = C;
        float U = d3d4 / (d3d4 + d6d5);
        return (C - B) * U + B;
    }
    float vc = d1 * d4 - d3 * d2;
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    f3 p = ab * v;
    f3 acw = ac * w;
    p = s.A + p;
    return p + acw;
}

// SimpleSimplex.TryTetrahedronTriangle (SimpleSimplex.cs:254-370)
BEPU_D bool ss_try_face(f3 A, f3 B, f3 C, f3 other, SSimplex* out, f3* point) {
    *out = SSimplex{zero3(), zero3(), zero3(), zero3(), S_EMPTY};
    *point = zero3();
    f3 ab = B - A, ac = C - A;
    f3 normal = cross3(ab, ac);
    f3 AD = other - A;
    float AdotN = dot3(A, normal), ADdotN = dot3(AD, normal);
    if (!(AdotN * ADdotN > 0)) return false;
    float CdotAB = -dot3(ab, C), CdotAC = -dot3(ac, C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        out->state = S_POINT;
        out->A = C;
        *point = C;
        return true;
    }
    float AdotAB = -dot3(ab, A), AdotAC = -dot3(ac, A);
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        out->state = S_SEGMENT;
        out->A = A;
        out->B = C;
        float V = AdotAC / (AdotAC - CdotAC);
        *point = ac * V + A;
        return true;
    }
    float BdotAB = -dot3(ab, B), BdotAC = -dot3(ac, B);
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        out->state = S_SEGMENT;
        out->A = B;
        out->B = C;
        float V = d3d4 / (d3d4 + d6d5);
        *point = (C - B) * V + B;
        return true;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    out->A = A;
    out->B = B;
    out->C = C;
    out->state = S_TRIANGLE;
    float denom = 1.f / (va + vb + vc);
    float w = vc * denom, v = vb * denom;
    f3 p = ab * v;
    f3 acw = ac * w;
    p = A + p;
    *point = p + acw;
    return true;
}

// SimpleSimplex.GetPointClosestToOrigin (SimpleSimplex.cs:39-67, 73-83, 204-250)
BEPU_D bool ss_closest(SSimplex& s, f3* point) {
    switch (s.state) {
    case S_POINT: *point = s.A; return false;
    case S_SEGMENT: {
        f3 seg = s.B - s.A;
        float dotA = dot3(seg, s.A);
        float V = -dotA / lensq3(seg);
        *point = seg * V + s.A;
        return false;
    }
    case S_TRIANGLE: *point = ss_triangle(s); return false;
    case S_TETRAHEDRON: {
        SSimplex best{zero3(), zero3(), zero3(), zero3(), S_EMPTY};
        *point = zero3();
        float minimum = FLT_MAX;
        SSimplex cand;
        f3 cp;
        float cd;
        if (ss_try_face(s.A, s.C, s.D, s.B, &cand, &cp)) {
            *point = cp;
            best = cand;
            minimum = lensq3(cp);
        }
        if (ss_try_face(s.C, s.B, s.D, s.A, &cand, &cp) && (cd = lensq3(cp)) < minimum) {
            *point = cp;
            best = cand;
            minimum = cd;
        }
        if (ss_try_face(s.B, s.A, s.D, s.C, &cand, &cp) && (cd = lensq3(cp)) < minimum) {
            *point = cp;
            best = cand;
            minimum = cd;
        }
        if (minimum < FLT_MAX) {
            s = best;
            return false;
        }
        return true;
    }
    default: *point = zero3(); return false;
    }
}

// SimpleSimplex.GetErrorTolerance (SimpleSimplex.cs:403-420)
BEPU_D float ss_error_tolerance(const SSimplex& s) {
    switch (s.state) {
    case S_POINT: return lensq3(s.A);
    case S_SEGMENT: return maxf_ref(lensq3(s.A), lensq3(s.B));
    case S_TRIANGLE: return maxf_ref(lensq3(s.A), maxf_ref(lensq3(s.B), lensq3(s.C)));
    case S_TETRAHEDRON: return maxf_ref(lensq3(s.A), maxf_ref(lensq3(s.B), maxf_ref(lensq3(s.C), lensq3(s.D))));
    default: return 1;
    }
}

// GJKToolbox.AreShapesIntersecting with a warm-started axis (GJKToolbox.cs:47-90)
template <class SB>
BEPU_D bool gjk_intersecting(const Shape& a, const SB& b, const RT& ta, const RT& tb, f3* axis) {
    RT lb = local_transform(ta, tb);
    SSimplex s{mk_extreme(a, b, *axis, lb), zero3(), zero3(), zero3(), S_POINT};  // AddNewSimplexPoint on Empty
    f3 cp;
    int count = 0;
    while (count++ < kMaxGJKIterations) {
        if (ss_closest(s, &cp) || lensq3(cp) <= ss_error_tolerance(s) * kBigEpsilon) return true;
        f3 d = -cp;
        f3 e = mk_extreme(a, b, d, lb);
        if (dot3(e, cp) > 0) {
            *axis = d;
            return false;
        }
        // SimpleSimplex.AddNewSimplexPoint (SimpleSimplex.cs:376-397)
        if (s.state == S_POINT) { s.state = S_SEGMENT; s.B = e; }
        else if (s.state == S_SEGMENT) { s.state = S_TRIANGLE; s.C = e; }
        else if (s.state == S_TRIANGLE) { s.state = S_TETRAHEDRON; s.D = e; }
    }
    return false;
}

// ---------------------------------------------------------------------------
// PairSimplex (GJK/PairSimplex.cs): the closest-point test.

struct PSimplex {
    f3 sa[4], sb[4];
    f3 A, B, C, D;
    int state;
    float U, V, W;
};

BEPU_D void ps_clear(PSimplex& s) {
    for (int k = 0; k < 4; ++k) {
        s.sa[k] = zero3();
        s.sb[k] = zero3();
    }
    s.A = s.B = s.C = s.D = zero3();
    s.state = S_EMPTY;
    s.U = s.V = s.W = 0;
}

// PairSimplex.GetPointOnSegmentClosestToOrigin (PairSimplex.cs:377-420)
BEPU_D f3 ps_segment(PSimplex& s) {
    f3 seg = s.B - s.A;
    float dotA = dot3(seg, s.A);
    if (dotA > 0) {
        s.state = S_POINT;
        s.U = 1;
        return s.A;
    }
    float dotB = dot3(seg, s.B);
    if (dotB > 0) {
        s.U = dotB / lensq3(seg);
        s.V = 1 - s.U;
        return seg * s.V + s.A;
    }
    s.A = s.B;
    s.sa[0] = s.sa[1];
    s.sb[0] = s.sb[1];
    s.state = S_POINT;
    s.U = 1;
    return s.A;
}

// PairSimplex.GetPointOnTriangleClosestToOrigin (PairSimplex.cs:426-565)
BEPU_D f3 ps_triangle(PSimplex& s) {
    f3 ab = s.B - s.A, ac = s.C - s.A;
    float AdotAB = -dot3(ab, s.A), AdotAC = -dot3(ac, s.A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        s.state = S_POINT;
        s.U = 1;
        return s.A;
    }
    float BdotAB = -dot3(ab, s.B), BdotAC = -dot3(ac, s.B);
    if (BdotAB >= 0.f && BdotAC <= BdotAB) {
        s.state = S_POINT;
        s.A = s.B;
        s.U = 1;
        s.sa[0] = s.sa[1];
        s.sb[0] = s.sb[1];
        return s.B;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    if (vc <= 0 && AdotAB > 0 && BdotAB < 0) {
        s.state = S_SEGMENT;
        s.V = AdotAB / (AdotAB - BdotAB);
        s.U = 1 - s.V;
        return ab * s.V + s.A;
    }
    float CdotAB = -dot3(ab, s.C), CdotAC = -dot3(ac, s.C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        s.state = S_POINT;
        s.A = s.C;
        s.sa[0] = s.sa[2];
        s.sb[0] = s.sb[2];
        s.U = 1;
        return s.A;
    }
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        s.state = S_SEGMENT;
        s.B = s.C;
        s.sa[1] = s.sa[2];
        s.sb[1] = s.sb[2];
        s.V = AdotAC / (AdotAC - CdotAC);
        s.U = 1 - s.V;
        return ac * s.V + s.A;
    }
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        s.state = S_SEGMENT;
        s.A = s.C;
        s.sa[0] = s.sa[2];
        s.sb[0] = s.sb[2];
        s.U = d3d4 / (d3d4 + d6d5);
        s.V = 1 - s.U;
        return (s.C - s.B) * s.U + s.B;
    }
    float denom = 1.f / (va + vb + vc);
    s.V = vb * denom;
    s.W = vc * denom;
    s.U = 1 - s.V - s.W;
    f3 p = ab * s.V;
    f3 acw = ac * s.W;
    p = s.A + p;
    return p + acw;
}

// Result of one face test of the tetrahedron case: the sub-simplex kept
// (state and up to three vertex indices of the tetrahedron), its barycentric
// weights and its closest point.  The winner is applied to the simplex once,
// instead of carrying a candidate copy of the whole simplex per face.
struct FaceResult {
    int state;
    int v0, v1, v2;
    float U, V, W;
    f3 point;
};

template <int I>
BEPU_D f3 vtx(const PSimplex& s) {
    return I == 0 ? s.A : (I == 1 ? s.B : (I == 2 ? s.C : s.D));
}
BEPU_D f3 sel4(int i, f3 a, f3 b, f3 c, f3 d) { return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : d)); }

// PairSimplex.TryTetrahedronTriangle (PairSimplex.cs:655-845): face (i, j, k) of
// the tetrahedron with `o` the opposite vertex.
template <int i, int j, int k, int o>
BEPU_D bool ps_try_face(const PSimplex& s, float tol, FaceResult* r) {
    f3 A = vtx<i>(s), B = vtx<j>(s), C = vtx<k>(s);
    f3 ab = B - A, ac = C - A;
    f3 normal = cross3(ab, ac);
    f3 AD = vtx<o>(s) - A;
    float AdotN = dot3(A, normal), ADdotN = dot3(AD, normal);
    if (!(AdotN * ADdotN >= -kEpsilon * tol)) return false;
    r->V = r->W = 0;
    float AdotAB = -dot3(ab, A), AdotAC = -dot3(ac, A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        *r = FaceResult{S_POINT, i, 0, 0, 1, 0, 0, A};
        return true;
    }
    float BdotAB = -dot3(ab, B), BdotAC = -dot3(ac, B);
    if (BdotAB >= 0.f && BdotAC <= BdotAB) {
        *r = FaceResult{S_POINT, j, 0, 0, 1, 0, 0, B};
        return true;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    if (vc <= 0 && AdotAB > 0 && BdotAB < 0) {
        float V = AdotAB / (AdotAB - BdotAB);
        *r = FaceResult{S_SEGMENT, i, j, 0, 1 - V, V, 0, ab * V + A};
        return true;
    }
    float CdotAB = -dot3(ab, C), CdotAC = -dot3(ac, C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        *r = FaceResult{S_POINT, k, 0, 0, 1, 0, 0, C};
        return true;
    }
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        float V = AdotAC / (AdotAC - CdotAC);
        *r = FaceResult{S_SEGMENT, i, k, 0, 1 - V, V, 0, ac * V + A};
        return true;
    }
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        float V = d3d4 / (d3d4 + d6d5);
        *r = FaceResult{S_SEGMENT, j, k, 0, 1 - V, V, 0, (C - B) * V + B};
        return true;
    }
    float denom = 1.f / (va + vb + vc);
    float W = vc * denom;
    float V = vb * denom;
    f3 p = ab * V;
    f3 acw = ac * W;
    p = A + p;
    *r = FaceResult{S_TRIANGLE, i, j, k, 1 - V - W, V, W, p + acw};
    return true;
}

// PairSimplex.GetPointOnTetrahedronClosestToOrigin (PairSimplex.cs:572-650).  The
// simplex becomes the winning face's sub-simplex; fields beyond its state are
// never read before they are written again.
BEPU_D bool ps_tetrahedron(PSimplex& s, float tol, f3* point) {
    FaceResult best, cand;
    best.state = S_EMPTY;
    *point = zero3();
    float minimum = FLT_MAX;
    float cd;
    // Faces ACD|B, BDC|A, ADB|C, ABC|D.  The first accepted face is taken
    // without a comparison (as the reference does), which matters only for a
    // non-finite candidate.
    if (ps_try_face<0, 2, 3, 1>(s, tol, &cand)) {
        best = cand;
        minimum = lensq3(cand.point);
    }
    if (ps_try_face<1, 3, 2, 0>(s, tol, &cand) && (cd = lensq3(cand.point)) < minimum) {
        best = cand;
        minimum = cd;
    }
    if (ps_try_face<0, 3, 1, 2>(s, tol, &cand) && (cd = lensq3(cand.point)) < minimum) {
        best = cand;
        minimum = cd;
    }
    if (ps_try_face<0, 1, 2, 3>(s, tol, &cand) && (cd = lensq3(cand.point)) < minimum) {
        best = cand;
        minimum = cd;
    }
    if (minimum < FLT_MAX) {
        *point = best.point;
        f3 A = sel4(best.v0, s.A, s.B, s.C, s.D), B = sel4(best.v1, s.A, s.B, s.C, s.D),
           C = sel4(best.v2, s.A, s.B, s.C, s.D);
        f3 a0 = sel4(best.v0, s.sa[0], s.sa[1], s.sa[2], s.sa[3]), a1 = sel4(best.v1, s.sa[0], s.sa[1], s.sa[2], s.sa[3]),
           a2 = sel4(best.v2, s.sa[0], s.sa[1], s.sa[2], s.sa[3]);
        f3 b0 = sel4(best.v0, s.sb[0], s.sb[1], s.sb[2], s.sb[3]), b1 = sel4(best.v1, s.sb[0], s.sb[1], s.sb[2], s.sb[3]),
           b2 = sel4(best.v2, s.sb[0], s.sb[1], s.sb[2], s.sb[3]);
        s.state = best.state;
        s.A = A;
        s.B = B;
        s.C = C;
        s.sa[0] = a0;
        s.sa[1] = a1;
        s.sa[2] = a2;
        s.sb[0] = b0;
        s.sb[1] = b1;
        s.sb[2] = b2;
        s.U = best.U;
        s.V = best.V;
        s.W = best.W;
        return false;
    }
    return true;
}

// GJKToolbox.GetClosestPoints with the pair tester's cached simplex: a Point
// simplex whose B-side local point is `simplex_b` (the shapes' origins for
// GeneralConvexPairTester, the triangle's centroid for
// TriangleConvexPairTester.cs:241-242; GJKToolbox.cs:131-182,
// PairSimplex.cs:163-193, :341-370, :861-965).  Closest points of the cores,
// in world space.
template <class SB>
BEPU_D bool gjk_closest_points(const Shape& a, const SB& b, const RT& ta, const RT& tb, f3* ca, f3* cb,
                               f3 simplex_b = f3{0, 0, 0}) {
    RT lb = local_transform(ta, tb);
    PSimplex s;
    ps_clear(s);
    float prev_distance = FLT_MAX;
    float tol = 0;
    s.state = S_POINT;
    s.sb[0] = qrotate(simplex_b, lb.q) + lb.pos;
    s.A = s.sa[0] - s.sb[0];
    f3 cp;
    int count = 0;
    bool intersecting = false;
    while (true) {
        bool inside = false;
        if (s.state == S_POINT) {
            cp = s.A;
            s.U = 1;
        } else if (s.state == S_SEGMENT) {
            cp = ps_segment(s);
        } else if (s.state == S_TRIANGLE) {
            cp = ps_triangle(s);
        } else {
            inside = ps_tetrahedron(s, tol, &cp);
        }
        if (inside || lensq3(cp) <= kEpsilon * tol) {
            intersecting = true;
            break;
        }
        if (++count > kMaxGJKIterations) break;
        // PairSimplex.GetNewSimplexPoint (PairSimplex.cs:861-920)
        f3 nd = -cp;
        f3 pa = support_local(a, nd);
        f3 pb = extreme_without_margin(b, cp, lb);
        f3 S = pa - pb;
        float dotS = dot3(S, nd);
        float dist = lensq3(cp);
        float progression = dotS + dist;
        if (count > kHighGJKIterations && dist - prev_distance < kDistanceConvergenceEpsilon * tol) break;
        if (dist < prev_distance) prev_distance = dist;
        if (s.state == S_POINT) {
            tol = maxf_ref(lensq3(s.A), lensq3(S));
            if (progression <= tol * kProgressionEpsilon) break;
            s.state = S_SEGMENT;
            s.B = S;
            s.sa[1] = pa;
            s.sb[1] = pb;
        } else if (s.state == S_SEGMENT) {
            tol = maxf_ref(maxf_ref(lensq3(s.A), lensq3(s.B)), lensq3(S));
            if (progression <= tol * kProgressionEpsilon) break;
            s.state = S_TRIANGLE;
            s.C = S;
            s.sa[2] = pa;
            s.sb[2] = pb;
        } else if (s.state == S_TRIANGLE) {
            tol = maxf_ref(maxf_ref(lensq3(s.A), lensq3(s.B)), maxf_ref(lensq3(s.C), lensq3(S)));
            if (progression <= tol * kProgressionEpsilon) break;
            s.state = S_TETRAHEDRON;
            s.D = S;
            s.sa[3] = pa;
            s.sb[3] = pb;
        }
    }
    f3 la = zero3(), lbp = zero3();
    if (!intersecting) {
        // PairSimplex.GetClosestPoints (PairSimplex.cs:925-965)
        if (s.state == S_POINT) {
            la = s.sa[0];
            lbp = s.sb[0];
        } else if (s.state == S_SEGMENT) {
            la = s.sa[0] * s.U + s.sa[1] * s.V;
            lbp = s.sb[0] * s.U + s.sb[1] * s.V;
        } else if (s.state == S_TRIANGLE) {
            la = (s.sa[0] * s.U + s.sa[1] * s.V) + s.sa[2] * s.W;
            lbp = (s.sb[0] * s.U + s.sb[1] * s.V) + s.sb[2] * s.W;
        }
    }
    *ca = rt_transform(la, ta);
    *cb = rt_transform(lbp, ta);
    return intersecting;
}

// ---------------------------------------------------------------------------
// MPR (MPRToolbox.cs)

// MPRToolbox.AreLocalShapesOverlapping (MPRToolbox.cs:413-598): the internal
// overload at :426 with its origin ray; the public one at :413 passes B's
// local position.
template <class SB>
BEPU_D bool mpr_overlapping_ray(const Shape& a, const SB& b, f3 ray, const RT& lb) {
    if (lensq3(ray) < kEpsilon) return true;
    f3 v0 = -ray;
    f3 n = ray;
    f3 v1 = mk_extreme(a, b, n, lb);
    n = cross3(v1, v0);
    if (lensq3(n) < kEpsilon) return !(dot3(v1, ray) < 0);
    f3 v2 = mk_extreme(a, b, n, lb);
    n = cross3(v1 - v0, v2 - v0);
    f3 v3;
    int count = 0;
    while (true) {
        v3 = mk_extreme(a, b, n, lb);
        if (count > kOuterIterationLimit) break;
        count++;
        if (dot3(cross3(v1, v3), v0) < 0) {
            v2 = v3;
            n = cross3(v1 - v0, v3 - v0);
            continue;
        }
        if (dot3(cross3(v3, v2), v0) < 0) {
            v1 = v3;
            n = cross3(v2 - v0, v3 - v0);
            continue;
        }
        break;
    }
    while (true) {
        n = cross3(v3 - v2, v1 - v2);
        float d = dot3(n, v1);
        if (d >= 0) return true;
        f3 v4 = mk_extreme(a, b, n, lb);
        float d2 = dot3(v4, n);
        if (d2 < 0) return false;
        if (d2 - d < kSurfaceEpsilon || count > kInnerIterationLimit) return false;
        count++;
        f3 t = cross3(v4, v0);
        if (dot3(v1, t) >= 0) {
            if (dot3(v2, t) >= 0) v1 = v4;
            else v3 = v4;
        } else {
            if (dot3(v3, t) >= 0) v2 = v4;
            else v1 = v4;
        }
    }
}

BEPU_D bool mpr_overlapping(const Shape& a, const Shape& b, const RT& lb) { return mpr_overlapping_ray(a, b, lb.pos, lb); }

// Toolbox.GetBarycentricCoordinates (BEPUutilities/Toolbox.cs:1665-1737)
BEPU_D void barycentric(f3 p, f3 a, f3 b, f3 c, float* wa, float* wb, float* wc) {
    f3 ab = b - a, ac = c - a;
    f3 n = cross3(ab, ac);
    float x = n.x < 0 ? -n.x : n.x;
    float y = n.y < 0 ? -n.y : n.y;
    float z = n.z < 0 ? -n.z : n.z;
    float nu, nv, den;
    if (x >= y && x >= z) {
        nu = (p.y - b.y) * (b.z - c.z) - (b.y - c.y) * (p.z - b.z);
        nv = (p.y - c.y) * (c.z - a.z) - (c.y - a.y) * (p.z - c.z);
        den = n.x;
    } else if (y >= z) {
        nu = (p.x - b.x) * (b.z - c.z) - (b.x - c.x) * (p.z - b.z);
        nv = (p.x - c.x) * (c.z - a.z) - (c.x - a.x) * (p.z - c.z);
        den = -n.y;
    } else {
        nu = (p.x - b.x) * (b.y - c.y) - (b.x - c.x) * (p.y - b.y);
        nv = (p.x - c.x) * (c.y - a.y) - (c.x - a.x) * (p.y - c.y);
        den = n.z;
    }
    // The thresholds are double literals in the reference: compare in double.
    if ((double)den < -1e-9 || (double)den > 1e-9) {
        den = 1 / den;
        *wa = nu * den;
        *wb = nv * den;
        *wc = 1 - *wa - *wb;
    } else {
        float d1 = distsq3(p, a), d2 = distsq3(p, b), d3 = distsq3(p, c);
        if (d1 < d2 && d1 < d3) { *wa = 1; *wb = 0; *wc = 0; }
        else if (d2 < d3) { *wa = 0; *wb = 1; *wc = 0; }
        else { *wa = 0; *wb = 0; *wc = 1; }
    }
}

// MPRToolbox.LocalSurfaceCast, both overloads (MPRToolbox.cs:610-880 and
// 886-1172): `with_position` selects the overload that expands each shape's
// support point separately and interpolates the contact position.
template <class SB>
BEPU_D void surface_cast(const Shape& a, const SB& b, const RT& lb, f3 dir, bool with_position, float* t, f3* normal,
                         f3* position) {
    f3 v1A = zero3(), v2A = zero3(), v3A = zero3(), v4A = zero3();
    f3 n = dir;
    f3 v1 = mk_extreme_sel(a, b, n, lb, with_position, &v1A);
    n = cross3(dir, v1);
    if (lensq3(n) < kEpsilon) {
        float rl = lensq3(dir);
        *normal = rl > kEpsilon * .01f ? divide3(dir, sqrtf(rl)) : zero3();
        float rate = dot3(*normal, dir);
        float distance = dot3(*normal, v1);
        *t = rate > 0 ? distance / rate : 0;
        if (with_position) *position = v1A;
        return;
    }
    f3 v2 = mk_extreme_sel(a, b, n, lb, with_position, &v2A);
    n = cross3(v1, v2);
    if (dot3(n, dir) > 0) {
        n = -n;
        f3 tmp = v1;
        v1 = v2;
        v2 = tmp;
        tmp = v1A;
        v1A = v2A;
        v2A = tmp;
    }
    f3 v3;
    int count = 0;
    while (true) {
        v3 = mk_extreme_sel(a, b, n, lb, with_position, &v3A);
        if (count > kOuterIterationLimit) {
            *t = FLT_MAX;
            *normal = f3{0, 1, 0};
            if (with_position) *position = zero3();
            return;
        }
        count++;
        if (dot3(cross3(v1, v3), dir) < 0) {
            v2 = v3;
            v2A = v3A;
            n = cross3(v1, v3);
            continue;
        }
        if (dot3(cross3(v3, v2), dir) < 0) {
            v1 = v3;
            v1A = v3A;
            n = cross3(v2, v3);
            continue;
        }
        break;
    }
    count = 0;
    while (true) {
        n = cross3(v1 - v2, v3 - v2);
        f3 v4 = mk_extreme_sel(a, b, n, lb, with_position, &v4A);
        float d = dot3(n, v1);
        float sd = dot3(v4, n);
        if (sd - d < kSurfaceEpsilon || count > kInnerIterationLimit) {
            float ls = lensq3(n);
            if (ls > kEpsilon * .01f) {
                *normal = divide3(n, sqrtf(ls));
                d = dot3(*normal, dir);
                sd = dot3(*normal, v1);
                *t = d > 0 ? sd / d : 0;
            } else {
                *normal = f3{0, 1, 0};
                *t = 0;
            }
            if (with_position) {
                float w1, w2, w3;
                f3 p = dir * *t;
                barycentric(p, v1, v2, v3, &w1, &w2, &w3);
                p = v1A * w1;
                p = v2A * w2 + p;
                p = v3A * w3 + p;
                *position = p;
            }
            return;
        }
        f3 tt = cross3(v4, dir);
        if (dot3(v1, tt) >= 0) {
            if (dot3(v2, tt) >= 0) {
                v1 = v4;
                v1A = v4A;
            } else {
                v3 = v4;
                v3A = v4A;
            }
        } else {
            if (dot3(v3, tt) >= 0) {
                v2 = v4;
                v2A = v4A;
            } else {
                v1 = v4;
                v1A = v4A;
            }
        }
        count++;
    }
}

// MPRToolbox.GetContact (MPRToolbox.cs:1201-1280) with RefinePenetration (:1282-1322).
BEPU_D bool mpr_contact(const Shape& a, const Shape& b, const RT& ta, const RT& tb, f3 axis, f3* pos, f3* nrm,
                        float* depth) {
    RT lb = local_transform(ta, tb);
    if (!mpr_overlapping(a, b, lb)) return false;
    // The casts of GetContact run through one call site: phase 0 casts along
    // the given axis, phase 1 along the origin offset (keeping the shallower
    // result), then RefinePenetration casts along the current normal with the
    // contact position until the depth settles or three refinements ran.
    float dep = FLT_MAX;
    f3 normal = f3{0, 1, 0};
    f3 p = zero3();
    int phase = 0, count = 0;
    while (true) {
        f3 dir;
        if (phase < 2) {
            f3 src = phase == 0 ? axis : lb.pos;
            float ls = lensq3(src);
            if (!(ls > kEpsilon)) {
                ++phase;
                continue;
            }
            dir = divide3(src, sqrtf(ls));
        } else {
            dir = normal;
        }
        float cd;
        f3 cn, cp;
        surface_cast(a, b, lb, dir, phase == 2, &cd, &cn, &cp);
        if (phase == 0) {
            dep = cd;
            normal = cn;
            phase = 1;
        } else if (phase == 1) {
            if (cd < dep) {
                normal = cn;
                dep = cd;
            }
            phase = 2;
        } else {
            p = cp;
            if (dep - cd <= kDepthRefinementEpsilon || ++count >= kMaxDepthRefinementIterations) {
                dep = cd;
                break;
            }
            dep = cd;
            normal = cn;
        }
    }
    m3 o = mat_from_quat(ta.q);
    *nrm = mat_transform(normal, o);
    *pos = mat_transform(p, o) + ta.pos;
    *depth = dep;
    return true;
}

// Pair tester state carried between steps (GeneralConvexPairTester.cs:15-50).
struct TesterState {
    int state;
    f3 separating_axis;
    f3 local_direction;
};

// GeneralConvexPairTester.DoDeepContact (GeneralConvexPairTester.cs:182-318)
BEPU_D bool deep_contact(const Shape& a, const Shape& b, const RT& ta, const RT& tb, f3 va, f3 vb, int previous,
                         TesterState& st, f3* pos, f3* nrm, float* depth, int* id) {
    if (previous == CT_SEPARATED) {
        st.local_direction = va - vb;
        if (lensq3(st.local_direction) < kEpsilon) st.local_direction = f3{0, 1, 0};
    }
    if (mpr_contact(a, b, ta, tb, st.local_direction, pos, nrm, depth)) {
        *id = -1;
        if (*depth < a.margin + b.margin) st.state = CT_SHALLOW;
        return true;
    }
    st.state = CT_SEPARATED;
    return false;
}

// GeneralConvexPairTester.GenerateContactCandidate / DoShallowContact
// (GeneralConvexPairTester.cs:60-179)
BEPU_D bool candidate(const Shape& a, const Shape& b, const RT& ta, const RT& tb, f3 va, f3 vb, TesterState& st,
                      f3* pos, f3* nrm, float* depth, int* id) {
    int previous = st.state;
    // The deep test has one call site (entered from the Deep state or when the
    // closest-point test finds the cores intersecting), which keeps one copy
    // of the MPR code in the kernel.
    bool deep = st.state == CT_DEEP;
    if (!deep) {
        if (st.state == CT_SEPARATED) {
            if (!gjk_intersecting(a, b, ta, tb, &st.separating_axis)) return false;
            st.state = CT_SHALLOW;
        }
        f3 ca, cb;
        gjk_closest_points(a, b, ta, tb, &ca, &cb);
        f3 disp = cb - ca;
        float d2 = lensq3(disp);
        if (d2 < kEpsilon) {
            st.state = CT_DEEP;
            deep = true;
        } else {
            st.local_direction = disp;
            float margin = a.margin + b.margin;
            if (d2 < margin * margin) {
                f3 p = margin > kEpsilon ? disp * (a.margin / margin) : zero3();
                *pos = ca + p;
                float distance = sqrtf(d2);
                *nrm = divide3(disp, distance);
                *depth = margin - distance;
                *id = 0;
                return true;
            }
            st.state = CT_SEPARATED;
            return false;
        }
    }
    return deep_contact(a, b, ta, tb, va, vb, previous, st, pos, nrm, depth, id);
}

}  // namespace cvx
}  // namespace bepu
