// refcpu_convex.cpp -- oracle restatement of the general convex pair path
// (hulls and capsules, and any pair involving them): support mappings, the
// Minkowski-difference helpers, boolean GJK, closest-point GJK, the MPR
// contact with its surface casts and depth refinement, and the
// GeneralConvexPairTester state machine.
//
// TEST INFRASTRUCTURE ONLY (see refcpu.h).  Each expression keeps the operand
// order of the reference so that FP32 results are bit-identical to what the
// C# code computes on an IEEE binary32 machine.
#include <cfloat>
#include <cstring>
#include <vector>
#include "../include/bepu_gpu.h"
#include "refcpu.h"

namespace ref {

namespace {

const float kEpsilon = 1e-7f;     // Toolbox.Epsilon (BEPUutilities/Toolbox.cs:25)
const float kBigEpsilon = 1e-5f;  // Toolbox.BigEpsilon (Toolbox.cs:20)
// GJKToolbox.MaximumGJKIterations / HighGJKIterations (GJK/GJKToolbox.cs:17,22)
const int kMaxGJKIterations = 15;
const int kHighGJKIterations = 8;
// PairSimplex.ProgressionEpsilon / DistanceConvergenceEpsilon (GJK/PairSimplex.cs:97,102)
const float kProgressionEpsilon = 1e-8f;
const float kDistanceConvergenceEpsilon = 1e-7f;
// MPRToolbox limits and epsilons (MPRToolbox.cs:24-99)
const int kInnerIterationLimit = 15;
const int kOuterIterationLimit = 15;
const float kSurfaceEpsilon = 1e-7f;
const float kDepthRefinementEpsilon = 1e-4f;
const int kMaximumDepthRefinementIterations = 3;

const V3 kZero{0, 0, 0};
const V3 kUp{0, 1, 0};

enum { S_EMPTY = 0, S_POINT, S_SEGMENT, S_TRIANGLE, S_TETRAHEDRON };  // SimplexState (PairSimplex.cs:12-19)

// ---------------------------------------------------------------------------
// Support mappings

}  // namespace

// GetLocalExtremePointWithoutMargin of each shape:
// BoxShape.cs:182-185, SphereShape.cs:95-98, CapsuleShape.cs:146-154, ConvexHullShape.cs:289-306.
V3 support_local(const ConvexShapeRef& s, const V3& d) {
    switch (s.kind) {
    case BEPU_SHAPE_BOX:
        return V3{signf_int(d.x) * (s.half.x - s.margin), signf_int(d.y) * (s.half.y - s.margin),
                  signf_int(d.z) * (s.half.z - s.margin)};
    case BEPU_SHAPE_CAPSULE:
        if (d.y > 0) return V3{0, s.half_length, 0};
        if (d.y < 0) return V3{0, -s.half_length, 0};
        return kZero;
    case BEPU_SHAPE_CONVEX_HULL: {
        float mx = dot(s.verts[0], d);
        int mi = 0;
        for (int i = 1; i < s.nverts; i++) {
            float dd = dot(s.verts[i], d);
            if (dd > mx) {
                mx = dd;
                mi = i;
            }
        }
        return s.verts[mi];
    }
    case kShapeTriangle: {
        // TriangleShape.GetLocalExtremePointWithoutMargin (TriangleShape.cs:166-185)
        float dA = dot(d, s.verts[0]), dB = dot(d, s.verts[1]), dC = dot(d, s.verts[2]);
        if (dA > dB && dA > dC) return s.verts[0];
        if (dB > dC) return s.verts[1];
        return s.verts[2];
    }
    default:  // sphere
        return kZero;
    }
}

// Quaternion.Concatenate (BEPUutilities/Quaternion.cs:101-118)
Q4 qconcat(const Q4& a, const Q4& b) {
    Q4 r;
    r.x = a.w * b.x + a.x * b.w + a.z * b.y - a.y * b.z;
    r.y = a.w * b.y + a.y * b.w + a.x * b.z - a.z * b.x;
    r.z = a.w * b.z + a.z * b.w + a.y * b.x - a.x * b.y;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}

// RigidTransform.Transform / TransformByInverse (BEPUutilities/RigidTransform.cs:135-156)
V3 rt_transform(const V3& p, const RT& t) { return add(qtransform(p, t.q), t.pos); }
V3 rt_transform_by_inverse(const V3& p, const RT& t) { return qtransform(sub(p, t.pos), qconj(t.q)); }

// MinkowskiToolbox.GetLocalTransform (MinkowskiToolbox.cs:19-29): B in A's space.
RT local_transform(const RT& a, const RT& b) {
    Q4 ca = qconj(a.q);
    RT r;
    r.q = qconcat(b.q, ca);
    r.pos = qtransform(sub(b.pos, a.pos), ca);
    return r;
}

namespace {

// ConvexShape.GetExtremePointWithoutMargin (ConvexShape.cs:71-80)
V3 extreme_without_margin(const ConvexShapeRef& s, const V3& direction, const RT& t) {
    V3 d = qtransform(direction, qconj(t.q));
    V3 p = support_local(s, d);
    p = qtransform(p, t.q);
    return add(p, t.pos);
}

// MinkowskiToolbox.ExpandMinkowskiSum, contribution form (MinkowskiToolbox.cs:133-148)
V3 expand_contribution(float ma, float mb, const V3& d) {
    float ls = length_sq(d);
    if (ls > kEpsilon) return mul(d, (ma + mb) / std::sqrt(ls));
    return kZero;
}

// MinkowskiToolbox.GetLocalMinkowskiExtremePoint, first overload (MinkowskiToolbox.cs:38-50)
V3 mk_extreme(const ConvexShapeRef& a, const ConvexShapeRef& b, const V3& d, const RT& lb) {
    V3 e = support_local(a, d);
    V3 v = extreme_without_margin(b, neg(d), lb);
    e = sub(e, v);
    v = expand_contribution(a.margin, b.margin, d);
    return add(e, v);
}

// MinkowskiToolbox.GetLocalMinkowskiExtremePoint with the A-side point
// (MinkowskiToolbox.cs:63-80) and ExpandMinkowskiSum's per-shape form (:160-176).
V3 mk_extreme_a(const ConvexShapeRef& a, const ConvexShapeRef& b, const V3& d, const RT& lb, V3* ea) {
    V3 pa = support_local(a, d);
    V3 pb = extreme_without_margin(b, neg(d), lb);
    float ls = length_sq(d);
    if (ls > kEpsilon) {
        ls = 1 / std::sqrt(ls);
        V3 c = mul(d, a.margin * ls);
        pa = add(pa, c);
        c = mul(d, b.margin * ls);
        pb = sub(pb, c);
    }
    *ea = pa;
    return sub(pa, pb);
}

// ---------------------------------------------------------------------------
// SimpleSimplex (GJK/SimpleSimplex.cs), used by the boolean test.

struct SSimplex {
    V3 A, B, C, D;
    int state;
};

// SimpleSimplex.GetPointOnSegmentClosestToOrigin (SimpleSimplex.cs:73-83)
V3 ss_segment(const SSimplex& s) {
    V3 seg = sub(s.B, s.A);
    float dotA = dot(seg, s.A);
    float V = -dotA / length_sq(seg);
    return add(mul(seg, V), s.A);
}

// SimpleSimplex.GetPointOnTriangleClosestToOrigin (SimpleSimplex.cs:108-199)
V3 ss_triangle(SSimplex& s) {
    V3 ab = sub(s.B, s.A), ac = sub(s.C, s.A);
    float d5 = -dot(ab, s.C), d6 = -dot(ac, s.C);
    if (d6 >= 0.f && d5 <= d6) {
        s.state = S_POINT;
        s.A = s.C;
        return s.A;
    }
    float d1 = -dot(ab, s.A), d2 = -dot(ac, s.A);
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 > 0.f && d6 < 0.f) {
        s.state = S_SEGMENT;
        s.B = s.C;
        float V = d2 / (d2 - d6);
        return add(mul(ac, V), s.A);
    }
    float d3 = -dot(ab, s.B), d4 = -dot(ac, s.B);
    float va = d3 * d6 - d5 * d4;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = d4 - d3) > 0.f && (d6d5 = d5 - d6) > 0.f) {
        s.state = S_SEGMENT;
        V3 B = s.B, C = s.C;
        s.A = C;
        float U = d3d4 / (d3d4 + d6d5);
        return add(mul(sub(C, B), U), B);
    }
    float vc = d1 * d4 - d3 * d2;
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    V3 p = mul(ab, v);
    V3 acw = mul(ac, w);
    p = add(s.A, p);
    return add(p, acw);
}

// SimpleSimplex.TryTetrahedronTriangle (SimpleSimplex.cs:254-370)
bool ss_try_face(const V3& A, const V3& B, const V3& C, const V3& other, SSimplex* out, V3* point) {
    std::memset(out, 0, sizeof(*out));
    *point = kZero;
    V3 ab = sub(B, A), ac = sub(C, A);
    V3 normal = cross(ab, ac);
    V3 AD = sub(other, A);
    float AdotN = dot(A, normal), ADdotN = dot(AD, normal);
    if (!(AdotN * ADdotN > 0)) return false;
    float CdotAB = -dot(ab, C), CdotAC = -dot(ac, C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        out->state = S_POINT;
        out->A = C;
        *point = C;
        return true;
    }
    float AdotAB = -dot(ab, A), AdotAC = -dot(ac, A);
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        out->state = S_SEGMENT;
        out->A = A;
        out->B = C;
        float V = AdotAC / (AdotAC - CdotAC);
        *point = add(mul(ac, V), A);
        return true;
    }
    float BdotAB = -dot(ab, B), BdotAC = -dot(ac, B);
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        out->state = S_SEGMENT;
        out->A = B;
        out->B = C;
        float V = d3d4 / (d3d4 + d6d5);
        *point = add(mul(sub(C, B), V), B);
        return true;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    out->A = A;
    out->B = B;
    out->C = C;
    out->state = S_TRIANGLE;
    float denom = 1.f / (va + vb + vc);
    float w = vc * denom, v = vb * denom;
    V3 p = mul(ab, v);
    V3 acw = mul(ac, w);
    p = add(A, p);
    *point = add(p, acw);
    return true;
}

// SimpleSimplex.GetPointOnTetrahedronClosestToOrigin (SimpleSimplex.cs:204-250)
bool ss_tetrahedron(SSimplex& s, V3* point) {
    SSimplex best;
    std::memset(&best, 0, sizeof(best));
    *point = kZero;
    float minimum = FLT_MAX;
    SSimplex cand;
    V3 cp;
    float cd;
    if (ss_try_face(s.A, s.C, s.D, s.B, &cand, &cp)) {
        *point = cp;
        best = cand;
        minimum = length_sq(cp);
    }
    if (ss_try_face(s.C, s.B, s.D, s.A, &cand, &cp) && (cd = length_sq(cp)) < minimum) {
        *point = cp;
        best = cand;
        minimum = cd;
    }
    if (ss_try_face(s.B, s.A, s.D, s.C, &cand, &cp) && (cd = length_sq(cp)) < minimum) {
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

// SimpleSimplex.GetPointClosestToOrigin (SimpleSimplex.cs:39-67)
bool ss_closest(SSimplex& s, V3* point) {
    switch (s.state) {
    case S_POINT: *point = s.A; break;
    case S_SEGMENT: *point = ss_segment(s); break;
    case S_TRIANGLE: *point = ss_triangle(s); break;
    case S_TETRAHEDRON: return ss_tetrahedron(s, point);
    default: *point = kZero; break;
    }
    return false;
}

// SimpleSimplex.AddNewSimplexPoint (SimpleSimplex.cs:376-397)
void ss_add(SSimplex& s, const V3& p) {
    switch (s.state) {
    case S_EMPTY: s.state = S_POINT; s.A = p; break;
    case S_POINT: s.state = S_SEGMENT; s.B = p; break;
    case S_SEGMENT: s.state = S_TRIANGLE; s.C = p; break;
    case S_TRIANGLE: s.state = S_TETRAHEDRON; s.D = p; break;
    default: break;
    }
}

// SimpleSimplex.GetErrorTolerance (SimpleSimplex.cs:403-420)
float ss_error_tolerance(const SSimplex& s) {
    switch (s.state) {
    case S_POINT: return length_sq(s.A);
    case S_SEGMENT: return maxf(length_sq(s.A), length_sq(s.B));
    case S_TRIANGLE: return maxf(length_sq(s.A), maxf(length_sq(s.B), length_sq(s.C)));
    case S_TETRAHEDRON:
        return maxf(length_sq(s.A), maxf(length_sq(s.B), maxf(length_sq(s.C), length_sq(s.D))));
    default: return 1;
    }
}

// ---------------------------------------------------------------------------
// PairSimplex (GJK/PairSimplex.cs), used by the closest-point test.

struct PSimplex {
    V3 sa[4], sb[4];  // SimplexA / SimplexB: A, B, C, D
    V3 A, B, C, D;
    int state;
    float U, V, W;
    RT lb;
    float error_tolerance;
    float prev_distance;
};

// PairSimplex(ref CachedSimplex, ref RigidTransform) for the cached simplex the
// pair tester hands in: State = Point, LocalSimplexA.A zero and
// LocalSimplexB.A = `pb` (zero for GeneralConvexPairTester.cs:329, the
// triangle's centroid for TriangleConvexPairTester.cs:241-242; PairSimplex.cs:163-193).
void ps_init_point(PSimplex& s, const RT& lb, const V3& pb = kZero) {
    std::memset(&s, 0, sizeof(s));
    s.prev_distance = FLT_MAX;
    s.error_tolerance = 0;
    s.lb = lb;
    s.state = S_POINT;
    s.sb[0] = qtransform(pb, lb.q);
    s.sb[0] = add(s.sb[0], lb.pos);
    s.A = sub(s.sa[0], s.sb[0]);
}

// PairSimplex.GetPointOnSegmentClosestToOrigin (PairSimplex.cs:377-420)
V3 ps_segment(PSimplex& s) {
    V3 seg = sub(s.B, s.A);
    float dotA = dot(seg, s.A);
    if (dotA > 0) {
        s.state = S_POINT;
        s.U = 1;
        return s.A;
    }
    float dotB = dot(seg, s.B);
    if (dotB > 0) {
        s.U = dotB / length_sq(seg);
        s.V = 1 - s.U;
        return add(mul(seg, s.V), s.A);
    }
    s.A = s.B;
    s.sa[0] = s.sa[1];
    s.sb[0] = s.sb[1];
    s.state = S_POINT;
    s.U = 1;
    return s.A;
}

// PairSimplex.GetPointOnTriangleClosestToOrigin (PairSimplex.cs:426-565)
V3 ps_triangle(PSimplex& s) {
    V3 ab = sub(s.B, s.A), ac = sub(s.C, s.A);
    float AdotAB = -dot(ab, s.A), AdotAC = -dot(ac, s.A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        s.state = S_POINT;
        s.U = 1;
        return s.A;
    }
    float BdotAB = -dot(ab, s.B), BdotAC = -dot(ac, s.B);
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
        return add(mul(ab, s.V), s.A);
    }
    float CdotAB = -dot(ab, s.C), CdotAC = -dot(ac, s.C);
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
        return add(mul(ac, s.V), s.A);
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
        return add(mul(sub(s.C, s.B), s.U), s.B);
    }
    float denom = 1.f / (va + vb + vc);
    s.V = vb * denom;
    s.W = vc * denom;
    s.U = 1 - s.V - s.W;
    V3 p = mul(ab, s.V);
    V3 acw = mul(ac, s.W);
    p = add(s.A, p);
    return add(p, acw);
}

// PairSimplex.TryTetrahedronTriangle (PairSimplex.cs:655-845)
bool ps_try_face(const V3& A, const V3& B, const V3& C, const V3& A1, const V3& B1, const V3& C1, const V3& A2,
                 const V3& B2, const V3& C2, float tol, const V3& other, PSimplex* out, V3* point) {
    std::memset(out, 0, sizeof(*out));
    *point = kZero;
    V3 ab = sub(B, A), ac = sub(C, A);
    V3 normal = cross(ab, ac);
    V3 AD = sub(other, A);
    float AdotN = dot(A, normal), ADdotN = dot(AD, normal);
    if (!(AdotN * ADdotN >= -kEpsilon * tol)) return false;
    float AdotAB = -dot(ab, A), AdotAC = -dot(ac, A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        out->state = S_POINT;
        out->A = A;
        out->U = 1;
        out->sa[0] = A1;
        out->sb[0] = A2;
        *point = A;
        return true;
    }
    float BdotAB = -dot(ab, B), BdotAC = -dot(ac, B);
    if (BdotAB >= 0.f && BdotAC <= BdotAB) {
        out->state = S_POINT;
        out->A = B;
        out->U = 1;
        out->sa[0] = B1;
        out->sb[0] = B2;
        *point = B;
        return true;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    if (vc <= 0 && AdotAB > 0 && BdotAB < 0) {
        out->state = S_SEGMENT;
        out->V = AdotAB / (AdotAB - BdotAB);
        out->U = 1 - out->V;
        out->A = A;
        out->B = B;
        out->sa[0] = A1;
        out->sb[0] = A2;
        out->sa[1] = B1;
        out->sb[1] = B2;
        *point = add(mul(ab, out->V), A);
        return true;
    }
    float CdotAB = -dot(ab, C), CdotAC = -dot(ac, C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        out->state = S_POINT;
        out->A = C;
        out->U = 1;
        out->sa[0] = C1;
        out->sb[0] = C2;
        *point = C;
        return true;
    }
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        out->state = S_SEGMENT;
        out->A = A;
        out->B = C;
        out->sa[0] = A1;
        out->sa[1] = C1;
        out->sb[0] = A2;
        out->sb[1] = C2;
        out->V = AdotAC / (AdotAC - CdotAC);
        out->U = 1 - out->V;
        *point = add(mul(ac, out->V), A);
        return true;
    }
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        out->state = S_SEGMENT;
        out->A = B;
        out->B = C;
        out->sa[0] = B1;
        out->sa[1] = C1;
        out->sb[0] = B2;
        out->sb[1] = C2;
        out->V = d3d4 / (d3d4 + d6d5);
        out->U = 1 - out->V;
        *point = add(mul(sub(C, B), out->V), B);
        return true;
    }
    out->A = A;
    out->B = B;
    out->C = C;
    out->sa[0] = A1;
    out->sa[1] = B1;
    out->sa[2] = C1;
    out->sb[0] = A2;
    out->sb[1] = B2;
    out->sb[2] = C2;
    out->state = S_TRIANGLE;
    float denom = 1.f / (va + vb + vc);
    out->W = vc * denom;
    out->V = vb * denom;
    out->U = 1 - out->V - out->W;
    V3 p = mul(ab, out->V);
    V3 acw = mul(ac, out->W);
    p = add(A, p);
    *point = add(p, acw);
    return true;
}

// PairSimplex.GetPointOnTetrahedronClosestToOrigin (PairSimplex.cs:572-650)
bool ps_tetrahedron(PSimplex& s, V3* point) {
    PSimplex best;
    std::memset(&best, 0, sizeof(best));
    *point = kZero;
    float minimum = FLT_MAX;
    PSimplex cand;
    V3 cp;
    float cd;
    if (ps_try_face(s.A, s.C, s.D, s.sa[0], s.sa[2], s.sa[3], s.sb[0], s.sb[2], s.sb[3], s.error_tolerance, s.B, &cand,
                    &cp)) {
        *point = cp;
        best = cand;
        minimum = length_sq(cp);
    }
    if (ps_try_face(s.B, s.D, s.C, s.sa[1], s.sa[3], s.sa[2], s.sb[1], s.sb[3], s.sb[2], s.error_tolerance, s.A, &cand,
                    &cp) &&
        (cd = length_sq(cp)) < minimum) {
        *point = cp;
        best = cand;
        minimum = cd;
    }
    if (ps_try_face(s.A, s.D, s.B, s.sa[0], s.sa[3], s.sa[1], s.sb[0], s.sb[3], s.sb[1], s.error_tolerance, s.C, &cand,
                    &cp) &&
        (cd = length_sq(cp)) < minimum) {
        *point = cp;
        best = cand;
        minimum = cd;
    }
    if (ps_try_face(s.A, s.B, s.C, s.sa[0], s.sa[1], s.sa[2], s.sb[0], s.sb[1], s.sb[2], s.error_tolerance, s.D, &cand,
                    &cp) &&
        (cd = length_sq(cp)) < minimum) {
        *point = cp;
        best = cand;
        minimum = cd;
    }
    if (minimum < FLT_MAX) {
        best.lb = s.lb;
        best.prev_distance = s.prev_distance;
        best.error_tolerance = s.error_tolerance;
        s = best;
        return false;
    }
    return true;
}

// PairSimplex.GetPointClosestToOrigin (PairSimplex.cs:341-370)
bool ps_closest(PSimplex& s, V3* point) {
    switch (s.state) {
    case S_POINT:
        *point = s.A;
        s.U = 1;
        break;
    case S_SEGMENT: *point = ps_segment(s); break;
    case S_TRIANGLE: *point = ps_triangle(s); break;
    case S_TETRAHEDRON: return ps_tetrahedron(s, point);
    default: *point = kZero; break;
    }
    return false;
}

// PairSimplex.GetNewSimplexPoint (PairSimplex.cs:861-920)
bool ps_new_point(PSimplex& s, const ConvexShapeRef& a, const ConvexShapeRef& b, int iteration, const V3& cp) {
    V3 nd = neg(cp);
    V3 pa = support_local(a, nd);
    V3 pb = extreme_without_margin(b, cp, s.lb);
    V3 S = sub(pa, pb);
    float dotS = dot(S, nd);
    float dist = length_sq(cp);
    float progression = dotS + dist;
    if (iteration > kHighGJKIterations && dist - s.prev_distance < kDistanceConvergenceEpsilon * s.error_tolerance)
        return true;
    if (dist < s.prev_distance) s.prev_distance = dist;
    switch (s.state) {
    case S_POINT:
        s.error_tolerance = maxf(length_sq(s.A), length_sq(S));
        if (progression <= s.error_tolerance * kProgressionEpsilon) return true;
        s.state = S_SEGMENT;
        s.B = S;
        s.sa[1] = pa;
        s.sb[1] = pb;
        return false;
    case S_SEGMENT:
        s.error_tolerance = maxf(maxf(length_sq(s.A), length_sq(s.B)), length_sq(S));
        if (progression <= s.error_tolerance * kProgressionEpsilon) return true;
        s.state = S_TRIANGLE;
        s.C = S;
        s.sa[2] = pa;
        s.sb[2] = pb;
        return false;
    case S_TRIANGLE:
        s.error_tolerance = maxf(maxf(length_sq(s.A), length_sq(s.B)), maxf(length_sq(s.C), length_sq(S)));
        if (progression <= s.error_tolerance * kProgressionEpsilon) return true;
        s.state = S_TETRAHEDRON;
        s.D = S;
        s.sa[3] = pa;
        s.sb[3] = pb;
        return false;
    default: return false;
    }
}

// PairSimplex.GetClosestPoints (PairSimplex.cs:925-965)
void ps_closest_points(const PSimplex& s, V3* ca, V3* cb) {
    switch (s.state) {
    case S_POINT:
        *ca = s.sa[0];
        *cb = s.sb[0];
        return;
    case S_SEGMENT:
        *ca = add(mul(s.sa[0], s.U), mul(s.sa[1], s.V));
        *cb = add(mul(s.sb[0], s.U), mul(s.sb[1], s.V));
        return;
    case S_TRIANGLE: {
        V3 p = add(mul(s.sa[0], s.U), mul(s.sa[1], s.V));
        *ca = add(p, mul(s.sa[2], s.W));
        p = add(mul(s.sb[0], s.U), mul(s.sb[1], s.V));
        *cb = add(p, mul(s.sb[2], s.W));
        return;
    }
    default:
        *ca = kZero;
        *cb = kZero;
    }
}

// ---------------------------------------------------------------------------
// MPR (MPRToolbox.cs)

// MPRToolbox.AreLocalShapesOverlapping (MPRToolbox.cs:413-598): the internal
// overload at :426 with its origin ray; the public one at :413 passes B's local position.
bool mpr_overlapping_ray(const ConvexShapeRef& a, const ConvexShapeRef& b, const V3& ray, const RT& lb) {
    if (length_sq(ray) < kEpsilon) return true;
    V3 v0 = neg(ray);
    V3 n = ray;
    V3 v1 = mk_extreme(a, b, n, lb);
    n = cross(v1, v0);
    if (length_sq(n) < kEpsilon) return !(dot(v1, ray) < 0);
    V3 v2 = mk_extreme(a, b, n, lb);
    n = cross(sub(v1, v0), sub(v2, v0));
    V3 v3;
    int count = 0;
    while (true) {
        v3 = mk_extreme(a, b, n, lb);
        if (count > kOuterIterationLimit) break;
        count++;
        if (dot(cross(v1, v3), v0) < 0) {
            v2 = v3;
            n = cross(sub(v1, v0), sub(v3, v0));
            continue;
        }
        if (dot(cross(v3, v2), v0) < 0) {
            v1 = v3;
            n = cross(sub(v2, v0), sub(v3, v0));
            continue;
        }
        break;
    }
    while (true) {
        n = cross(sub(v3, v2), sub(v1, v2));
        float d = dot(n, v1);
        if (d >= 0) return true;
        V3 v4 = mk_extreme(a, b, n, lb);
        float d2 = dot(v4, n);
        if (d2 < 0) return false;
        if (d2 - d < kSurfaceEpsilon || count > kInnerIterationLimit) return false;
        count++;
        V3 t = cross(v4, v0);
        if (dot(v1, t) >= 0) {
            if (dot(v2, t) >= 0) v1 = v4;
            else v3 = v4;
        } else {
            if (dot(v3, t) >= 0) v2 = v4;
            else v1 = v4;
        }
    }
}

bool mpr_overlapping(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& lb) {
    return mpr_overlapping_ray(a, b, lb.pos, lb);
}

// Toolbox.GetBarycentricCoordinates (BEPUutilities/Toolbox.cs:1665-1737)
void barycentric(const V3& p, const V3& a, const V3& b, const V3& c, float* wa, float* wb, float* wc) {
    V3 ab = sub(b, a), ac = sub(c, a);
    V3 n = cross(ab, ac);
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
    // The thresholds are double literals: the comparison happens in double.
    if ((double)den < -1e-9 || (double)den > 1e-9) {
        den = 1 / den;
        *wa = nu * den;
        *wb = nv * den;
        *wc = 1 - *wa - *wb;
    } else {
        float d1 = distance_sq(p, a), d2 = distance_sq(p, b), d3 = distance_sq(p, c);
        if (d1 < d2 && d1 < d3) { *wa = 1; *wb = 0; *wc = 0; }
        else if (d2 < d3) { *wa = 0; *wb = 1; *wc = 0; }
        else { *wa = 0; *wb = 0; *wc = 1; }
    }
}

// MPRToolbox.LocalSurfaceCast, both overloads (MPRToolbox.cs:610-880 without the
// position, 886-1172 with it).  The overloads differ in how the Minkowski
// support point is formed (difference plus margin expansion versus expanding
// each shape's point) and so in their rounding; `want_position` selects.
void surface_cast(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& lb, const V3& dir, bool want_position,
                  float* t, V3* normal, V3* position) {
    V3 v1A{}, v2A{}, v3A{}, v4A{};
    auto support = [&](const V3& d, V3* ea) {
        return want_position ? mk_extreme_a(a, b, d, lb, ea) : mk_extreme(a, b, d, lb);
    };
    V3 n = dir;
    V3 v1 = support(n, &v1A);
    n = cross(dir, v1);
    if (length_sq(n) < kEpsilon) {
        float rl = length_sq(dir);
        if (rl > kEpsilon * .01f) *normal = divide(dir, std::sqrt(rl));
        else *normal = kZero;
        float rate = dot(*normal, dir);
        float distance = dot(*normal, v1);
        *t = rate > 0 ? distance / rate : 0;
        if (want_position) *position = v1A;
        return;
    }
    V3 v2 = support(n, &v2A);
    n = cross(v1, v2);
    if (dot(n, dir) > 0) {
        n = neg(n);
        V3 tmp = v1;
        v1 = v2;
        v2 = tmp;
        tmp = v1A;
        v1A = v2A;
        v2A = tmp;
    }
    V3 v3;
    int count = 0;
    while (true) {
        v3 = support(n, &v3A);
        if (count > kOuterIterationLimit) {
            *t = FLT_MAX;
            *normal = kUp;
            if (want_position) *position = kZero;
            return;
        }
        count++;
        if (dot(cross(v1, v3), dir) < 0) {
            v2 = v3;
            v2A = v3A;
            n = cross(v1, v3);
            continue;
        }
        if (dot(cross(v3, v2), dir) < 0) {
            v1 = v3;
            v1A = v3A;
            n = cross(v2, v3);
            continue;
        }
        break;
    }
    count = 0;
    while (true) {
        n = cross(sub(v1, v2), sub(v3, v2));
        V3 v4 = support(n, &v4A);
        float d = dot(n, v1);
        float sd = dot(v4, n);
        if (sd - d < kSurfaceEpsilon || count > kInnerIterationLimit) {
            float ls = length_sq(n);
            if (ls > kEpsilon * .01f) {
                *normal = divide(n, std::sqrt(ls));
                d = dot(*normal, dir);
                sd = dot(*normal, v1);
                *t = d > 0 ? sd / d : 0;
            } else {
                *normal = kUp;
                *t = 0;
            }
            if (want_position) {
                float w1, w2, w3;
                V3 p = mul(dir, *t);
                barycentric(p, v1, v2, v3, &w1, &w2, &w3);
                p = mul(v1A, w1);
                p = add(mul(v2A, w2), p);
                p = add(mul(v3A, w3), p);
                *position = p;
            }
            return;
        }
        V3 tt = cross(v4, dir);
        if (dot(v1, tt) >= 0) {
            if (dot(v2, tt) >= 0) {
                v1 = v4;
                v1A = v4A;
            } else {
                v3 = v4;
                v3A = v4A;
            }
        } else {
            if (dot(v3, tt) >= 0) {
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

// MPRToolbox.RefinePenetration (MPRToolbox.cs:1282-1322)
void refine_penetration(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& lb, float initial_depth,
                        const V3& initial_normal, float* depth, V3* normal, V3* position) {
    int count = 0;
    *normal = initial_normal;
    *depth = initial_depth;
    float cd;
    V3 cn;
    while (true) {
        surface_cast(a, b, lb, *normal, true, &cd, &cn, position);
        if (*depth - cd <= kDepthRefinementEpsilon || ++count >= kMaximumDepthRefinementIterations) {
            *depth = cd;
            break;
        }
        *depth = cd;
        *normal = cn;
    }
}

}  // namespace

// GJKToolbox.AreShapesIntersecting with a warm-started axis (GJKToolbox.cs:47-90)
bool gjk_intersecting(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, V3* axis) {
    RT lb = local_transform(ta, tb);
    SSimplex s;
    std::memset(&s, 0, sizeof(s));
    s.state = S_EMPTY;
    ss_add(s, mk_extreme(a, b, *axis, lb));
    V3 cp;
    int count = 0;
    while (count++ < kMaxGJKIterations) {
        if (ss_closest(s, &cp) || length_sq(cp) <= ss_error_tolerance(s) * kBigEpsilon) return true;
        V3 d = neg(cp);
        V3 e = mk_extreme(a, b, d, lb);
        if (dot(e, cp) > 0) {
            *axis = d;
            return false;
        }
        ss_add(s, e);
    }
    return false;
}

// GJKToolbox.GetClosestPoints with the pair tester's cached simplex (GJKToolbox.cs:131-182):
// closest points of the shapes' cores, returned in world space.
bool gjk_closest_points(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, V3* ca, V3* cb) {
    return gjk_closest_points_from(a, b, ta, tb, kZero, ca, cb);
}
bool gjk_closest_points_from(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb,
                             const V3& simplex_b, V3* ca, V3* cb) {
    RT lb = local_transform(ta, tb);
    PSimplex s;
    ps_init_point(s, lb, simplex_b);
    V3 cp;
    int count = 0;
    bool intersecting = false;
    while (true) {
        if (ps_closest(s, &cp) || length_sq(cp) <= kEpsilon * s.error_tolerance) {
            *ca = kZero;
            *cb = kZero;
            intersecting = true;
            break;
        }
        if (++count > kMaxGJKIterations) break;
        if (ps_new_point(s, a, b, count, cp)) break;
    }
    if (!intersecting) ps_closest_points(s, ca, cb);
    *ca = rt_transform(*ca, ta);
    *cb = rt_transform(*cb, ta);
    return intersecting;
}

// MPRToolbox.GetContact (MPRToolbox.cs:1201-1280)
bool mpr_contact(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& axis,
                 ContactData* c) {
    RT lb = local_transform(ta, tb);
    if (!mpr_overlapping(a, b, lb)) {
        std::memset(c, 0, sizeof(*c));
        return false;
    }
    float depth;
    V3 normal, dir;
    float ls = length_sq(axis);
    if (ls > kEpsilon) {
        dir = divide(axis, std::sqrt(ls));
        surface_cast(a, b, lb, dir, false, &depth, &normal, nullptr);
    } else {
        depth = FLT_MAX;
        normal = kUp;
    }
    ls = length_sq(lb.pos);
    if (ls > kEpsilon) {
        dir = divide(lb.pos, std::sqrt(ls));
        float dc;
        V3 nc;
        surface_cast(a, b, lb, dir, false, &dc, &nc, nullptr);
        if (dc < depth) {
            normal = nc;
            depth = dc;
        }
    }
    V3 pos;
    refine_penetration(a, b, lb, depth, normal, &depth, &normal, &pos);
    c->id = -1;
    M3 o = m3_from_quat(ta.q);
    c->normal = m3_transform(normal, o);
    c->pos = add(m3_transform(pos, o), ta.pos);
    c->depth = depth;
    return true;
}

namespace {

// GeneralConvexPairTester.DoShallowContact / DoDeepContact (GeneralConvexPairTester.cs:106-318)
bool deep_contact(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& va,
                  const V3& vb, int previous, ConvexTesterState* st, ContactData* c) {
    if (previous == CT_SEPARATED) {
        // Both collidables belong to entities: the relative linear velocity.
        st->local_direction = sub(va, vb);
        if (length_sq(st->local_direction) < kEpsilon) st->local_direction = kUp;
    }
    if (mpr_contact(a, b, ta, tb, st->local_direction, c)) {
        if (c->depth < a.margin + b.margin) st->state = CT_SHALLOW;
        return true;
    }
    st->state = CT_SEPARATED;
    return false;
}

bool shallow_contact(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& va,
                     const V3& vb, int previous, ConvexTesterState* st, ContactData* c) {
    V3 ca, cb;
    gjk_closest_points(a, b, ta, tb, &ca, &cb);
    V3 disp = sub(cb, ca);
    float d2 = length_sq(disp);
    if (d2 < kEpsilon) {
        st->state = CT_DEEP;
        return deep_contact(a, b, ta, tb, va, vb, previous, st, c);
    }
    st->local_direction = disp;
    float margin = a.margin + b.margin;
    if (d2 < margin * margin) {
        std::memset(c, 0, sizeof(*c));
        if (margin > kEpsilon) c->pos = mul(disp, a.margin / margin);
        else c->pos = kZero;
        c->pos = add(ca, c->pos);
        float distance = std::sqrt(d2);
        c->normal = divide(disp, distance);
        c->depth = margin - distance;
        return true;
    }
    st->state = CT_SEPARATED;
    std::memset(c, 0, sizeof(*c));
    return false;
}

}  // namespace

// GeneralConvexPairTester.GenerateContactCandidate (GeneralConvexPairTester.cs:60-104)
bool convex_candidate(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& va,
                      const V3& vb, ConvexTesterState* st, ContactData* c) {
    int previous = st->state;
    switch (st->state) {
    case CT_SEPARATED:
        if (gjk_intersecting(a, b, ta, tb, &st->separating_axis)) {
            st->state = CT_SHALLOW;
            return shallow_contact(a, b, ta, tb, va, vb, previous, st, c);
        }
        std::memset(c, 0, sizeof(*c));
        return false;
    case CT_SHALLOW: return shallow_contact(a, b, ta, tb, va, vb, previous, st, c);
    case CT_DEEP: return deep_contact(a, b, ta, tb, va, vb, previous, st, c);
    default: std::memset(c, 0, sizeof(*c)); return false;
    }
}

// ---------------------------------------------------------------------------
// Casts for continuous collision detection (GJK/GJKToolbox.cs:195-495) and
// their simplex, RaySimplex (GJK/RaySimplex.cs).  The reduction functions run
// on the simplex shifted to the current hit location (`s`, the reference's
// `this`) and reduce the unshifted simplex (`out`) in place.

namespace {

struct RSimplex {
    V3 A, B, C, D;
    int state;
};

// RaySimplex.GetPointOnSegmentClosestToOrigin (RaySimplex.cs:76-115).  Past
// B the reduced simplex is B, and the point returned is the shifted A, as
// the reference returns it.
V3 rs_segment(const RSimplex& s, RSimplex& out) {
    V3 seg = sub(s.B, s.A);
    float dotA = dot(seg, s.A);
    if (dotA > 0) {
        out.state = S_POINT;
        return s.A;
    }
    float dotB = dot(seg, s.B);
    if (dotB > 0) {
        float V = -dotA / length_sq(seg);
        return add(mul(seg, V), s.A);
    }
    out.A = out.B;
    out.state = S_POINT;
    return s.A;
}

// RaySimplex.GetPointOnTriangleClosestToOrigin (RaySimplex.cs:122-240).  In
// the region of C the reduced simplex is C and the point returned is the
// shifted A, as the reference returns it.
V3 rs_triangle(const RSimplex& s, RSimplex& out) {
    V3 ab = sub(s.B, s.A), ac = sub(s.C, s.A);
    float AdotAB = -dot(ab, s.A), AdotAC = -dot(ac, s.A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        out.state = S_POINT;
        return s.A;
    }
    float BdotAB = -dot(ab, s.B), BdotAC = -dot(ac, s.B);
    if (BdotAB >= 0.f && BdotAC <= BdotAB) {
        out.state = S_POINT;
        out.A = out.B;
        return s.B;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    if (vc <= 0 && AdotAB > 0 && BdotAB < 0) {
        out.state = S_SEGMENT;
        float V = AdotAB / (AdotAB - BdotAB);
        return add(mul(ab, V), s.A);
    }
    float CdotAB = -dot(ab, s.C), CdotAC = -dot(ac, s.C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        out.state = S_POINT;
        out.A = out.C;
        return s.A;
    }
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        out.state = S_SEGMENT;
        out.B = out.C;
        float V = AdotAC / (AdotAC - CdotAC);
        return add(mul(ac, V), s.A);
    }
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        out.state = S_SEGMENT;
        out.A = out.C;
        float U = d3d4 / (d3d4 + d6d5);
        return add(mul(sub(s.C, s.B), U), s.B);
    }
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    V3 point = mul(ab, v);
    V3 acw = mul(ac, w);
    point = add(s.A, point);
    return add(point, acw);
}

// RaySimplex.TryTetrahedronTriangle (RaySimplex.cs:320-470): the face ABC
// (shifted), its unshifted vertices sA..sC, and the fourth vertex.
bool rs_try_face(const V3& A, const V3& B, const V3& C, const V3& sA, const V3& sB, const V3& sC, const V3& other,
                 RSimplex* cand, V3* point) {
    std::memset(cand, 0, sizeof(*cand));
    *point = kZero;
    V3 ab = sub(B, A), ac = sub(C, A);
    V3 normal = cross(ab, ac);
    V3 AD = sub(other, A);
    float AdotN = dot(A, normal), ADdotN = dot(AD, normal);
    if (!(AdotN * ADdotN >= 0)) return false;
    float AdotAB = -dot(ab, A), AdotAC = -dot(ac, A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        cand->state = S_POINT;
        cand->A = sA;
        *point = A;
        return true;
    }
    float BdotAB = -dot(ab, B), BdotAC = -dot(ac, B);
    if (BdotAB >= 0.f && BdotAC <= BdotAB) {
        cand->state = S_POINT;
        cand->A = sB;
        *point = B;
        return true;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    if (vc <= 0 && AdotAB > 0 && BdotAB < 0) {
        cand->state = S_SEGMENT;
        cand->A = sA;
        cand->B = sB;
        float V = AdotAB / (AdotAB - BdotAB);
        *point = add(mul(ab, V), A);
        return true;
    }
    float CdotAB = -dot(ab, C), CdotAC = -dot(ac, C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        cand->state = S_POINT;
        cand->A = sC;
        *point = C;
        return true;
    }
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        cand->state = S_SEGMENT;
        cand->A = sA;
        cand->B = sC;
        float V = AdotAC / (AdotAC - CdotAC);
        *point = add(mul(ac, V), A);
        return true;
    }
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        cand->state = S_SEGMENT;
        cand->A = sB;
        cand->B = sC;
        float V = d3d4 / (d3d4 + d6d5);
        *point = add(mul(sub(C, B), V), B);
        return true;
    }
    cand->state = S_TRIANGLE;
    cand->A = sA;
    cand->B = sB;
    cand->C = sC;
    float denom = 1.f / (va + vb + vc);
    float w = vc * denom, v = vb * denom;
    V3 p = mul(ab, v);
    V3 acw = mul(ac, w);
    p = add(A, p);
    *point = add(p, acw);
    return true;
}

// RaySimplex.GetPointOnTetrahedronClosestToOrigin (RaySimplex.cs:247-317):
// the closest of the faces DAC, DCB, DBA (their origin-side tests) and ABC;
// true when none qualifies (the origin is enclosed; the point is zero).
bool rs_tetrahedron(const RSimplex& s, RSimplex& out, V3* point) {
    RSimplex best;
    std::memset(&best, 0, sizeof(best));
    *point = kZero;
    float best_d = FLT_MAX;
    RSimplex cand;
    V3 cp;
    float d;
    if (rs_try_face(s.A, s.C, s.D, out.A, out.C, out.D, s.B, &cand, &cp)) {
        *point = cp;
        best = cand;
        best_d = length_sq(cp);
    }
    if (rs_try_face(s.C, s.B, s.D, out.C, out.B, out.D, s.A, &cand, &cp) && (d = length_sq(cp)) < best_d) {
        *point = cp;
        best = cand;
        best_d = d;
    }
    if (rs_try_face(s.B, s.A, s.D, out.B, out.A, out.D, s.C, &cand, &cp) && (d = length_sq(cp)) < best_d) {
        *point = cp;
        best = cand;
        best_d = d;
    }
    if (rs_try_face(s.A, s.B, s.C, out.A, out.B, out.C, s.D, &cand, &cp) && (d = length_sq(cp)) < best_d) {
        *point = cp;
        best = cand;
        best_d = d;
    }
    if (best_d < FLT_MAX) {
        out = best;
        return false;
    }
    return true;
}

// RaySimplex.GetPointClosestToOrigin (RaySimplex.cs:40-69), called on the
// shifted simplex with the unshifted one to reduce.
V3 rs_closest(const RSimplex& s, RSimplex& out) {
    V3 p;
    switch (s.state) {
    case S_POINT: return s.A;
    case S_SEGMENT: return rs_segment(s, out);
    case S_TRIANGLE: return rs_triangle(s, out);
    case S_TETRAHEDRON: rs_tetrahedron(s, out, &p); return p;
    default: return kZero;
    }
}

// RaySimplex.AddNewSimplexPoint (RaySimplex.cs:478-520): the simplex grows
// by `p`; returns it shifted to the hit location (hit - vertex).
RSimplex rs_add(RSimplex& s, const V3& p, const V3& hit) {
    RSimplex sh;
    std::memset(&sh, 0, sizeof(sh));
    switch (s.state) {
    case S_EMPTY:
        s.state = S_POINT;
        s.A = p;
        sh.A = sub(hit, s.A);
        break;
    case S_POINT:
        s.state = S_SEGMENT;
        s.B = p;
        sh.A = sub(hit, s.A);
        sh.B = sub(hit, s.B);
        break;
    case S_SEGMENT:
        s.state = S_TRIANGLE;
        s.C = p;
        sh.A = sub(hit, s.A);
        sh.B = sub(hit, s.B);
        sh.C = sub(hit, s.C);
        break;
    case S_TRIANGLE:
        s.state = S_TETRAHEDRON;
        s.D = p;
        sh.A = sub(hit, s.A);
        sh.B = sub(hit, s.B);
        sh.C = sub(hit, s.C);
        sh.D = sub(hit, s.D);
        break;
    default: break;
    }
    sh.state = s.state;
    return sh;
}

// RaySimplex.GetErrorTolerance (RaySimplex.cs:527-560)
float rs_error_tolerance(const RSimplex& s, const V3& o) {
    switch (s.state) {
    case S_POINT: return distance_sq(s.A, o);
    case S_SEGMENT: return maxf(distance_sq(s.A, o), distance_sq(s.B, o));
    case S_TRIANGLE: return maxf(distance_sq(s.A, o), maxf(distance_sq(s.B, o), distance_sq(s.C, o)));
    case S_TETRAHEDRON:
        return maxf(distance_sq(s.A, o), maxf(distance_sq(s.B, o), maxf(distance_sq(s.C, o), distance_sq(s.D, o))));
    default: return 0;
    }
}

// ConvexShape.GetLocalExtremePoint (ConvexShape.cs:104-113): with the margin.
V3 local_extreme_with_margin(const ConvexShapeRef& s, V3 d) {
    V3 p = support_local(s, d);
    float ls = length_sq(d);
    if (ls > kEpsilon) {
        d = mul(d, s.margin / std::sqrt(ls));
        p = add(p, d);
    }
    return p;
}

}  // namespace

// GJKToolbox.SphereCast (GJKToolbox.cs:387-459): a ray from `origin` along
// `dir` (lengths in units of dir) swept by a sphere of `radius` against the
// posed shape; *t receives the hit parameter, also where the cast fails (0
// after too many iterations, as the reference's fresh RayHit).
bool sphere_cast(V3 origin, V3 dir, float radius, const ConvexShapeRef& shape, const RT& xf, float max_length,
                 float* t, V3* normal) {
    V3 hn{0, 0, 0};
    origin = sub(origin, xf.pos);
    Q4 cj = qconj(xf.q);
    origin = qtransform(origin, cj);
    dir = qtransform(dir, cj);
    float T = 0;
    V3 loc = origin;
    V3 v = loc;
    RSimplex simplex;
    std::memset(&simplex, 0, sizeof(simplex));
    int count = 0;
    while (length_sq(v) >= kEpsilon * rs_error_tolerance(simplex, origin)) {
        if (++count > kMaxGJKIterations) {
            *t = 0;
            return false;
        }
        V3 p = support_local(shape, v);
        V3 contribution = expand_contribution(shape.margin, radius, v);
        p = add(p, contribution);
        V3 w = sub(loc, p);
        float vw = dot(v, w);
        if (vw > 0) {
            float vdir = dot(v, dir);
            T = T - vw / vdir;
            if (vdir >= 0 || T > max_length) {
                *t = T;
                return false;
            }
            loc = add(mul(dir, T), origin);
            hn = v;
        }
        RSimplex shifted = rs_add(simplex, p, loc);
        v = rs_closest(shifted, simplex);
    }
    *t = T;
    if (normal) *normal = qtransform(hn, xf.q);
    return true;
}

// GJKToolbox.RayCast (GJKToolbox.cs:195-265): the ray against the shape with
// its margin.
bool ray_cast(V3 origin, V3 dir, const ConvexShapeRef& shape, const RT& xf, float max_length, float* t,
              V3* normal) {
    V3 hn{0, 0, 0};
    origin = sub(origin, xf.pos);
    Q4 cj = qconj(xf.q);
    origin = qtransform(origin, cj);
    dir = qtransform(dir, cj);
    float T = 0;
    V3 loc = origin;
    V3 off = loc;
    RSimplex simplex;
    std::memset(&simplex, 0, sizeof(simplex));
    int count = 0;
    while (length_sq(off) >= kEpsilon * rs_error_tolerance(simplex, origin)) {
        if (++count > kMaxGJKIterations) {
            *t = 0;
            return false;
        }
        V3 e = local_extreme_with_margin(shape, off);
        V3 eo = sub(loc, e);
        float vw = dot(off, eo);
        if (vw > 0) {
            float cd = dot(off, dir);
            if (cd >= 0) {
                *t = 0;
                return false;
            }
            T = T - vw / cd;
            if (T > max_length) {
                *t = 0;
                return false;
            }
            loc = add(mul(dir, T), origin);
            hn = off;
        }
        RSimplex shifted = rs_add(simplex, e, loc);
        off = rs_closest(shifted, simplex);
    }
    *t = T;
    if (normal) *normal = qtransform(hn, xf.q);
    return true;
}

// GJKToolbox.CCDSphereCast (GJKToolbox.cs:470-495): the sphere cast, retried
// with the radius scaled by `scaling` while it starts embedded (a hit at 0),
// up to three times, then a plain ray cast.
bool ccd_sphere_cast(const V3& origin, const V3& dir, float radius, const ConvexShapeRef& shape, const RT& xf,
                     float max_length, float scaling, float* t) {
    V3 n;
    return ccd_sphere_cast_normal(origin, dir, radius, shape, xf, max_length, scaling, t, &n);
}
bool ccd_sphere_cast_normal(const V3& origin, const V3& dir, float radius, const ConvexShapeRef& shape, const RT& xf,
                            float max_length, float scaling, float* t, V3* normal) {
    int iterations = 0;
    for (;;) {
        float T;
        if (sphere_cast(origin, dir, radius, shape, xf, max_length, &T, normal) && T > 0) {
            *t = T;
            return true;
        }
        if (T > max_length || T < 0) {
            *t = T;
            return false;
        }
        radius *= scaling;
        iterations++;
        if (iterations > 3) {
            bool hit = ray_cast(origin, dir, shape, xf, max_length, &T, normal);
            *t = T;
            return hit && T > 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Triangles of static meshes: the TriangleShape support (above, in
// support_local), Toolbox.GetClosestPointOnTriangleToPoint and the two
// TrianglePairTesters.  The convex is shape A at the identity, the triangle
// (already in the convex's local space) shape B at the identity.

V3 local_extreme_point(const ConvexShapeRef& s, const V3& d) { return local_extreme_with_margin(s, d); }

ConvexShapeRef triangle_shape(const TriangleRef& t) {
    ConvexShapeRef r;
    std::memset(&r, 0, sizeof(r));
    r.kind = kShapeTriangle;
    r.verts = t.v;
    r.nverts = 3;
    r.margin = t.margin;
    return r;
}

int closest_point_on_triangle(const V3& a, const V3& b, const V3& c, const V3& p, V3* out) {
    V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 < 0) {
        *out = a;
        return VR_A;
    }
    V3 bp = sub(p, b);
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
        *out = b;
        return VR_B;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        float v = d1 / (d1 - d3);
        *out = add(mul(ab, v), a);
        return VR_AB;
    }
    V3 cp = sub(p, c);
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) {
        *out = c;
        return VR_C;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        float w = d2 / (d2 - d6);
        *out = add(mul(ac, w), a);
        return VR_AC;
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *out = add(mul(sub(c, b), w), b);
        return VR_BC;
    }
    float denom = 1 / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    *out = add(add(a, mul(ab, v)), mul(ac, w));
    return VR_ABC;
}

namespace {

const RT kIdentityRT{V3{0, 0, 0}, Q4{0, 0, 0, 1}};
constexpr int kEscapeAttemptPeriod = 10;  // TriangleConvexPairTester.cs:16

// TriangleConvexPairTester.GetVoronoiRegion (TriangleConvexPairTester.cs:619-686)
int tri_voronoi(const TriangleRef& t, const V3& p) {
    V3 ab = sub(t.v[1], t.v[0]), ac = sub(t.v[2], t.v[0]), ap = sub(p, t.v[0]);
    float APdotAB = dot(ap, ab), APdotAC = dot(ap, ac);
    if (APdotAC <= 0.f && APdotAB <= 0) return VR_A;
    V3 bp = sub(p, t.v[1]);
    float BPdotAB = dot(ab, bp), BPdotAC = dot(ac, bp);
    if (BPdotAB >= 0.f && BPdotAC <= BPdotAB) return VR_B;
    float vc = APdotAB * BPdotAC - BPdotAB * APdotAC;
    if (vc <= 0 && APdotAB > 0 && BPdotAB < 0) return VR_AB;
    V3 cp = sub(p, t.v[2]);
    float CPdotAB = dot(ab, cp), CPdotAC = dot(ac, cp);
    if (CPdotAC >= 0.f && CPdotAB <= CPdotAC) return VR_C;
    float vb = CPdotAB * APdotAC - APdotAB * CPdotAC;
    if (vb <= 0.f && APdotAC > 0.f && CPdotAC < 0.f) return VR_AC;
    float va = BPdotAB * CPdotAC - CPdotAB * BPdotAC;
    if (va <= 0.f && (BPdotAC - BPdotAB) > 0.f && (CPdotAB - CPdotAC) > 0.f) return VR_BC;
    return VR_ABC;
}

// TryToEscape, both overloads (TriangleConvexPairTester.cs:529-545)
void tri_escape(TriTesterState* st) {
    if (++st->escapes == kEscapeAttemptPeriod) {
        st->escapes = 0;
        st->state = TT_PLANE;
    }
}
void tri_escape_at(TriTesterState* st, const TriangleRef& t, const V3& position) {
    if (++st->escapes == kEscapeAttemptPeriod && tri_voronoi(t, position) == VR_ABC) {
        st->escapes = 0;
        st->state = TT_PLANE;
    }
}

// TryInnerSphereContact (TriangleConvexPairTester.cs:548-613)
bool tri_inner_sphere(float min_radius, float core_scaling, const TriangleRef& t, ContactData* c) {
    std::memset(c, 0, sizeof(*c));
    V3 closest;
    closest_point_on_triangle(t.v[0], t.v[1], t.v[2], kZero, &closest);
    float length = length_sq(closest);
    float minimumRadius = min_radius * (core_scaling + .01f);
    if (length < minimumRadius * minimumRadius) {
        V3 n = cross(sub(t.v[1], t.v[0]), sub(t.v[2], t.v[0]));
        float d = dot(closest, n);
        if ((t.sidedness == SIDED_CLOCKWISE && d > 0) || (t.sidedness == SIDED_COUNTERCLOCKWISE && d < 0)) return false;
        length = std::sqrt(length);
        c->pos = closest;
        if (length > kEpsilon) {
            c->normal = divide(closest, length);
        } else {
            float nl = length_sq(n);
            if (length_sq(n) > kEpsilon) {
                n = divide(n, std::sqrt(nl));
                c->normal = t.sidedness == SIDED_CLOCKWISE ? n : neg(n);
            } else {
                std::memset(c, 0, sizeof(*c));
                return false;
            }
        }
        c->depth = min_radius - length;
        c->id = -1;
        return true;
    }
    return false;
}

// The surface casts of the deep test: a normal facing against a one-sided
// triangle loses its component along the triangle normal and the depth is
// scaled to the corrected direction (TriangleConvexPairTester.cs:372-459).
// A degenerate correction sets the *contact's* depth and normal, not the
// candidate's, as the reference does.
void tri_cast_corrected(const ConvexShapeRef& cvx, const ConvexShapeRef& tri, int sided, const V3& tn, const V3& dir,
                        float* cd, V3* cn, float* depth, V3* normal) {
    surface_cast(cvx, tri, kIdentityRT, dir, false, cd, cn, nullptr);
    float d = dot(tn, *cn);
    if ((sided == SIDED_CLOCKWISE && d > 0) || (sided == SIDED_COUNTERCLOCKWISE && d < 0)) {
        V3 previous = *cn;
        d = dot(*cn, tn);
        V3 p = mul(*cn, d);
        *cn = sub(*cn, p);
        float length = length_sq(*cn);
        if (length > kEpsilon) {
            *cn = divide(*cn, std::sqrt(length));
            d = dot(*cn, previous);
            *cd *= d;
        } else {
            *depth = FLT_MAX;
            *normal = kZero;
        }
    }
}

int tri_deep(const ConvexShapeRef& cvx, float min_radius, float core_scaling, const TriangleRef& t,
             TriTesterState* st, ContactData out[2]);

// DoExternalNear (TriangleConvexPairTester.cs:230-298)
int tri_near(const ConvexShapeRef& cvx, float min_radius, float core_scaling, const TriangleRef& t,
             TriTesterState* st, ContactData out[2]) {
    const ConvexShapeRef tri = triangle_shape(t);
    V3 centroid = add(add(t.v[0], t.v[1]), t.v[2]);
    centroid = mul(centroid, .33333333f);
    V3 ca, cb;
    if (gjk_closest_points_from(cvx, tri, kIdentityRT, kIdentityRT, centroid, &ca, &cb)) {
        st->state = TT_DEEP;
        return tri_deep(cvx, min_radius, core_scaling, t, st, out);
    }
    V3 disp = sub(cb, ca);
    float d2 = length_sq(disp);
    float margin = cvx.margin + t.margin;
    if (d2 < margin * margin) {
        ContactData c;
        std::memset(&c, 0, sizeof(c));
        if (t.sidedness != SIDED_DOUBLE) {
            V3 n = cross(sub(t.v[1], t.v[0]), sub(t.v[2], t.v[0]));
            float d = dot(n, disp);
            if (t.sidedness == SIDED_CLOCKWISE && d > 0) return 0;
            if (t.sidedness == SIDED_COUNTERCLOCKWISE && d < 0) return 0;
        }
        if (margin > kEpsilon) c.pos = mul(disp, cvx.margin / margin);
        else c.pos = kZero;
        c.pos = add(ca, c.pos);
        c.normal = disp;
        float distance = std::sqrt(d2);
        c.normal = divide(c.normal, distance);
        c.depth = margin - distance;
        out[0] = c;
        tri_escape_at(st, t, c.pos);
        return 1;
    }
    st->state = TT_SEPARATED;
    return 0;
}

// DoExternalSeparated (TriangleConvexPairTester.cs:215-228)
int tri_separated(const ConvexShapeRef& cvx, float min_radius, float core_scaling, const TriangleRef& t,
                  TriTesterState* st, ContactData out[2]) {
    if (gjk_intersecting(cvx, triangle_shape(t), kIdentityRT, kIdentityRT, &st->axis)) {
        st->state = TT_NEAR;
        return tri_near(cvx, min_radius, core_scaling, t, st, out);
    }
    tri_escape(st);
    return 0;
}

// DoDeepContact (TriangleConvexPairTester.cs:300-527)
int tri_deep(const ConvexShapeRef& cvx, float min_radius, float core_scaling, const TriangleRef& t,
             TriTesterState* st, ContactData out[2]) {
    const ConvexShapeRef tri = triangle_shape(t);
    const int sided = t.sidedness;
    V3 center = add(add(t.v[0], t.v[1]), t.v[2]);
    center = mul(center, 1.f / 3.f);
    int n_out = 0;
    ContactData c;
    std::memset(&c, 0, sizeof(c));
    if (mpr_overlapping_ray(cvx, tri, center, kIdentityRT)) {
        V3 tn = cross(sub(t.v[1], t.v[0]), sub(t.v[2], t.v[0]));
        float ls = length_sq(tn);
        if (ls < kEpsilon * .01f) {
            surface_cast(cvx, tri, kIdentityRT, center, true, &c.depth, &c.normal, &c.pos);
        } else {
            tn = divide(tn, std::sqrt(ls));
            V3 AO = sub(center, t.v[0]), BO = sub(center, t.v[1]), CO = sub(center, t.v[2]);
            V3 AB = sub(t.v[1], t.v[0]), BC = sub(t.v[2], t.v[1]), CA = sub(t.v[0], t.v[2]);
            float d = dot(AO, AB);
            V3 ABn = mul(AB, d / length_sq(AB));
            ABn = normalize(sub(AO, ABn));
            d = dot(BO, BC);
            V3 BCn = mul(BC, d / length_sq(BC));
            BCn = normalize(sub(BO, BCn));
            d = dot(CO, CA);
            V3 CAn = mul(CA, d / length_sq(CA));
            CAn = normalize(sub(CO, CAn));
            float cd;
            V3 cn;
            tri_cast_corrected(cvx, tri, sided, tn, ABn, &c.depth, &c.normal, &c.depth, &c.normal);
            tri_cast_corrected(cvx, tri, sided, tn, BCn, &cd, &cn, &c.depth, &c.normal);
            if (cd < c.depth) {
                c.normal = cn;
                c.depth = cd;
            }
            tri_cast_corrected(cvx, tri, sided, tn, CAn, &cd, &cn, &c.depth, &c.normal);
            if (cd < c.depth) {
                c.normal = cn;
                c.depth = cd;
            }
            if (sided != SIDED_CLOCKWISE) {
                surface_cast(cvx, tri, kIdentityRT, tn, false, &cd, &cn, nullptr);
                if (cd < c.depth) {
                    c.normal = cn;
                    c.depth = cd;
                }
            }
            if (sided != SIDED_COUNTERCLOCKWISE) {
                tn = neg(tn);
                surface_cast(cvx, tri, kIdentityRT, tn, false, &cd, &cn, nullptr);
                if (cd < c.depth) {
                    c.normal = cn;
                    c.depth = cd;
                }
            }
        }
        refine_penetration(cvx, tri, kIdentityRT, c.depth, c.normal, &c.depth, &c.normal, &c.pos);
        bool skip = false;
        if (sided != SIDED_DOUBLE && dot(tn, c.normal) < 0) skip = true;
        if (!skip) {
            c.id = -1;
            if (c.depth < cvx.margin + t.margin) st->state = TT_NEAR;
            out[n_out++] = c;
        }
    }
    if (tri_inner_sphere(min_radius, core_scaling, t, &c)) out[n_out++] = c;
    if (n_out > 0) return n_out;
    st->state = TT_SEPARATED;
    return 0;
}

// DoPlaneTest (TriangleConvexPairTester.cs:48-211)
int tri_plane(const ConvexShapeRef& cvx, float min_radius, float core_scaling, const TriangleRef& t,
              TriTesterState* st, ContactData out[2]) {
    V3 ab = sub(t.v[1], t.v[0]), ac = sub(t.v[2], t.v[0]);
    V3 rn = cross(ac, ab);
    float dotA = dot(t.v[0], rn);
    if (t.sidedness == SIDED_DOUBLE) {
        if (dotA < 0) {
            rn = neg(rn);
            dotA = -dotA;
        }
    } else if (t.sidedness == SIDED_COUNTERCLOCKWISE) {
        rn = neg(rn);
        dotA = -dotA;
    }
    V3 extreme = support_local(cvx, rn);
    if (tri_voronoi(t, extreme) != VR_ABC) {
        st->state = TT_SEPARATED;
        return tri_separated(cvx, min_radius, core_scaling, t, st, out);
    }
    float dotE = dot(extreme, rn);
    float tt = (dotA - dotE) / length_sq(rn);
    V3 offset = mul(rn, tt);
    float dsq = length_sq(offset);
    float marginSum = t.margin + cvx.margin;
    if (tt <= 0 || dsq < marginSum * marginSum) {
        ContactData c;
        std::memset(&c, 0, sizeof(c));
        if (marginSum > kEpsilon) c.pos = mul(offset, cvx.margin / marginSum);
        else c.pos = kZero;
        c.pos = add(extreme, c.pos);
        float nl = length(rn);
        c.normal = divide(rn, nl);
        float distance = nl * tt;
        c.depth = marginSum - distance;
        int n_out = 0;
        if (c.depth > marginSum) {
            ContactData alt;
            if (tri_inner_sphere(min_radius, core_scaling, t, &alt)) out[n_out++] = alt;
            int previous = st->state;
            st->state = TT_NEAR;
            ContactData alts[2];
            if (tri_near(cvx, min_radius, core_scaling, t, st, alts) > 0) {
                alt = alts[0];
                if (alt.depth + .01f < c.depth) {
                    out[n_out++] = alt;
                } else {
                    out[n_out++] = c;
                    st->state = previous;
                }
            } else {
                st->state = previous;
                return 0;
            }
        } else {
            out[n_out++] = c;
        }
        return n_out;
    }
    return 0;
}

}  // namespace

int triangle_convex_candidates(const ConvexShapeRef& convex, float min_radius, float core_scaling,
                               const TriangleRef& t, TriTesterState* st, ContactData out[2]) {
    switch (st->state) {
    case TT_PLANE: return tri_plane(convex, min_radius, core_scaling, t, st, out);
    case TT_SEPARATED: return tri_separated(convex, min_radius, core_scaling, t, st, out);
    case TT_NEAR: return tri_near(convex, min_radius, core_scaling, t, st, out);
    case TT_DEEP: return tri_deep(convex, min_radius, core_scaling, t, st, out);
    default: return 0;
    }
}

int triangle_convex_region(const TriangleRef& t, const ContactData& c) {
    float dotA = -dot(t.v[0], c.normal), dotB = -dot(t.v[1], c.normal), dotC = -dot(t.v[2], c.normal);
    const float faceEpsilon = .01f, edgeEpsilon = .01f;
    // Extreme vertex e, second vertex s, third vertex o: a face normal when
    // the extreme and third dots nearly tie, else the edge (e, s) when the
    // normal is nearly perpendicular to it, else the vertex e.
    auto edge_or_vertex = [&](float de, float dthird, const V3& ve, const V3& vs, int edge_region, int vertex_region) {
        if (std::fabs(de - dthird) < faceEpsilon) return (int)VR_ABC;
        V3 e = sub(vs, ve);
        float ed = dot(e, c.normal);
        if (ed * ed < length_sq(e) * edgeEpsilon) return edge_region;
        return vertex_region;
    };
    if (dotA > dotB && dotA > dotC) {
        if (dotB > dotC) return edge_or_vertex(dotA, dotC, t.v[0], t.v[1], VR_AB, VR_A);
        return edge_or_vertex(dotA, dotB, t.v[0], t.v[2], VR_AC, VR_A);
    } else if (dotB > dotC) {
        if (dotC > dotA) return edge_or_vertex(dotB, dotA, t.v[1], t.v[2], VR_BC, VR_B);
        return edge_or_vertex(dotB, dotC, t.v[1], t.v[0], VR_AB, VR_B);
    } else {
        if (dotA > dotB) return edge_or_vertex(dotC, dotB, t.v[2], t.v[0], VR_AC, VR_C);
        return edge_or_vertex(dotC, dotA, t.v[2], t.v[1], VR_BC, VR_C);
    }
}

int triangle_sphere_candidates(float radius, const TriangleRef& t, ContactData* out, int* region) {
    std::memset(out, 0, sizeof(*out));
    V3 n = cross(sub(t.v[1], t.v[0]), sub(t.v[2], t.v[0]));
    if (length_sq(n) < kEpsilon * .01f) {
        n = add(add(t.v[0], t.v[1]), t.v[2]);
        n = mul(n, 1 / 3.f);
        if (length_sq(n) < kEpsilon * .01f) n = kUp;
    }
    float d = dot(n, t.v[0]);
    if (t.sidedness == SIDED_DOUBLE) {
        if (d < 0) n = neg(n);
    } else if (t.sidedness == SIDED_CLOCKWISE) {
        if (d > 0) return 0;
    } else {
        if (d < 0) return 0;
    }
    V3 closest;
    *region = closest_point_on_triangle(t.v[0], t.v[1], t.v[2], kZero, &closest);
    float lsq = length_sq(closest);
    float marginSum = t.margin + radius;
    if (lsq <= marginSum * marginSum) {
        if (lsq < kEpsilon) {
            out->normal = normalize(neg(n));
            out->depth = marginSum;
            return 1;
        }
        lsq = std::sqrt(lsq);
        out->normal = divide(closest, lsq);
        out->depth = marginSum - lsq;
        out->pos = closest;
        return 1;
    }
    return 0;
}

}  // namespace ref

// ---------------------------------------------------------------------------
// C entry points for the tests: single queries on two posed shapes.
// A shape is {kind, d0, d1, d2, margin}: box half extents (margin included),
// capsule half length in d0, hull vertices in `verts` (n of them, xyz).
// A pose is {px, py, pz, qx, qy, qz, qw}.

namespace {

ref::ConvexShapeRef make_shape(const float* s, const float* verts, int n, std::vector<ref::V3>* store) {
    ref::ConvexShapeRef r;
    std::memset(&r, 0, sizeof(r));
    r.kind = (int)s[0];
    r.half = ref::V3{s[1], s[2], s[3]};
    r.half_length = s[1];
    r.margin = s[4];
    if (r.kind == BEPU_SHAPE_CONVEX_HULL) {
        store->resize(n);
        for (int i = 0; i < n; ++i) (*store)[i] = ref::V3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        r.verts = store->data();
        r.nverts = n;
    }
    return r;
}
ref::RT make_pose(const float* p) { return ref::RT{ref::V3{p[0], p[1], p[2]}, ref::Q4{p[3], p[4], p[5], p[6]}}; }
void put_contact(const ref::ContactData& c, float* out) {
    out[0] = c.pos.x; out[1] = c.pos.y; out[2] = c.pos.z;
    out[3] = c.normal.x; out[4] = c.normal.y; out[5] = c.normal.z;
    out[6] = c.depth;
    out[7] = (float)c.id;
}

}  // namespace

extern "C" {

int refcpu_gjk_intersecting(const float* sa, const float* va, int na, const float* pa, const float* sb,
                            const float* vb, int nb, const float* pb, float* axis_io) {
    std::vector<ref::V3> ha, hb;
    ref::V3 axis{axis_io[0], axis_io[1], axis_io[2]};
    bool r = ref::gjk_intersecting(make_shape(sa, va, na, &ha), make_shape(sb, vb, nb, &hb), make_pose(pa),
                                   make_pose(pb), &axis);
    axis_io[0] = axis.x; axis_io[1] = axis.y; axis_io[2] = axis.z;
    return r ? 1 : 0;
}

int refcpu_gjk_closest(const float* sa, const float* va, int na, const float* pa, const float* sb, const float* vb,
                       int nb, const float* pb, float* out6) {
    std::vector<ref::V3> ha, hb;
    ref::V3 ca, cb;
    bool r = ref::gjk_closest_points(make_shape(sa, va, na, &ha), make_shape(sb, vb, nb, &hb), make_pose(pa),
                                     make_pose(pb), &ca, &cb);
    out6[0] = ca.x; out6[1] = ca.y; out6[2] = ca.z;
    out6[3] = cb.x; out6[4] = cb.y; out6[5] = cb.z;
    return r ? 1 : 0;
}

int refcpu_mpr_contact(const float* sa, const float* va, int na, const float* pa, const float* sb, const float* vb,
                       int nb, const float* pb, const float* axis, float* out8) {
    std::vector<ref::V3> ha, hb;
    ref::ContactData c;
    bool r = ref::mpr_contact(make_shape(sa, va, na, &ha), make_shape(sb, vb, nb, &hb), make_pose(pa), make_pose(pb),
                              ref::V3{axis[0], axis[1], axis[2]}, &c);
    put_contact(c, out8);
    return r ? 1 : 0;
}

// One GenerateContactCandidate call; state_io / sep_axis_io / local_dir_io carry
// the tester state across calls.
int refcpu_convex_candidate(const float* sa, const float* va, int na, const float* pa, const float* sb,
                            const float* vb, int nb, const float* pb, const float* vel_a, const float* vel_b,
                            int* state_io, float* sep_axis_io, float* local_dir_io, float* out8) {
    std::vector<ref::V3> ha, hb;
    ref::ConvexTesterState st;
    st.state = *state_io;
    st.separating_axis = ref::V3{sep_axis_io[0], sep_axis_io[1], sep_axis_io[2]};
    st.local_direction = ref::V3{local_dir_io[0], local_dir_io[1], local_dir_io[2]};
    ref::ContactData c;
    bool r = ref::convex_candidate(make_shape(sa, va, na, &ha), make_shape(sb, vb, nb, &hb), make_pose(pa),
                                   make_pose(pb), ref::V3{vel_a[0], vel_a[1], vel_a[2]},
                                   ref::V3{vel_b[0], vel_b[1], vel_b[2]}, &st, &c);
    *state_io = st.state;
    sep_axis_io[0] = st.separating_axis.x; sep_axis_io[1] = st.separating_axis.y; sep_axis_io[2] = st.separating_axis.z;
    local_dir_io[0] = st.local_direction.x; local_dir_io[1] = st.local_direction.y; local_dir_io[2] = st.local_direction.z;
    put_contact(c, out8);
    return r ? 1 : 0;
}

// The casts on one posed shape: `kind` 0 SphereCast, 1 RayCast, 2
// CCDSphereCast (with `scaling`); ray origin o3 and direction d3; returns
// hit (1) or not (0), *t the hit parameter either way.
int refcpu_cast(int kind, const float* s, const float* verts, int n, const float* pose, const float* o3,
                const float* d3, float radius, float max_length, float scaling, float* t) {
    std::vector<ref::V3> h;
    ref::ConvexShapeRef shape = make_shape(s, verts, n, &h);
    ref::V3 o{o3[0], o3[1], o3[2]}, d{d3[0], d3[1], d3[2]};
    bool r;
    if (kind == 0) r = ref::sphere_cast(o, d, radius, shape, make_pose(pose), max_length, t);
    else if (kind == 1) r = ref::ray_cast(o, d, shape, make_pose(pose), max_length, t);
    else r = ref::ccd_sphere_cast(o, d, radius, shape, make_pose(pose), max_length, scaling, t);
    return r ? 1 : 0;
}

// TriangleConvexPairTester.GenerateContactCandidates on a triangle already in
// the convex's local space (tri9: three vertices); state_io = {state,
// escapes} and axis_io carry the tester across calls; out16: two contacts of
// 8 floats.  Returns the number of candidates.
int refcpu_triangle_convex(const float* s, const float* verts, int n, float min_radius, float core_scaling,
                           const float* tri9, int sidedness, float tri_margin, int* state_io, float* axis_io,
                           float* out16) {
    std::vector<ref::V3> h;
    ref::TriangleRef t;
    for (int k = 0; k < 3; ++k) t.v[k] = ref::V3{tri9[3 * k], tri9[3 * k + 1], tri9[3 * k + 2]};
    t.sidedness = sidedness;
    t.margin = tri_margin;
    ref::TriTesterState st{state_io[0], state_io[1], ref::V3{axis_io[0], axis_io[1], axis_io[2]}};
    ref::ContactData c[2];
    std::memset(c, 0, sizeof(c));
    int r = ref::triangle_convex_candidates(make_shape(s, verts, n, &h), min_radius, core_scaling, t, &st, c);
    state_io[0] = st.state;
    state_io[1] = st.escapes;
    axis_io[0] = st.axis.x; axis_io[1] = st.axis.y; axis_io[2] = st.axis.z;
    put_contact(c[0], out16);
    put_contact(c[1], out16 + 8);
    return r;
}

// TriangleConvexPairTester.GetRegion of a contact normal (normal3).
int refcpu_triangle_region(const float* tri9, const float* normal3) {
    ref::TriangleRef t;
    for (int k = 0; k < 3; ++k) t.v[k] = ref::V3{tri9[3 * k], tri9[3 * k + 1], tri9[3 * k + 2]};
    t.sidedness = 0;
    t.margin = 0;
    ref::ContactData c;
    std::memset(&c, 0, sizeof(c));
    c.normal = ref::V3{normal3[0], normal3[1], normal3[2]};
    return ref::triangle_convex_region(t, c);
}

// TriangleSpherePairTester.GenerateContactCandidates; *region the Voronoi region.
int refcpu_triangle_sphere(float radius, const float* tri9, int sidedness, float tri_margin, float* out8,
                           int* region) {
    ref::TriangleRef t;
    for (int k = 0; k < 3; ++k) t.v[k] = ref::V3{tri9[3 * k], tri9[3 * k + 1], tri9[3 * k + 2]};
    t.sidedness = sidedness;
    t.margin = tri_margin;
    ref::ContactData c;
    int r = ref::triangle_sphere_candidates(radius, t, &c, region);
    put_contact(c, out8);
    return r;
}

// Toolbox.GetClosestPointOnTriangleToPoint: returns the region, out3 the point.
int refcpu_closest_point_triangle(const float* tri9, const float* p3, float* out3) {
    ref::V3 v[3];
    for (int k = 0; k < 3; ++k) v[k] = ref::V3{tri9[3 * k], tri9[3 * k + 1], tri9[3 * k + 2]};
    ref::V3 o;
    int r = ref::closest_point_on_triangle(v[0], v[1], v[2], ref::V3{p3[0], p3[1], p3[2]}, &o);
    out3[0] = o.x; out3[1] = o.y; out3[2] = o.z;
    return r;
}

}  // extern "C"
