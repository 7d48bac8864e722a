// cast.h -- the GJK casts of continuous collision detection on the device:
// the ray simplex (GJK/RaySimplex.cs), GJKToolbox.SphereCast, RayCast and
// CCDSphereCast (GJK/GJKToolbox.cs:195-495).  One thread runs one cast; every
// expression keeps the reference's operand order so the results are
// bit-identical under -ffp-contract=off with correctly rounded division and
// square root.
#pragma once
#include "convex.h"

namespace bepu {
namespace cvx {

// The simplex of a cast.  The reduction functions run on the simplex shifted
// to the current hit location (`s`) and reduce the unshifted one (`out`).
struct RSimplex {
    f3 A, B, C, D;
    int state;
};

BEPU_D RSimplex rs_empty() {
    RSimplex r;
    r.A = r.B = r.C = r.D = zero3();
    r.state = S_EMPTY;
    return r;
}

// RaySimplex.GetPointOnSegmentClosestToOrigin (RaySimplex.cs:76-115).  Past
// B the reduced simplex is B, and the point returned is the shifted A, as
// the reference returns it.
BEPU_D f3 rs_segment(const RSimplex& s, RSimplex& out) {
    f3 seg = s.B - s.A;
    float dotA = dot3(seg, s.A);
    if (dotA > 0) {
        out.state = S_POINT;
        return s.A;
    }
    float dotB = dot3(seg, s.B);
    if (dotB > 0) {
        float V = -dotA / lensq3(seg);
        return seg * V + s.A;
    }
    out.A = out.B;
    out.state = S_POINT;
    return s.A;
}

// RaySimplex.GetPointOnTriangleClosestToOrigin (RaySimplex.cs:122-240).  In
// the region of C the reduced simplex is C and the point returned is the
// shifted A, as the reference returns it.
BEPU_D f3 rs_triangle(const RSimplex& s, RSimplex& out) {
    f3 ab = s.B - s.A, ac = s.C - s.A;
    float AdotAB = -dot3(ab, s.A), AdotAC = -dot3(ac, s.A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        out.state = S_POINT;
        return s.A;
    }
    float BdotAB = -dot3(ab, s.B), BdotAC = -dot3(ac, s.B);
    if (BdotAB >= 0.f && BdotAC <= BdotAB) {
        out.state = S_POINT;
        out.A = out.B;
        return s.B;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    if (vc <= 0 && AdotAB > 0 && BdotAB < 0) {
        out.state = S_SEGMENT;
        float V = AdotAB / (AdotAB - BdotAB);
        return ab * V + s.A;
    }
    float CdotAB = -dot3(ab, s.C), CdotAC = -dot3(ac, s.C);
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
        return ac * V + s.A;
    }
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        out.state = S_SEGMENT;
        out.A = out.C;
        float U = d3d4 / (d3d4 + d6d5);
        return (s.C - s.B) * U + s.B;
    }
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    f3 point = ab * v;
    f3 acw = ac * w;
    point = s.A + point;
    return point + acw;
}

// RaySimplex.TryTetrahedronTriangle (RaySimplex.cs:320-470): the face ABC
// (shifted), its unshifted vertices sA..sC, and the fourth vertex.
BEPU_D bool rs_try_face(f3 A, f3 B, f3 C, f3 sA, f3 sB, f3 sC, f3 other, RSimplex& cand, f3& point) {
    cand = rs_empty();
    point = zero3();
    f3 ab = B - A, ac = C - A;
    f3 normal = cross3(ab, ac);
    f3 AD = other - A;
    float AdotN = dot3(A, normal), ADdotN = dot3(AD, normal);
    if (!(AdotN * ADdotN >= 0)) return false;
    float AdotAB = -dot3(ab, A), AdotAC = -dot3(ac, A);
    if (AdotAC <= 0.f && AdotAB <= 0) {
        cand.state = S_POINT;
        cand.A = sA;
        point = A;
        return true;
    }
    float BdotAB = -dot3(ab, B), BdotAC = -dot3(ac, B);
    if (BdotAB >= 0.f && BdotAC <= BdotAB) {
        cand.state = S_POINT;
        cand.A = sB;
        point = B;
        return true;
    }
    float vc = AdotAB * BdotAC - BdotAB * AdotAC;
    if (vc <= 0 && AdotAB > 0 && BdotAB < 0) {
        cand.state = S_SEGMENT;
        cand.A = sA;
        cand.B = sB;
        float V = AdotAB / (AdotAB - BdotAB);
        point = ab * V + A;
        return true;
    }
    float CdotAB = -dot3(ab, C), CdotAC = -dot3(ac, C);
    if (CdotAC >= 0.f && CdotAB <= CdotAC) {
        cand.state = S_POINT;
        cand.A = sC;
        point = C;
        return true;
    }
    float vb = CdotAB * AdotAC - AdotAB * CdotAC;
    if (vb <= 0.f && AdotAC > 0.f && CdotAC < 0.f) {
        cand.state = S_SEGMENT;
        cand.A = sA;
        cand.B = sC;
        float V = AdotAC / (AdotAC - CdotAC);
        point = ac * V + A;
        return true;
    }
    float va = BdotAB * CdotAC - CdotAB * BdotAC;
    float d3d4, d6d5;
    if (va <= 0.f && (d3d4 = BdotAC - BdotAB) > 0.f && (d6d5 = CdotAB - CdotAC) > 0.f) {
        cand.state = S_SEGMENT;
        cand.A = sB;
        cand.B = sC;
        float V = d3d4 / (d3d4 + d6d5);
        point = (C - B) * V + B;
        return true;
    }
    cand.state = S_TRIANGLE;
    cand.A = sA;
    cand.B = sB;
    cand.C = sC;
    float denom = 1.f / (va + vb + vc);
    float w = vc * denom, v = vb * denom;
    f3 p = ab * v;
    f3 acw = ac * w;
    p = A + p;
    point = p + acw;
    return true;
}

// RaySimplex.GetPointOnTetrahedronClosestToOrigin (RaySimplex.cs:247-317):
// the closest of the faces DAC, DCB, DBA and ABC that pass their test.
BEPU_D f3 rs_tetrahedron(const RSimplex& s, RSimplex& out) {
    RSimplex best = rs_empty(), cand;
    f3 point = zero3(), cp;
    float best_d = FLT_MAX, d;
    if (rs_try_face(s.A, s.C, s.D, out.A, out.C, out.D, s.B, cand, cp)) {
        point = cp;
        best = cand;
        best_d = lensq3(cp);
    }
    if (rs_try_face(s.C, s.B, s.D, out.C, out.B, out.D, s.A, cand, cp) && (d = lensq3(cp)) < best_d) {
        point = cp;
        best = cand;
        best_d = d;
    }
    if (rs_try_face(s.B, s.A, s.D, out.B, out.A, out.D, s.C, cand, cp) && (d = lensq3(cp)) < best_d) {
        point = cp;
        best = cand;
        best_d = d;
    }
    if (rs_try_face(s.A, s.B, s.C, out.A, out.B, out.C, s.D, cand, cp) && (d = lensq3(cp)) < best_d) {
        point = cp;
        best = cand;
        best_d = d;
    }
    if (best_d < FLT_MAX) out = best;
    return point;
}

// RaySimplex.GetPointClosestToOrigin (RaySimplex.cs:40-69), on the shifted
// simplex with the unshifted one to reduce.
BEPU_D f3 rs_closest(const RSimplex& s, RSimplex& out) {
    switch (s.state) {
    case S_POINT: return s.A;
    case S_SEGMENT: return rs_segment(s, out);
    case S_TRIANGLE: return rs_triangle(s, out);
    case S_TETRAHEDRON: return rs_tetrahedron(s, out);
    default: return zero3();
    }
}

// RaySimplex.AddNewSimplexPoint (RaySimplex.cs:478-520): the simplex grows
// by `p`; returns it shifted to the hit location (hit - vertex).
BEPU_D RSimplex rs_add(RSimplex& s, f3 p, f3 hit) {
    RSimplex sh = rs_empty();
    switch (s.state) {
    case S_EMPTY:
        s.state = S_POINT;
        s.A = p;
        sh.A = hit - s.A;
        break;
    case S_POINT:
        s.state = S_SEGMENT;
        s.B = p;
        sh.A = hit - s.A;
        sh.B = hit - s.B;
        break;
    case S_SEGMENT:
        s.state = S_TRIANGLE;
        s.C = p;
        sh.A = hit - s.A;
        sh.B = hit - s.B;
        sh.C = hit - s.C;
        break;
    case S_TRIANGLE:
        s.state = S_TETRAHEDRON;
        s.D = p;
        sh.A = hit - s.A;
        sh.B = hit - s.B;
        sh.C = hit - s.C;
        sh.D = hit - s.D;
        break;
    default: break;
    }
    sh.state = s.state;
    return sh;
}

// RaySimplex.GetErrorTolerance (RaySimplex.cs:527-560)
BEPU_D float rs_error_tolerance(const RSimplex& s, f3 o) {
    switch (s.state) {
    case S_POINT: return distsq3(s.A, o);
    case S_SEGMENT: return maxf_ref(distsq3(s.A, o), distsq3(s.B, o));
    case S_TRIANGLE: return maxf_ref(distsq3(s.A, o), maxf_ref(distsq3(s.B, o), distsq3(s.C, o)));
    case S_TETRAHEDRON:
        return maxf_ref(distsq3(s.A, o), maxf_ref(distsq3(s.B, o), maxf_ref(distsq3(s.C, o), distsq3(s.D, o))));
    default: return 0;
    }
}

// ConvexShape.GetLocalExtremePoint (ConvexShape.cs:104-113): with the margin.
template <class S>
BEPU_D f3 local_extreme_with_margin(const S& s, f3 d) {
    f3 p = support_local(s, d);
    float ls = lensq3(d);
    if (ls > kEpsilon) {
        d = d * (s.margin / sqrtf(ls));
        p = p + d;
    }
    return p;
}

// GJKToolbox.SphereCast (GJKToolbox.cs:387-459): a ray from `origin` along
// `dir` (lengths in units of dir) swept by a sphere of `radius` against the
// posed shape; t receives the hit parameter, also where the cast fails (0
// after too many iterations); `normal` (optional) RayHit.Normal, the search
// direction of the last advance in world space.
template <class S>
BEPU_D bool sphere_cast(f3 origin, f3 dir, float radius, const S& shape, const RT& xf, float max_length, float& t,
                        f3* normal = nullptr) {
    f3 hn = zero3();
    origin = origin - xf.pos;
    const q4 cj = qconj(xf.q);
    origin = qrotate(origin, cj);
    dir = qrotate(dir, cj);
    float T = 0;
    f3 loc = origin;
    f3 v = loc;
    RSimplex simplex = rs_empty();
    int count = 0;
    while (lensq3(v) >= kEpsilon * rs_error_tolerance(simplex, origin)) {
        if (++count > kMaxGJKIterations) {
            t = 0;
            return false;
        }
        f3 p = support_local(shape, v);
        const float ls = lensq3(v);
        const f3 contribution = ls > kEpsilon ? v * ((shape.margin + radius) / sqrtf(ls)) : zero3();
        p = p + contribution;
        const f3 w = loc - p;
        const float vw = dot3(v, w);
        if (vw > 0) {
            const float vdir = dot3(v, dir);
            T = T - vw / vdir;
            if (vdir >= 0 || T > max_length) {
                t = T;
                return false;
            }
            loc = dir * T + origin;
            hn = v;
        }
        const RSimplex shifted = rs_add(simplex, p, loc);
        v = rs_closest(shifted, simplex);
    }
    t = T;
    if (normal) *normal = qrotate(hn, xf.q);
    return true;
}

// GJKToolbox.RayCast (GJKToolbox.cs:195-265): the ray against the shape with
// its margin.
template <class S>
BEPU_D bool ray_cast(f3 origin, f3 dir, const S& shape, const RT& xf, float max_length, float& t,
                     f3* normal = nullptr) {
    f3 hn = zero3();
    origin = origin - xf.pos;
    const q4 cj = qconj(xf.q);
    origin = qrotate(origin, cj);
    dir = qrotate(dir, cj);
    float T = 0;
    f3 loc = origin;
    f3 off = loc;
    RSimplex simplex = rs_empty();
    int count = 0;
    while (lensq3(off) >= kEpsilon * rs_error_tolerance(simplex, origin)) {
        if (++count > kMaxGJKIterations) {
            t = 0;
            return false;
        }
        const f3 e = local_extreme_with_margin(shape, off);
        const f3 eo = loc - e;
        const float vw = dot3(off, eo);
        if (vw > 0) {
            const float cd = dot3(off, dir);
            if (cd >= 0) {
                t = 0;
                return false;
            }
            T = T - vw / cd;
            if (T > max_length) {
                t = 0;
                return false;
            }
            loc = dir * T + origin;
            hn = off;
        }
        const RSimplex shifted = rs_add(simplex, e, loc);
        off = rs_closest(shifted, simplex);
    }
    t = T;
    if (normal) *normal = qrotate(hn, xf.q);
    return true;
}

// GJKToolbox.CCDSphereCast (GJKToolbox.cs:470-495): the sphere cast, retried
// with the radius scaled by `scaling` while it starts embedded (a hit at 0),
// up to three times, then a plain ray cast.
template <class S>
BEPU_D bool ccd_sphere_cast(f3 origin, f3 dir, float radius, const S& shape, const RT& xf, float max_length,
                            float scaling, float& t, f3* normal = nullptr) {
    for (int iterations = 1;; ++iterations) {
        float T;
        if (sphere_cast(origin, dir, radius, shape, xf, max_length, T, normal) && T > 0) {
            t = T;
            return true;
        }
        if (T > max_length || T < 0) {
            t = T;
            return false;
        }
        radius *= scaling;
        if (iterations > 3) {
            const bool hit = ray_cast(origin, dir, shape, xf, max_length, T, normal);
            t = T;
            return hit && T > 0;
        }
    }
}

}  // namespace cvx
}  // namespace bepu
