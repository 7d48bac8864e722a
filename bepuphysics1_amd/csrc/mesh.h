// mesh.h -- static meshes on the device: the mesh shape tables, the
// per-pair tester table, and the two triangle pair testers
// (TriangleConvexPairTester, TriangleSpherePairTester).  A triangle is tested
// in the local space of the convex it meets (the convex is shape A at the
// identity, the triangle shape B at the identity), as
// TriangleMeshConvexContactManifold.Update hands it over.  Every expression
// keeps the reference's operand order.
#pragma once
#include "convex.h"
#include "world.h"

namespace bepu {

// TriangleSidedness (BEPUutilities/TriangleSidedness.cs)
enum { SIDED_DOUBLE = 0, SIDED_CLOCKWISE = 1, SIDED_COUNTERCLOCKWISE = 2 };
// VoronoiRegion (BEPUutilities/VoronoiRegion.cs)
enum { VR_A = 0, VR_B, VR_C, VR_AB, VR_AC, VR_BC, VR_ABC };
// TriangleConvexPairTester.CollisionState (TriangleConvexPairTester.cs:691-697)
enum { TT_PLANE = 0, TT_SEPARATED = 1, TT_NEAR = 2, TT_DEEP = 3 };

// One entry of a mesh pair's tester table (activePairTesters,
// TriangleMeshConvexContactManifold.cs:19-20): the triangle's vertex indices
// and its TriangleConvexPairTester state.
struct MeshTester {
    int32_t ia, ib, ic, state;
    int32_t escapes;
    float ax, ay, az;  // localSeparatingAxis
};
static_assert(sizeof(MeshTester) == 32, "MeshTester layout");

// Per-pair capacities of the mesh manifold update.  A step that needs more
// sets the overflow flag (the sticky capacity error).
constexpr int kMeshTesters = 32;    // live triangle testers of one mesh pair
constexpr int kMeshCandidates = 32; // contact candidates and boundary contacts per step
constexpr int kMeshBlocked = 64;    // blocked vertices and blocked edges per step
constexpr int kMeshStack = 64;      // depth of the tree walk (the host checks each tree)

// A mesh shape's record in DeviceWorld::mesh_info (4 float4): the rows of
// its world transform's linear part with the vertex, index and node bases
// in .w, and its translation with the triangle count in .w.
struct MeshInfo {
    m3 m;
    f3 t;
    uint32_t vbase, ibase, nbase, ntri;
};
BEPU_D MeshInfo mesh_info(const DeviceWorld& w, uint32_t shape) {
    const float4 r0 = w.mesh_info[4 * shape], r1 = w.mesh_info[4 * shape + 1], r2 = w.mesh_info[4 * shape + 2],
                 r3 = w.mesh_info[4 * shape + 3];
    MeshInfo i;
    i.m = m3{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r2.x, r2.y, r2.z};
    i.t = f3_of(r3);
    i.vbase = __float_as_uint(r0.w);
    i.ibase = __float_as_uint(r1.w);
    i.nbase = __float_as_uint(r2.w);
    i.ntri = __float_as_uint(r3.w);
    return i;
}

// AffineTransform.Transform (BEPUutilities/AffineTransform.cs:148-152)
BEPU_D f3 affine_transform(f3 p, const m3& m, f3 t) { return mat_transform(p, m) + t; }

// A terrain shape's record (6 float4 in terrain_info).
struct TerrainInfo {
    m3 m;
    f3 t;
    uint32_t hbase, width, length, organization;
};
BEPU_D TerrainInfo terrain_info(const DeviceWorld& w, uint32_t shape) {
    const float4* r = w.terrain_info + 6 * (size_t)shape;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    TerrainInfo i;
    i.m = m3{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r2.x, r2.y, r2.z};
    i.t = f3_of(r3);
    i.hbase = __float_as_uint(r0.w);
    i.width = __float_as_uint(r1.w);
    i.length = __float_as_uint(r2.w);
    i.organization = __float_as_uint(r3.w);
    return i;
}
BEPU_D float terrain_height(const DeviceWorld& w, const TerrainInfo& ti, int column, int row) {
    return w.terrain_heights[ti.hbase + (size_t)column * ti.length + row];
}

// TerrainShape.GetLocalIndices (TerrainShape.cs:626-663): the (column, row)
// of the three vertices of triangle `tri` (2 per quad, quads row by row).
BEPU_D void terrain_local_indices(const TerrainInfo& ti, int tri, int col[3], int row[3]) {
    const int quad = tri / 2;
    const int r = quad / (int)ti.width;
    const int c = quad - r * (int)ti.width;
    const bool blur = ti.organization == 0;  // BottomLeftUpperRight
    if ((tri & 1) == 0) {
        col[0] = c; row[0] = r;
        col[1] = c + 1; row[1] = r;
        col[2] = blur ? c : c + 1; row[2] = r + 1;
    } else if (blur) {
        col[0] = c; row[0] = r + 1;
        col[1] = c + 1; row[1] = r;
        col[2] = c + 1; row[2] = r + 1;
    } else {
        col[0] = c; row[0] = r;
        col[1] = c + 1; row[1] = r + 1;
        col[2] = c; row[2] = r + 1;
    }
}

// TerrainShape.GetPosition (TerrainShape.cs:432-452): clamped indices, world space.
BEPU_D f3 terrain_position(const DeviceWorld& w, const TerrainInfo& ti, int c, int r) {
    if (c <= 0) c = 0;
    else if (c >= (int)ti.width) c = (int)ti.width - 1;
    if (r <= 0) r = 0;
    else if (r >= (int)ti.length) r = (int)ti.length - 1;
    return affine_transform(f3{(float)c, terrain_height(w, ti, c, r), (float)r}, ti.m, ti.t);
}

// TerrainShape.GetTriangle (TerrainShape.cs:607-622) with GetFirstTriangle /
// GetSecondTriangle (:557-604): the world-space triangle.
BEPU_D void terrain_world_triangle(const DeviceWorld& w, const TerrainInfo& ti, int tri, f3 v[3]) {
    const int quad = tri / 2;
    const int r = quad / (int)ti.width;
    const int c = quad - r * (int)ti.width;
    const bool blur = ti.organization == 0;
    if ((tri & 1) == 0) {
        v[0] = terrain_position(w, ti, c, r);
        v[1] = terrain_position(w, ti, c + 1, r);
        v[2] = blur ? terrain_position(w, ti, c, r + 1) : terrain_position(w, ti, c + 1, r + 1);
    } else if (blur) {
        v[0] = terrain_position(w, ti, c, r + 1);
        v[1] = terrain_position(w, ti, c + 1, r + 1);
        v[2] = terrain_position(w, ti, c + 1, r);
    } else {
        v[0] = terrain_position(w, ti, c, r);
        v[1] = terrain_position(w, ti, c, r + 1);
        v[2] = terrain_position(w, ti, c + 1, r + 1);
    }
}

// TerrainShape.GetOverlaps (TerrainShape.cs:497-543): the quads whose cells
// the local box covers (bounds truncated towards zero) and whose four
// heights reach its vertical range, columns outer, rows inner; `visit(tri)`
// receives both triangles of each (2q, 2q + 1 for q = 2 (i + j width)).
template <class F>
BEPU_D void terrain_overlaps(const DeviceWorld& w, const TerrainInfo& ti, f3 lmin, f3 lmax, F&& visit) {
    const int width = (int)ti.width;
    const int minX = max((int)lmin.x, 0), minY = max((int)lmin.z, 0);
    const int maxX = min((int)lmax.x, width - 2), maxY = min((int)lmax.z, (int)ti.length - 2);
    for (int i = minX; i <= maxX; i++)
        for (int j = minY; j <= maxY; j++) {
            const float y1 = terrain_height(w, ti, i, j), y2 = terrain_height(w, ti, i + 1, j),
                        y3 = terrain_height(w, ti, i, j + 1), y4 = terrain_height(w, ti, i + 1, j + 1);
            float highest = y1, lowest = y1;
            if (y2 > highest) highest = y2; else if (y2 < lowest) lowest = y2;
            if (y3 > highest) highest = y3; else if (y3 < lowest) lowest = y3;
            if (y4 > highest) highest = y4; else if (y4 < lowest) lowest = y4;
            if (lmax.y < lowest || lmin.y > highest) continue;
            const int quad = (i + j * width) << 1;
            visit(quad);
            visit(quad | 1);
        }
}

// Toolbox.GetTriangleBoundingBox (BEPUutilities/Toolbox.cs:1516-1580), one axis.
BEPU_D void tri_axis(float a, float b, float c, float* mn, float* mx) {
    if (a > b && a > c) {
        *mx = a;
        *mn = b > c ? c : b;
    } else if (b > c) {
        *mx = b;
        *mn = a > c ? c : a;
    } else {
        *mx = c;
        *mn = a > b ? b : a;
    }
}

// Toolbox.GetClosestPointOnTriangleToPoint (BEPUutilities/Toolbox.cs:338-419)
BEPU_D int closest_on_triangle(f3 a, f3 b, f3 c, f3 p, f3* out) {
    f3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0 && d2 < 0) {
        *out = a;
        return VR_A;
    }
    f3 bp = p - b;
    float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
        *out = b;
        return VR_B;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        float v = d1 / (d1 - d3);
        *out = ab * v + a;
        return VR_AB;
    }
    f3 cp = p - c;
    float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0 && d5 <= d6) {
        *out = c;
        return VR_C;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        float wt = d2 / (d2 - d6);
        *out = ac * wt + a;
        return VR_AC;
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        float wt = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *out = (c - b) * wt + b;
        return VR_BC;
    }
    float denom = 1 / (va + vb + vc);
    float v = vb * denom, wt = vc * denom;
    *out = (a + ab * v) + ac * wt;
    return VR_ABC;
}

// One contact candidate (ContactData): position, normal, depth, id.
struct MeshContact {
    f3 pos, normal;
    float depth;
    int id;
};

// What a triangle test needs of the convex: its shape, MinimumRadius and
// MotionSettings.CoreShapeScaling.
struct TriConvex {
    cvx::Shape shape;
    float min_radius;
    float core_scaling;
};

namespace tri {

// TriangleConvexPairTester.GetVoronoiRegion (TriangleConvexPairTester.cs:619-686)
BEPU_D int voronoi(const cvx::TriShape& t, f3 p) {
    f3 ab = t.v1 - t.v0, ac = t.v2 - t.v0, ap = p - t.v0;
    float APdotAB = dot3(ap, ab), APdotAC = dot3(ap, ac);
    if (APdotAC <= 0.f && APdotAB <= 0) return VR_A;
    f3 bp = p - t.v1;
    float BPdotAB = dot3(ab, bp), BPdotAC = dot3(ac, bp);
    if (BPdotAB >= 0.f && BPdotAC <= BPdotAB) return VR_B;
    float vc = APdotAB * BPdotAC - BPdotAB * APdotAC;
    if (vc <= 0 && APdotAB > 0 && BPdotAB < 0) return VR_AB;
    f3 cp = p - t.v2;
    float CPdotAB = dot3(ab, cp), CPdotAC = dot3(ac, cp);
    if (CPdotAC >= 0.f && CPdotAB <= CPdotAC) return VR_C;
    float vb = CPdotAB * APdotAC - APdotAB * CPdotAC;
    if (vb <= 0.f && APdotAC > 0.f && CPdotAC < 0.f) return VR_AC;
    float va = BPdotAB * CPdotAC - CPdotAB * BPdotAC;
    if (va <= 0.f && (BPdotAC - BPdotAB) > 0.f && (CPdotAB - CPdotAC) > 0.f) return VR_BC;
    return VR_ABC;
}

constexpr int kEscapeAttemptPeriod = 10;  // TriangleConvexPairTester.cs:16

// TryToEscape, both overloads (TriangleConvexPairTester.cs:529-545)
BEPU_D void escape(MeshTester& st) {
    if (++st.escapes == kEscapeAttemptPeriod) {
        st.escapes = 0;
        st.state = TT_PLANE;
    }
}
BEPU_D void escape_at(MeshTester& st, const cvx::TriShape& t, f3 position) {
    if (++st.escapes == kEscapeAttemptPeriod && voronoi(t, position) == VR_ABC) {
        st.escapes = 0;
        st.state = TT_PLANE;
    }
}

// TryInnerSphereContact (TriangleConvexPairTester.cs:548-613)
BEPU_D bool inner_sphere(const TriConvex& cv, const cvx::TriShape& t, int sided, MeshContact* c) {
    *c = MeshContact{f3{0, 0, 0}, f3{0, 0, 0}, 0, 0};
    f3 closest;
    closest_on_triangle(t.v0, t.v1, t.v2, f3{0, 0, 0}, &closest);
    float length = lensq3(closest);
    float minimumRadius = cv.min_radius * (cv.core_scaling + .01f);
    if (length < minimumRadius * minimumRadius) {
        f3 n = cross3(t.v1 - t.v0, t.v2 - t.v0);
        float d = dot3(closest, n);
        if ((sided == SIDED_CLOCKWISE && d > 0) || (sided == SIDED_COUNTERCLOCKWISE && d < 0)) return false;
        length = sqrtf(length);
        c->pos = closest;
        if (length > cvx::kEpsilon) {
            c->normal = divide3(closest, length);
        } else {
            float nl = lensq3(n);
            if (lensq3(n) > cvx::kEpsilon) {
                n = divide3(n, sqrtf(nl));
                c->normal = sided == SIDED_CLOCKWISE ? n : -n;
            } else {
                *c = MeshContact{f3{0, 0, 0}, f3{0, 0, 0}, 0, 0};
                return false;
            }
        }
        c->depth = cv.min_radius - length;
        c->id = -1;
        return true;
    }
    return false;
}

const cvx::RT kIdentityRT{f3{0, 0, 0}, q4{0, 0, 0, 1}};

// DoDeepContact (TriangleConvexPairTester.cs:300-527).  Its surface casts
// (along the three edge normals, with a one-sided triangle's correction, then
// along both face normals, then RefinePenetration's casts along the current
// normal) run through one call site.
__device__ __forceinline__ int deep(const TriConvex& cv, const cvx::TriShape& t, int sided, MeshTester& st,
                                 MeshContact out[2]) {
    f3 center = (t.v0 + t.v1) + t.v2;
    center = center * (1.f / 3.f);
    int n_out = 0;
    MeshContact c{f3{0, 0, 0}, f3{0, 0, 0}, 0, 0};
    if (cvx::mpr_overlapping_ray(cv.shape, t, center, kIdentityRT)) {
        f3 tn = cross3(t.v1 - t.v0, t.v2 - t.v0);
        float ls = lensq3(tn);
        const bool degenerate = ls < cvx::kEpsilon * .01f;
        f3 dirs[5];
        if (!degenerate) {
            tn = divide3(tn, sqrtf(ls));
            f3 AO = center - t.v0, BO = center - t.v1, CO = center - t.v2;
            f3 AB = t.v1 - t.v0, BC = t.v2 - t.v1, CA = t.v0 - t.v2;
            float d = dot3(AO, AB);
            f3 ABn = AB * (d / lensq3(AB));
            dirs[0] = normalize3(AO - ABn);
            d = dot3(BO, BC);
            f3 BCn = BC * (d / lensq3(BC));
            dirs[1] = normalize3(BO - BCn);
            d = dot3(CO, CA);
            f3 CAn = CA * (d / lensq3(CA));
            dirs[2] = normalize3(CO - CAn);
            dirs[3] = tn;
            dirs[4] = -tn;
        }
        // phase 0..4: the casts of the search (phase 0 alone when degenerate,
        // along the centre offset with the position); 5: refinement.
        int phase = degenerate ? -1 : 0, refinements = 0;
        while (true) {
            if (phase == 3 && sided == SIDED_CLOCKWISE) phase = 4;
            if (phase == 4 && sided == SIDED_COUNTERCLOCKWISE) phase = 5;
            if (phase == 4) tn = -tn;
            const bool refine = phase == 5;
            const bool with_position = refine || phase == -1;
            const f3 dir = phase == -1 ? center : (refine ? c.normal : dirs[phase]);
            float cd;
            f3 cn, cp;
            cvx::surface_cast(cv.shape, t, kIdentityRT, dir, with_position, &cd, &cn, &cp);
            if (phase == -1) {
                c.depth = cd;
                c.normal = cn;
                c.pos = cp;
                phase = 5;
                continue;
            }
            if (refine) {
                // MPRToolbox.RefinePenetration (MPRToolbox.cs:1282-1322)
                c.pos = cp;
                if (c.depth - cd <= cvx::kDepthRefinementEpsilon || ++refinements >= cvx::kMaxDepthRefinementIterations) {
                    c.depth = cd;
                    break;
                }
                c.depth = cd;
                c.normal = cn;
                continue;
            }
            if (phase <= 2) {
                float d = dot3(tn, cn);
                if ((sided == SIDED_CLOCKWISE && d > 0) || (sided == SIDED_COUNTERCLOCKWISE && d < 0)) {
                    // a normal facing against a one-sided triangle loses its
                    // component along the face normal; the depth follows it
                    f3 previous = cn;
                    d = dot3(cn, tn);
                    f3 p = cn * d;
                    cn = cn - p;
                    float length = lensq3(cn);
                    if (length > cvx::kEpsilon) {
                        cn = divide3(cn, sqrtf(length));
                        d = dot3(cn, previous);
                        cd *= d;
                    } else {
                        // the reference resets the contact here, not the candidate
                        c.depth = FLT_MAX;
                        c.normal = f3{0, 0, 0};
                        if (phase == 0) {
                            cd = c.depth;
                            cn = c.normal;
                        }
                    }
                }
            }
            if (phase == 0 || cd < c.depth) {
                c.normal = cn;
                c.depth = cd;
            }
            phase = phase + 1;
        }
        bool skip = sided != SIDED_DOUBLE && dot3(tn, c.normal) < 0;
        if (!skip) {
            c.id = -1;
            if (c.depth < cv.shape.margin + t.margin) st.state = TT_NEAR;
            out[n_out++] = c;
        }
    }
    if (inner_sphere(cv, t, sided, &c)) out[n_out++] = c;
    if (n_out > 0) return n_out;
    st.state = TT_SEPARATED;
    return 0;
}

// DoExternalNear (TriangleConvexPairTester.cs:230-298)
__device__ __forceinline__ int near(const TriConvex& cv, const cvx::TriShape& t, int sided, MeshTester& st,
                                 MeshContact out[2]) {
    f3 centroid = (t.v0 + t.v1) + t.v2;
    centroid = centroid * .33333333f;
    f3 ca, cb;
    if (cvx::gjk_closest_points(cv.shape, t, kIdentityRT, kIdentityRT, &ca, &cb, centroid)) {
        st.state = TT_DEEP;
        return deep(cv, t, sided, st, out);
    }
    f3 disp = cb - ca;
    float d2 = lensq3(disp);
    float margin = cv.shape.margin + t.margin;
    if (d2 < margin * margin) {
        MeshContact c{f3{0, 0, 0}, f3{0, 0, 0}, 0, 0};
        if (sided != SIDED_DOUBLE) {
            f3 n = cross3(t.v1 - t.v0, t.v2 - t.v0);
            float d = dot3(n, disp);
            if (sided == SIDED_CLOCKWISE && d > 0) return 0;
            if (sided == SIDED_COUNTERCLOCKWISE && d < 0) return 0;
        }
        if (margin > cvx::kEpsilon) c.pos = disp * (cv.shape.margin / margin);
        else c.pos = f3{0, 0, 0};
        c.pos = ca + c.pos;
        c.normal = disp;
        float distance = sqrtf(d2);
        c.normal = divide3(c.normal, distance);
        c.depth = margin - distance;
        out[0] = c;
        escape_at(st, t, c.pos);
        return 1;
    }
    st.state = TT_SEPARATED;
    return 0;
}

// DoExternalSeparated (TriangleConvexPairTester.cs:215-228)
__device__ __forceinline__ int separated(const TriConvex& cv, const cvx::TriShape& t, int sided, MeshTester& st,
                                      MeshContact out[2]) {
    f3 axis{st.ax, st.ay, st.az};
    const bool hit = cvx::gjk_intersecting(cv.shape, t, kIdentityRT, kIdentityRT, &axis);
    st.ax = axis.x;
    st.ay = axis.y;
    st.az = axis.z;
    if (hit) {
        st.state = TT_NEAR;
        return near(cv, t, sided, st, out);
    }
    escape(st);
    return 0;
}

// DoPlaneTest (TriangleConvexPairTester.cs:48-211)
BEPU_D int plane(const TriConvex& cv, const cvx::TriShape& t, int sided, MeshTester& st, MeshContact out[2]) {
    f3 ab = t.v1 - t.v0, ac = t.v2 - t.v0;
    f3 rn = cross3(ac, ab);
    float dotA = dot3(t.v0, rn);
    if (sided == SIDED_DOUBLE) {
        if (dotA < 0) {
            rn = -rn;
            dotA = -dotA;
        }
    } else if (sided == SIDED_COUNTERCLOCKWISE) {
        rn = -rn;
        dotA = -dotA;
    }
    f3 extreme = cvx::support_local(cv.shape, rn);
    if (voronoi(t, extreme) != VR_ABC) {
        st.state = TT_SEPARATED;
        return separated(cv, t, sided, st, out);
    }
    float dotE = dot3(extreme, rn);
    float tt = (dotA - dotE) / lensq3(rn);
    f3 offset = rn * tt;
    float dsq = lensq3(offset);
    float marginSum = t.margin + cv.shape.margin;
    if (tt <= 0 || dsq < marginSum * marginSum) {
        MeshContact c{f3{0, 0, 0}, f3{0, 0, 0}, 0, 0};
        if (marginSum > cvx::kEpsilon) c.pos = offset * (cv.shape.margin / marginSum);
        else c.pos = f3{0, 0, 0};
        c.pos = extreme + c.pos;
        float nl = len3(rn);
        c.normal = divide3(rn, nl);
        float distance = nl * tt;
        c.depth = marginSum - distance;
        int n_out = 0;
        if (c.depth > marginSum) {
            MeshContact alt;
            if (inner_sphere(cv, t, sided, &alt)) out[n_out++] = alt;
            int previous = st.state;
            st.state = TT_NEAR;
            MeshContact alts[2];
            if (near(cv, t, sided, st, alts) > 0) {
                alt = alts[0];
                if (alt.depth + .01f < c.depth) {
                    out[n_out++] = alt;
                } else {
                    out[n_out++] = c;
                    st.state = previous;
                }
            } else {
                st.state = previous;
                return 0;
            }
        } else {
            out[n_out++] = c;
        }
        return n_out;
    }
    return 0;
}

}  // namespace tri

// TriangleConvexPairTester.GenerateContactCandidates (TriangleConvexPairTester.cs:23-46)
BEPU_D int triangle_convex(const TriConvex& cv, const cvx::TriShape& t, int sided, MeshTester& st,
                           MeshContact out[2]) {
    switch (st.state) {
    case TT_PLANE: return tri::plane(cv, t, sided, st, out);
    case TT_SEPARATED: return tri::separated(cv, t, sided, st, out);
    case TT_NEAR: return tri::near(cv, t, sided, st, out);
    case TT_DEEP: return tri::deep(cv, t, sided, st, out);
    default: return 0;
    }
}

// TriangleConvexPairTester.GetRegion (TriangleConvexPairTester.cs:708-897):
// the extreme vertex along the normal, the edge to the second one when the
// normal is nearly perpendicular to it, a face when the extreme and third
// vertices nearly tie.
BEPU_D int triangle_convex_region(const cvx::TriShape& t, f3 normal) {
    float dotA = -dot3(t.v0, normal), dotB = -dot3(t.v1, normal), dotC = -dot3(t.v2, normal);
    const float faceEpsilon = .01f, edgeEpsilon = .01f;
    float de, dthird;
    f3 ve, vs;
    int edge_region, vertex_region;
    if (dotA > dotB && dotA > dotC) {
        de = dotA;
        ve = t.v0;
        vertex_region = VR_A;
        if (dotB > dotC) { dthird = dotC; vs = t.v1; edge_region = VR_AB; }
        else { dthird = dotB; vs = t.v2; edge_region = VR_AC; }
    } else if (dotB > dotC) {
        de = dotB;
        ve = t.v1;
        vertex_region = VR_B;
        if (dotC > dotA) { dthird = dotA; vs = t.v2; edge_region = VR_BC; }
        else { dthird = dotC; vs = t.v0; edge_region = VR_AB; }
    } else {
        de = dotC;
        ve = t.v2;
        vertex_region = VR_C;
        if (dotA > dotB) { dthird = dotB; vs = t.v0; edge_region = VR_AC; }
        else { dthird = dotA; vs = t.v1; edge_region = VR_BC; }
    }
    if (fabsf(de - dthird) < faceEpsilon) return VR_ABC;
    f3 e = vs - ve;
    float ed = dot3(e, normal);
    if (ed * ed < lensq3(e) * edgeEpsilon) return edge_region;
    return vertex_region;
}

// TriangleSpherePairTester.GenerateContactCandidates (TriangleSpherePairTester.cs:20-94):
// at most one candidate; *region the tester's lastRegion.
BEPU_D int triangle_sphere(float radius, const cvx::TriShape& t, int sided, MeshContact* out, int* region) {
    *out = MeshContact{f3{0, 0, 0}, f3{0, 0, 0}, 0, 0};
    f3 n = cross3(t.v1 - t.v0, t.v2 - t.v0);
    if (lensq3(n) < cvx::kEpsilon * .01f) {
        n = (t.v0 + t.v1) + t.v2;
        n = n * (1 / 3.f);
        if (lensq3(n) < cvx::kEpsilon * .01f) n = f3{0, 1, 0};
    }
    float d = dot3(n, t.v0);
    if (sided == SIDED_DOUBLE) {
        if (d < 0) n = -n;
    } else if (sided == SIDED_CLOCKWISE) {
        if (d > 0) return 0;
    } else {
        if (d < 0) return 0;
    }
    f3 closest;
    *region = closest_on_triangle(t.v0, t.v1, t.v2, f3{0, 0, 0}, &closest);
    float lsq = lensq3(closest);
    float marginSum = t.margin + radius;
    if (lsq <= marginSum * marginSum) {
        if (lsq < cvx::kEpsilon) {
            out->normal = normalize3(-n);
            out->depth = marginSum;
            return 1;
        }
        lsq = sqrtf(lsq);
        out->normal = divide3(closest, lsq);
        out->depth = marginSum - lsq;
        out->pos = closest;
        return 1;
    }
    return 0;
}

}  // namespace bepu
