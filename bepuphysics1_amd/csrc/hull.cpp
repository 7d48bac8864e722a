// hull.cpp -- host-side construction of a convex hull shape from a point
// cloud: the incremental hull of BEPUutilities.ConvexHullHelper, the volume
// distribution of InertiaHelper.ComputeShapeDistribution, and the recentring
// of ConvexHullShape.ComputeDescription.  Run once per hull shape; the vertex
// order it produces is the order the support mapping scans (ties keep the
// first vertex), so it follows the reference step for step.
#include "hull.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <unordered_set>

namespace bepu {

namespace {

const float kEpsilon = 1e-7f;     // Toolbox.Epsilon
const float kBigEpsilon = 1e-5f;  // Toolbox.BigEpsilon

// RawList.FastRemoveAt (BEPUutilities/DataStructures/RawList.cs:79-91)
inline void fast_remove_at(std::vector<int>& v, size_t i) {
    size_t last = v.size() - 1;
    if (i < last) v[i] = v[last];
    v.pop_back();
}

// ConvexHullHelper.FindNormal (ConvexHullHelper.cs:475-482)
inline f3 find_normal(const std::vector<int>& idx, const std::vector<f3>& p, size_t t) {
    f3 a = p[idx[t]];
    f3 ab = p[idx[t + 1]] - a;
    f3 ac = p[idx[t + 2]] - a;
    return cross3(ac, ab);
}

// ConvexHullHelper.IsTriangleVisibleFromPoint (ConvexHullHelper.cs:484-499)
inline bool visible(const std::vector<int>& idx, const std::vector<f3>& p, size_t t, f3 point) {
    f3 a = p[idx[t]];
    f3 n = cross3(p[idx[t + 2]] - a, p[idx[t + 1]] - a);
    return dot3(point - a, n) >= 0;
}

// ConvexHullHelper.VerifyWindings (ConvexHullHelper.cs:501-514)
void verify_windings(std::vector<int>& idx, const std::vector<f3>& p, f3 centroid) {
    for (size_t k = 0; k < idx.size(); k += 3)
        if (visible(idx, p, k, centroid)) std::swap(idx[k + 1], idx[k + 2]);
}

// ConvexHullHelper.MaintainEdge (ConvexHullHelper.cs:214-238): an edge seen
// twice is interior to the removed region and cancels.
void maintain_edge(int a, int b, std::vector<int>& edges) {
    bool contained = false;
    size_t index = 0;
    for (size_t k = 0; k < edges.size(); k += 2)
        if ((edges[k] == a && edges[k + 1] == b) || (edges[k] == b && edges[k + 1] == a)) {
            contained = true;
            index = k;
        }
    if (!contained) {
        edges.push_back(a);
        edges.push_back(b);
    } else {
        fast_remove_at(edges, index + 1);
        fast_remove_at(edges, index);
    }
}

// ConvexHullHelper.GetExtremePoint (ConvexHullHelper.cs:240-255): index into `outside`.
size_t extreme_outside(f3 dir, const std::vector<f3>& p, const std::vector<int>& outside) {
    float best = -FLT_MAX;
    size_t bi = 0;
    for (size_t i = 0; i < outside.size(); ++i) {
        float d = dot3(p[outside[i]], dir);
        if (d > best) {
            best = d;
            bi = i;
        }
    }
    return bi;
}

// ConvexHullHelper.GetExtremePoints (ConvexHullHelper.cs:257-280)
void extreme_points(f3 dir, const std::vector<f3>& p, float* maxd, float* mind, int* maxi, int* mini) {
    *maxi = 0;
    *mini = 0;
    float d = dot3(p[0], dir);
    *mind = d;
    *maxd = d;
    for (size_t i = 1; i < p.size(); ++i) {
        d = dot3(p[i], dir);
        if (d > *maxd) {
            *maxd = d;
            *maxi = (int)i;
        } else if (d < *mind) {
            *mind = d;
            *mini = (int)i;
        }
    }
}

// ConvexHullHelper.ComputeInitialTetrahedron (ConvexHullHelper.cs:282-438)
bool initial_tetrahedron(const std::vector<f3>& p, std::vector<int>& outside, std::vector<int>& tri, f3* centroid) {
    float minX = FLT_MAX, maxX = -FLT_MAX;
    int minXi = 0, maxXi = 0;
    for (size_t i = 0; i < p.size(); ++i) {
        f3 v = p[i];
        if (v.x > maxX) {
            maxX = v.x;
            maxXi = (int)i;
        } else if (v.x < minX) {
            minX = v.x;
            minXi = (int)i;
        }
    }
    int a = minXi, b = maxXi;
    if (a == b) return false;
    f3 ab = p[b] - p[a];
    f3 dir = cross3(ab, f3{0, 1, 0});
    if (lensq3(dir) < kEpsilon) dir = cross3(ab, f3{1, 0, 0});
    float maxd, mind;
    int maxi, mini;
    extreme_points(dir, p, &maxd, &mind, &maxi, &mini);
    float d = dot3(dir, p[a]);
    int c = std::fabs(d - mind) > std::fabs(d - maxd) ? mini : maxi;
    if (a == c || b == c) return false;
    f3 ac = p[c] - p[a];
    dir = cross3(ab, ac);
    extreme_points(dir, p, &maxd, &mind, &maxi, &mini);
    d = dot3(dir, p[a]);
    int dd = std::fabs(d - mind) > std::fabs(d - maxd) ? mini : maxi;
    if (a == dd || b == dd || c == dd) return false;
    tri = {a, b, c, a, b, dd, a, c, dd, b, c, dd};
    f3 cen = p[a] + p[b];
    cen = cen + p[c];
    cen = cen + p[dd];
    *centroid = cen * 0.25f;
    for (size_t i = 0; i < tri.size(); i += 3) {
        f3 vA = p[tri[i]], vB = p[tri[i + 1]], vC = p[tri[i + 2]];
        f3 cr = cross3(vC - vA, vB - vA);
        float vol = dot3(vA - *centroid, cr);
        if (std::fabs(vol) < kBigEpsilon) return false;
        if (vol < 0) std::swap(tri[i], tri[i + 1]);
    }
    int t4[4] = {a, b, c, dd};
    std::sort(t4, t4 + 4);
    int ti = 0;
    for (size_t i = 0; i < p.size(); ++i) {
        if (ti < 4 && (int)i == t4[ti]) ++ti;
        else outside.push_back((int)i);
    }
    return true;
}

// ConvexHullHelper.RemoveInsidePoints (ConvexHullHelper.cs:440-473)
void remove_inside_points(const std::vector<f3>& p, const std::vector<int>& tri, std::vector<int>& outside) {
    std::vector<int> inside(outside);
    outside.clear();
    for (size_t i = 0; i < tri.size() && !inside.empty(); i += 3) {
        f3 n = find_normal(tri, p, i);
        f3 a = p[tri[i]];
        for (int j = (int)inside.size() - 1; j >= 0; --j) {
            if (dot3(p[inside[j]] - a, n) > 0) {
                outside.push_back(inside[j]);
                fast_remove_at(inside, (size_t)j);
            }
        }
    }
}

// ConvexHullHelper.GetConvexHull (ConvexHullHelper.cs:114-212): triangle list.
bool convex_hull(const std::vector<f3>& p, std::vector<int>& tri) {
    if (p.empty()) return false;
    std::vector<int> outside;
    f3 inside_point;
    if (!initial_tetrahedron(p, outside, tri, &inside_point)) return false;
    remove_inside_points(p, tri, outside);
    std::vector<int> edges, fresh;
    // Every pass consumes one outside point; the bound only guards against a
    // rounding cycle on pathological input.
    size_t guard = 4 * p.size() + 16;
    while (!outside.empty()) {
        if (guard-- == 0) return false;
        for (size_t k = 0; k < tri.size(); k += 3) {
            f3 n = find_normal(tri, p, k);
            size_t oi = extreme_outside(n, p, outside);
            int mi = outside[oi];
            f3 m = p[mi];
            if (dot3(n, m - p[tri[k]]) > 0) {
                fast_remove_at(outside, oi);
                edges.clear();
                for (int t = (int)tri.size() - 3; t >= 0; t -= 3) {
                    if (visible(tri, p, (size_t)t, m)) {
                        maintain_edge(tri[t], tri[t + 1], edges);
                        maintain_edge(tri[t], tri[t + 2], edges);
                        maintain_edge(tri[t + 1], tri[t + 2], edges);
                        fast_remove_at(tri, (size_t)t + 2);
                        fast_remove_at(tri, (size_t)t + 1);
                        fast_remove_at(tri, (size_t)t);
                    }
                }
                for (size_t e = 0; e < edges.size(); e += 2) {
                    fresh.push_back(edges[e]);
                    fresh.push_back(edges[e + 1]);
                    fresh.push_back(mi);
                }
                verify_windings(fresh, p, inside_point);
                tri.insert(tri.end(), fresh.begin(), fresh.end());
                fresh.clear();
                remove_inside_points(p, tri, outside);
                break;
            }
        }
    }
    return true;
}

// InertiaHelper.ComputeShapeDistribution (InertiaHelper.cs:407-503) with
// InertiaHelper.GetPointContribution (InertiaHelper.cs:36-51) for the shift
// to the centroid.
void shape_distribution(const std::vector<f3>& v, const std::vector<int>& tri, f3* center, float* volume, m3* vd) {
    float a = 0, b = 0, c = 0, ao = 0, bo = 0, co = 0;
    f3 summed{0, 0, 0};
    float scaled = 0;
    for (size_t i = 0; i < tri.size(); i += 3) {
        f3 v2 = v[tri[i]], v3 = v[tri[i + 1]], v4 = v[tri[i + 2]];
        float s = v2.x * (v3.z * v4.y - v3.y * v4.z) - v3.x * (v2.z * v4.y - v2.y * v4.z) +
                  v4.x * (v2.z * v3.y - v2.y * v3.z);
        scaled += s;
        f3 tc = v2 + v3;
        tc = tc + v4;
        tc = tc * s;
        summed = tc + summed;
        a += s * (v2.y * v2.y + v2.y * v3.y + v3.y * v3.y + v2.y * v4.y + v3.y * v4.y + v4.y * v4.y +
                  v2.z * v2.z + v2.z * v3.z + v3.z * v3.z + v2.z * v4.z + v3.z * v4.z + v4.z * v4.z);
        b += s * (v2.x * v2.x + v2.x * v3.x + v3.x * v3.x + v2.x * v4.x + v3.x * v4.x + v4.x * v4.x +
                  v2.z * v2.z + v2.z * v3.z + v3.z * v3.z + v2.z * v4.z + v3.z * v4.z + v4.z * v4.z);
        c += s * (v2.x * v2.x + v2.x * v3.x + v3.x * v3.x + v2.x * v4.x + v3.x * v4.x + v4.x * v4.x +
                  v2.y * v2.y + v2.y * v3.y + v3.y * v3.y + v2.y * v4.y + v3.y * v4.y + v4.y * v4.y);
        ao += s * (2 * v2.y * v2.z + v3.y * v2.z + v4.y * v2.z + v2.y * v3.z + 2 * v3.y * v3.z + v4.y * v3.z +
                   v2.y * v4.z + v3.y * v4.z + 2 * v4.y * v4.z);
        bo += s * (2 * v2.x * v2.z + v3.x * v2.z + v4.x * v2.z + v2.x * v3.z + 2 * v3.x * v3.z + v4.x * v3.z +
                   v2.x * v4.z + v3.x * v4.z + 2 * v4.x * v4.z);
        co += s * (2 * v2.x * v2.y + v3.x * v2.y + v4.x * v2.y + v2.x * v3.y + 2 * v3.x * v3.y + v4.x * v3.y +
                   v2.x * v4.y + v3.x * v4.y + 2 * v4.x * v4.y);
    }
    if (scaled < kEpsilon) {
        *vd = m3{0, 0, 0, 0, 0, 0, 0, 0, 0};
        *volume = 0;
        *center = f3{0, 0, 0};
        return;
    }
    *center = summed * (0.25f / scaled);
    *volume = scaled / 6;
    float density = 1 / *volume;
    float diag = density / 60;
    float off = -density / 120;
    a *= diag;
    b *= diag;
    c *= diag;
    ao *= off;
    bo *= off;
    co *= off;
    m3 r{a, co, bo, co, b, ao, bo, ao, c};
    f3 pt = *center - f3{0, 0, 0};
    float xx = 1.0f * pt.x * pt.x, yy = 1.0f * pt.y * pt.y, zz = 1.0f * pt.z * pt.z;
    float m12 = -1.0f * pt.x * pt.y, m13 = -1.0f * pt.x * pt.z, m23 = -1.0f * pt.y * pt.z;
    m3 add{yy + zz, m12, m13, m12, xx + zz, m23, m13, m23, xx + yy};
    *vd = m3{r.m11 - add.m11, r.m12 - add.m12, r.m13 - add.m13, r.m21 - add.m21, r.m22 - add.m22,
             r.m23 - add.m23, r.m31 - add.m31, r.m32 - add.m32, r.m33 - add.m33};
}

}  // namespace

// InertiaHelper.ComputeMinimumRadius (InertiaHelper.cs:529-570): the
// smallest distance from `center` to the plane of a hull triangle (the input
// points, the hull's triangle list), skipping degenerate triangles.  The
// reference throws for a negative distance (a centre outside the hull), which
// the hull of the points around its own centroid does not produce.
static float minimum_radius(const std::vector<f3>& p, const std::vector<int>& tri, f3 center) {
    float best = FLT_MAX;
    for (size_t i = 0; i + 2 < tri.size(); i += 3) {
        const f3 a = p[tri[i]], b = p[tri[i + 1]], c = p[tri[i + 2]];
        f3 n = cross3(c - a, b - a);
        const float ls = lensq3(n);
        if (!(ls > 1e-10f)) continue;
        n = divide3(n, sqrtf(ls));
        const float d = dot3(n, a - center);
        if (d < best) best = d;
    }
    return best;
}

bool compute_hull_description(const f3* points, uint32_t n, HullDescription* out) {
    std::vector<f3> p(points, points + n);
    for (const f3& q : p)
        if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z)) return false;
    std::vector<int> tri;
    if (!convex_hull(p, tri)) return false;
    // Unique surface vertices in the order of a backwards scan of the triangle
    // list (ConvexHullHelper.cs:88-112).
    out->vertices.clear();
    std::unordered_set<int> seen;
    for (int i = (int)tri.size() - 1; i >= 0; --i)
        if (seen.insert(tri[i]).second) out->vertices.push_back(p[tri[i]]);
    shape_distribution(p, tri, &out->center, &out->volume, &out->volume_distribution);
    if (!(out->volume > 0)) return false;
    for (f3& v : out->vertices) v = v - out->center;
    out->minimum_radius = minimum_radius(p, tri, out->center);
    return true;
}

}  // namespace bepu
