// k_narrowphase.hip -- per-pair contact generation and persistent manifold
// update on gfx950.
//
// One thread per broadphase pair.  The pair's previous state is found in last
// step's key -> slot table; the updated state is written to this step's slot
// and table.  Box-box candidate contacts (up to 8 before pruning) live in LDS,
// one column per thread, so the clipper never spills to scratch.
//
// Colliders (paths relative to the reference repository root):
//   box-box     BoxBoxCollider.AreBoxesColliding (BoxBoxCollider.cs:1494-2386)
//   box-sphere  BoxSphereTester.AreShapesColliding (BoxSphereTester.cs:23-85)
//   sphere      SphereTester.AreSpheresColliding (SphereTester.cs:23-56)
// Manifolds: BoxContactManifold.Update (BoxContactManifold.cs:125-206),
// SphereContactManifold / BoxSphereContactManifold.Update (…Manifold.cs:55-81).
#include "manifold.h"
#include "mesh.h"
#include "wave.h"

namespace bepu {

constexpr int kNpBlock = 128;
constexpr float kEps = 1e-7f;      // Toolbox.Epsilon
constexpr float kBigEps = 1e-5f;   // Toolbox.BigEpsilon

// Candidate contact list of one thread, stored in LDS columns:
// entry k of thread t is pd[k * stride + t] (position xyz, depth) and id[...],
// stride = the workgroup size.
struct LdsList {
    float4* pd;
    int* id;
    int n;
    int stride;
    __device__ float4& P(int k) { return pd[k * stride]; }
    __device__ int& I(int k) { return id[k * stride]; }
    // TinyStructList semantics: Add beyond 8 entries is ignored, RemoveAt shifts.
    __device__ void add(float4 p, int i) {
        if (n < 8) { P(n) = p; I(n) = i; ++n; }
    }
    __device__ void remove_at(int i) {
        for (int k = i; k < n - 1; ++k) { P(k) = P(k + 1); I(k) = I(k + 1); }
        --n;
    }
};

// ---------------------------------------------------------------------------
// Box-box: edge-edge contacts (BoxBoxCollider.cs:2392-2698)

__device__ __forceinline__ int edge_id(int a, int b) { return (a + 1) * 571 + (b + 1) * 577; }

// GetEdgeData (BoxBoxCollider.cs:2700-2864) as sign tables.
__device__ __forceinline__ void edge_data(int index, int axis, f3 h, f3* s, f3* e, int* sid, int* eid) {
    // For axis 0 the edge runs along x; index bits pick the signs of the other two axes.
    float u = (index & 2) ? 1.0f : -1.0f;  // first of the remaining axes
    float v = (index & 1) ? 1.0f : -1.0f;  // second
    if (axis == 0) {
        *s = f3{-h.x, u * h.y, v * h.z}; *e = f3{h.x, u * h.y, v * h.z};
        *sid = index; *eid = 4 + index;
    } else if (axis == 1) {
        *s = f3{u * h.x, -h.y, v * h.z}; *e = f3{u * h.x, h.y, v * h.z};
        *sid = ((index & 2) << 1) | (index & 1); *eid = *sid + 2;
    } else {
        *s = f3{u * h.x, v * h.y, -h.z}; *e = f3{u * h.x, v * h.y, h.z};
        *sid = (index << 1); *eid = *sid + 1;
    }
}

__device__ void pick_edges(f3 m, f3 h, f3* s1, f3* e1, int* s1id, int* e1id, f3* s2, f3* e2, int* s2id, int* e2id) {
    float d[4];
    int axis;
    if (fabsf(m.x) < kEps) {
        d[0] = -h.y * m.y - h.z * m.z; d[1] = -h.y * m.y + h.z * m.z;
        d[2] = h.y * m.y - h.z * m.z;  d[3] = h.y * m.y + h.z * m.z;
        axis = 0;
    } else if (fabsf(m.y) < kEps) {
        d[0] = -h.x * m.x - h.z * m.z; d[1] = -h.x * m.x + h.z * m.z;
        d[2] = h.x * m.x - h.z * m.z;  d[3] = h.x * m.x + h.z * m.z;
        axis = 1;
    } else {
        d[0] = -h.x * m.x - h.y * m.y; d[1] = -h.x * m.x + h.y * m.y;
        d[2] = h.x * m.x - h.y * m.y;  d[3] = h.x * m.x + h.y * m.y;
        axis = 2;
    }
    // FindHighestIndices (BoxBoxCollider.cs:2866-2890)
    int hi = 0;
    float hv = d[0];
    for (int i = 1; i < 4; i++)
        if (d[i] > hv) { hi = i; hv = d[i]; }
    int se = 0;
    float sv = -3.402823466e+38f;
    for (int i = 0; i < 4; i++)
        if (i != hi && d[i] > sv) { se = i; sv = d[i]; }
    edge_data(hi, axis, h, s1, e1, s1id, e1id);
    edge_data(se, axis, h, s2, e2, s2id, e2id);
}

// GetClosestPointsBetweenSegments (BoxBoxCollider.cs:2902-2990)
__device__ bool seg_closest(f3 p1, f3 q1, f3 p2, f3 q2, f3* c1, f3* c2) {
    f3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    float a = lensq3(d1), e = lensq3(d2), f = dot3(d2, r);
    float s, t;
    if (a <= kEps && e <= kEps) return false;
    if (a <= kEps) {
        s = 0.0f;
        t = f / e;
        if (t < 0 || t > 1) return false;
    } else {
        float c = dot3(d1, r);
        if (e <= kEps) {
            t = 0.0f;
            s = clampf_ref(-c / a, 0.0f, 1.0f);
        } else {
            float b = dot3(d1, d2);
            float denom = a * e - b * b;
            if (denom != 0.0f) {
                s = (b * f - c * e) / denom;
                if (s < 0 || s > 1) return false;
            } else {
                s = .5f;
            }
            t = (b * s + f) / e;
            if (t < 0 || t > 1) return false;
        }
    }
    *c1 = d1 * s + p1;
    *c2 = d2 * t + p2;
    return true;
}

__device__ void edge_edge(f3 ha, f3 hb, f3 pA, const m3& oA, f3 pB, const m3& oB, f3 mtd, LdsList& out) {
    f3 mA = mat_transform_t(-mtd, oA);
    f3 mB = mat_transform_t(mtd, oB);
    f3 as[2], ae[2], bs[2], be[2];
    int asid[2], aeid[2], bsid[2], beid[2];
    pick_edges(mA, ha, &as[0], &ae[0], &asid[0], &aeid[0], &as[1], &ae[1], &asid[1], &aeid[1]);
    pick_edges(mB, hb, &bs[0], &be[0], &bsid[0], &beid[0], &bs[1], &be[1], &bsid[1], &beid[1]);
    for (int k = 0; k < 2; ++k) {
        as[k] = mat_transform(as[k], oA) + pA;
        ae[k] = mat_transform(ae[k], oA) + pA;
        bs[k] = mat_transform(bs[k], oB) + pB;
        be[k] = mat_transform(be[k], oB) + pB;
    }
    out.n = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            f3 onA, onB;
            if (seg_closest(as[i], ae[i], bs[j], be[j], &onA, &onB)) {
                float d = dot3(onA - onB, mtd);
                if (d < 0) out.add(f4_of(onA, d), edge_id(asid[i], aeid[i]) * 2549 + edge_id(bsid[j], beid[j]) * 2857);
            }
        }
}

// ---------------------------------------------------------------------------
// Box-box: face contacts (BoxBoxCollider.cs:3459-3487, 4517-5278, 6007-6289)

struct Face {
    int id[4];
    f3 v[4];
    f3 n;
    float width, height;
};

__device__ __forceinline__ f3 xform(f3 c, const m3& o, f3 p) {  // Matrix.Transform (Matrix.cs:717-725)
    return f3{c.x * o.m11 + c.y * o.m21 + c.z * o.m31 + p.x, c.x * o.m12 + c.y * o.m22 + c.z * o.m32 + p.y,
              c.x * o.m13 + c.y * o.m23 + c.z * o.m33 + p.z};
}

// GetNearestFace (BoxBoxCollider.cs:6042-6177)
__device__ void nearest_face(f3 p, const m3& o, f3 mtd, f3 h, Face& f) {
    f.id[0] = f.id[1] = f.id[2] = f.id[3] = 0;
    f.v[0] = f.v[1] = f.v[2] = f.v[3] = f3{0, 0, 0};
    f.n = f3{0, 0, 0};
    f.width = f.height = 0;
    float xd = o.m11 * mtd.x + o.m12 * mtd.y + o.m13 * mtd.z;
    float yd = o.m21 * mtd.x + o.m22 * mtd.y + o.m23 * mtd.z;
    float zd = o.m31 * mtd.x + o.m32 * mtd.y + o.m33 * mtd.z;
    float ax = fabsf(xd), ay = fabsf(yd), az = fabsf(zd);
    float hw = h.x, hh = h.y, hl = h.z;
    int bit;
    if (ax > ay && ax > az) {
        if (xd < 0) { hw = -hw; bit = 0; } else bit = 1;
        f.v[0] = xform(f3{hw, hh, hl}, o, p); f.v[1] = xform(f3{hw, -hh, hl}, o, p);
        f.v[2] = xform(f3{hw, -hh, -hl}, o, p); f.v[3] = xform(f3{hw, hh, -hl}, o, p);
        f.n = xd < 0 ? f3{-o.m11, -o.m12, -o.m13} : f3{o.m11, o.m12, o.m13};
        f.width = hh * 2; f.height = hl * 2;
        f.id[0] = bit + 2 + 4; f.id[1] = bit + 4; f.id[2] = bit + 2; f.id[3] = bit;
    } else if (ay > ax && ay > az) {
        if (yd < 0) { hh = -hh; bit = 0; } else bit = 2;
        f.v[0] = xform(f3{hw, hh, hl}, o, p); f.v[1] = xform(f3{-hw, hh, hl}, o, p);
        f.v[2] = xform(f3{-hw, hh, -hl}, o, p); f.v[3] = xform(f3{hw, hh, -hl}, o, p);
        f.n = yd < 0 ? f3{-o.m21, -o.m22, -o.m23} : f3{o.m21, o.m22, o.m23};
        f.width = hw * 2; f.height = hl * 2;
        f.id[0] = 1 + bit + 4; f.id[1] = bit + 4; f.id[2] = 1 + bit; f.id[3] = bit;
    } else if (az > ax && az > ay) {
        if (zd < 0) { hl = -hl; bit = 0; } else bit = 4;
        f.v[0] = xform(f3{hw, hh, hl}, o, p); f.v[1] = xform(f3{-hw, hh, hl}, o, p);
        f.v[2] = xform(f3{-hw, -hh, hl}, o, p); f.v[3] = xform(f3{hw, -hh, hl}, o, p);
        f.n = zd < 0 ? f3{-o.m31, -o.m32, -o.m33} : f3{o.m31, o.m32, o.m33};
        f.width = hw * 2; f.height = hh * 2;
        f.id[0] = 1 + 2 + bit; f.id[1] = 2 + bit; f.id[2] = 1 + bit; f.id[3] = bit;
    }
}

struct ClipEdge {
    f3 a, b, perp;
    int id;
};

// BoxFace.GetEdge (BoxBoxCollider.cs:6220-6268)
__device__ ClipEdge clip_edge(const Face& f, int i) {
    ClipEdge e;
    int i1 = (i + 1) & 3, i2 = (i + 2) & 3;
    e.a = f.v[i];
    e.b = f.v[i1];
    e.id = edge_id(f.id[i], f.id[i1]);
    f3 dir = normalize3(e.b - e.a);
    e.perp = cross3(dir, f.n);
    if (dot3(e.perp, f.v[i2] - e.a) > 0) e.perp = -e.perp;
    return e;
}

// ComputeIntersection (BoxBoxCollider.cs:6007-6040)
__device__ bool intersect_edge(f3 a1, f3 a2, const ClipEdge& ce, f3* out) {
    f3 off = ce.a - a1;
    f3 dir = a2 - a1;
    float t = dot3(off, ce.perp) / dot3(dir, ce.perp);
    if (t < 0 || t > 1) return false;
    *out = dir * t + a1;
    off = *out - ce.a;
    dir = ce.b - ce.a;
    t = dot3(dir, off);
    if (t < 0 || t > lensq3(dir)) return false;
    return true;
}

// Keeps entries [from, n) of `l` whose depth against the clip face is <= 0,
// recomputing their depth (the three "Compute depths" passes of ClipFacesDirect).
// Compacts in place: an entry moves only to a lower index, so the kept entries
// keep their order, as a rebuilt list would.
__device__ void refilter_depths(LdsList& l, int from, f3 mtd, float clipFaceDot) {
    int j = from;
    for (int i = from; i < l.n; i++) {
        float4 it = l.P(i);
        int id = l.I(i);
        float depth = dot3(f3_of(it), mtd) - clipFaceDot;
        if (depth <= 0) {
            l.P(j) = make_float4(it.x, it.y, it.z, depth);
            l.I(j) = id;
            ++j;
        }
    }
    l.n = j;
}

// ClipFacesDirect, non-ALLOWUNSAFE (BoxBoxCollider.cs:4517-5278)
__device__ void clip_faces(const Face& cf, const Face& face, f3 mtd, LdsList& out) {
    out.n = 0;
    float icw = 1 / cf.width, ich = 1 / cf.height;
    float icw2 = icw * icw, ich2 = ich * ich;
    f3 clipX = (cf.v[3] - cf.v[2]) * icw2;
    f3 clipY = (cf.v[1] - cf.v[2]) * ich2;
    float ifw = 1 / face.width, ifh = 1 / face.height;
    float ifw2 = ifw * ifw, ifh2 = ifh * ifh;
    f3 faceX = (face.v[3] - face.v[2]) * ifw2;
    f3 faceY = (face.v[1] - face.v[2]) * ifh2;
    f3 cc = cf.v[0] + cf.v[2];
    float ccx = dot3(cc, clipX) * .5f, ccy = dot3(cc, clipY) * .5f;
    f3 fc = face.v[0] + face.v[2];
    float fcx = dot3(fc, faceX) * .5f, fcy = dot3(fc, faceY) * .5f;
    float ex = .5f + .01f * icw, ey = .5f + .01f * ich;
    float cMaxX = ccx + ex, cMaxY = ccy + ey, cMinX = ccx - ex, cMinY = ccy - ey;
    ex = .5f + .01f * ifw;
    ey = .5f + .01f * ifh;
    float fMaxX = fcx + ex, fMaxY = fcy + ey, fMinX = fcx - ex, fMinY = fcy - ey;
    // Bit k of each mask: face vertex k (resp. clip vertex k) is inside that bound.
    unsigned maxX = 0, minX = 0, maxY = 0, minY = 0, cin = 0;
    for (int i = 0; i < 4; ++i) {
        float dx = dot3(clipX, face.v[i]);
        float dy = dot3(clipY, face.v[i]);
        maxX |= (dx < cMaxX) << i; minX |= (dx > cMinX) << i;
        maxY |= (dy < cMaxY) << i; minY |= (dy > cMinY) << i;
        float ex2 = dot3(faceX, cf.v[i]);
        float ey2 = dot3(faceY, cf.v[i]);
        cin |= ((ex2 < fMaxX) && (ex2 > fMinX) && (ey2 < fMaxY) && (ey2 > fMinY)) << i;
    }
    unsigned fin = maxX & minX & maxY & minY;
    for (int i = 0; i < 4; ++i)
        if (fin & (1u << i)) out.add(f4_of(face.v[i], 0.0f), face.id[i]);
    float clipFaceDot = dot3(cf.v[0], mtd);
    refilter_depths(out, 0, mtd, clipFaceDot);
    int prev = out.n;
    if (prev >= 4) return;
    float b = dot3(face.v[0], face.n);
    for (int i = 0; i < 4; ++i)
        if (cin & (1u << i)) {
            float a = dot3(cf.v[i], face.n);
            out.add(f4_of(cf.v[i] - face.n * (a - b), 0.0f), cf.id[i] + 8);
        }
    refilter_depths(out, prev, mtd, clipFaceDot);
    prev = out.n;
    if (prev >= 4) return;
    // Clip edge k is bounded by: 0 -> maxY, 1 -> minX, 2 -> minY, 3 -> maxX.
    const unsigned F[4] = {maxY, minX, minY, maxX};
    for (int k = 0; k < 4; ++k) {
        ClipEdge ce = clip_edge(cf, k);
        unsigned fl = F[k];
        for (int i = 0; i < 4; ++i) {
            if (fl & (1u << i)) continue;
            // neighbours of vertex i in the reference's order (vertex 4 on clip edge 2 is reversed)
            int n1 = (i == 0) ? 1 : (i == 1) ? 0 : (i == 2) ? 1 : (k == 2 ? 2 : 0);
            int n2 = (i == 0) ? 3 : (i == 1) ? 2 : (i == 2) ? 3 : (k == 2 ? 0 : 2);
            for (int q = 0; q < 2; ++q) {
                int j = q == 0 ? n1 : n2;
                if (!(fl & (1u << j)) || out.n >= 8) continue;
                int lo = min(i, j), hi = max(i, j);
                int s = (lo == 0 && hi == 3) ? 3 : lo;
                int e = (lo == 0 && hi == 3) ? 0 : hi;
                f3 v;
                if (intersect_edge(face.v[s], face.v[e], ce, &v))
                    out.add(f4_of(v, 0.0f), edge_id(face.id[s], face.id[e]) * 2549 + ce.id * 2857);
            }
        }
    }
    refilter_depths(out, prev, mtd, clipFaceDot);
}

// PruneContactsMaxDistance, non-ALLOWUNSAFE (BoxBoxCollider.cs:3612-3741), in
// place: the four picks are read before the list is rewritten with them.
__device__ void prune4(f3 mtd, LdsList& in) {
    int n = in.n;
    int deepest = 0;
    for (int i = 1; i < n; i++)
        if (in.P(i).w > in.P(deepest).w) deepest = i;
    f3 dp = f3_of(in.P(deepest));
    int furthest = 0;
    float fd = distsq3(dp, f3_of(in.P(0)));
    for (int i = 1; i < n; i++) {
        float d = distsq3(dp, f3_of(in.P(i)));
        if (d > fd) { fd = d; furthest = i; }
    }
    f3 yAxis = cross3(mtd, f3_of(in.P(furthest)) - dp);
    int mn = 0, mx = 0;
    float minY = dot3(f3_of(in.P(0)), yAxis), maxY = minY;
    for (int i = 1; i < n; i++) {
        float d = dot3(yAxis, f3_of(in.P(i)));
        if (d < minY) { minY = d; mn = i; }
        else if (d > maxY) { maxY = d; mx = i; }
    }
    float4 p0 = in.P(deepest), p1 = in.P(furthest), p2 = in.P(mn), p3 = in.P(mx);
    int i0 = in.I(deepest), i1 = in.I(furthest), i2 = in.I(mn), i3 = in.I(mx);
    in.P(0) = p0; in.I(0) = i0;
    in.P(1) = p1; in.I(1) = i1;
    in.P(2) = p2; in.I(2) = i2;
    in.P(3) = p3; in.I(3) = i3;
    in.n = 4;
}

// Separating-axis tables for the nine edge axes A.i x B.j (BoxBoxCollider.cs:1795-2352):
// rarb = sum of h[h_k] * |R[p_k]| (left to right), tl = t[tu]*R[q1] - t[tv]*R[q2].
// Separation is only tested when |R[j][i]| < 1 (the reference reads the transposed entry).
struct EdgeAxis {
    int8_t h[4];
    int8_t p[4];
    int8_t tu, q1, tv, q2;
};
__constant__ EdgeAxis c_edge_axes[9] = {
    {{1, 2, 4, 5}, {6, 3, 2, 1}, 2, 3, 1, 6},  // A.X x B.X
    {{1, 2, 3, 5}, {7, 4, 2, 0}, 2, 4, 1, 7},  // A.X x B.Y
    {{1, 2, 3, 4}, {8, 5, 1, 0}, 2, 5, 1, 8},  // A.X x B.Z
    {{0, 2, 4, 5}, {6, 0, 5, 4}, 0, 6, 2, 0},  // A.Y x B.X
    {{0, 2, 3, 5}, {7, 1, 5, 3}, 0, 7, 2, 1},  // A.Y x B.Y
    {{0, 2, 3, 4}, {8, 2, 4, 3}, 0, 8, 2, 2},  // A.Y x B.Z
    {{0, 1, 4, 5}, {3, 0, 8, 7}, 1, 0, 0, 3},  // A.Z x B.X
    {{0, 1, 3, 5}, {4, 1, 8, 6}, 1, 1, 0, 4},  // A.Z x B.Y
    {{0, 1, 3, 4}, {5, 2, 7, 6}, 1, 2, 0, 5},  // A.Z x B.Z
};

__device__ __forceinline__ float m3_at(const m3& m, int k) {
    switch (k) {
    case 0: return m.m11; case 1: return m.m12; case 2: return m.m13;
    case 3: return m.m21; case 4: return m.m22; case 5: return m.m23;
    case 6: return m.m31; case 7: return m.m32; default: return m.m33;
    }
}

// BoxBoxCollider.AreBoxesColliding (BoxBoxCollider.cs:1494-2386).  Fills `out`
// with up to 4 contacts; returns whether the boxes collide.
__device__ bool box_box(f3 ha, f3 hb, f3 pA, q4 qA, f3 pB, q4 qB, f3* axis_out, LdsList& out) {
    m3 aO = mat_from_quat(qA);
    m3 bO = mat_from_quat(qB);
    f3 t = pB - pA;
    float minD = -3.402823466e+38f;
    f3 minAxis{0, 0, 0};
    int feature = 2;
    m3 R, aR;
    R.m11 = aO.m11 * bO.m11 + aO.m12 * bO.m12 + aO.m13 * bO.m13;
    R.m12 = aO.m11 * bO.m21 + aO.m12 * bO.m22 + aO.m13 * bO.m23;
    R.m13 = aO.m11 * bO.m31 + aO.m12 * bO.m32 + aO.m13 * bO.m33;
    aR.m11 = fabsf(R.m11) + kEps; aR.m12 = fabsf(R.m12) + kEps; aR.m13 = fabsf(R.m13) + kEps;
    float tX = t.x;
    t.x = t.x * aO.m11 + t.y * aO.m12 + t.z * aO.m13;
    // Face test: returns false on separation, else folds the candidate in.
    auto face = [&](float tv, float rarb, f3 r, int feat) -> bool {
        if (tv > rarb || tv < -rarb) return false;
        float td;
        f3 ax;
        if (tv > 0) { td = tv - rarb; ax = -r; }
        else { td = -tv - rarb; ax = r; }
        if (td > minD) { minD = td; minAxis = ax; feature = feat; }
        return true;
    };
    if (!face(t.x, ha.x + hb.x * aR.m11 + hb.y * aR.m12 + hb.z * aR.m13, mat_row(aO, 0), 0)) return false;
    R.m21 = aO.m21 * bO.m11 + aO.m22 * bO.m12 + aO.m23 * bO.m13;
    R.m22 = aO.m21 * bO.m21 + aO.m22 * bO.m22 + aO.m23 * bO.m23;
    R.m23 = aO.m21 * bO.m31 + aO.m22 * bO.m32 + aO.m23 * bO.m33;
    aR.m21 = fabsf(R.m21) + kEps; aR.m22 = fabsf(R.m22) + kEps; aR.m23 = fabsf(R.m23) + kEps;
    float tY = t.y;
    t.y = tX * aO.m21 + t.y * aO.m22 + t.z * aO.m23;
    if (!face(t.y, ha.y + hb.x * aR.m21 + hb.y * aR.m22 + hb.z * aR.m23, mat_row(aO, 1), 0)) return false;
    R.m31 = aO.m31 * bO.m11 + aO.m32 * bO.m12 + aO.m33 * bO.m13;
    R.m32 = aO.m31 * bO.m21 + aO.m32 * bO.m22 + aO.m33 * bO.m23;
    R.m33 = aO.m31 * bO.m31 + aO.m32 * bO.m32 + aO.m33 * bO.m33;
    aR.m31 = fabsf(R.m31) + kEps; aR.m32 = fabsf(R.m32) + kEps; aR.m33 = fabsf(R.m33) + kEps;
    t.z = tX * aO.m31 + tY * aO.m32 + t.z * aO.m33;
    if (!face(t.z, ha.z + hb.x * aR.m31 + hb.y * aR.m32 + hb.z * aR.m33, mat_row(aO, 2), 0)) return false;
    const float antiB = .01f;
    minD += antiB;
    if (!face(t.x * R.m11 + t.y * R.m21 + t.z * R.m31, hb.x + ha.x * aR.m11 + ha.y * aR.m21 + ha.z * aR.m31,
              mat_row(bO, 0), 1)) return false;
    if (!face(t.x * R.m12 + t.y * R.m22 + t.z * R.m32, hb.y + ha.x * aR.m12 + ha.y * aR.m22 + ha.z * aR.m32,
              mat_row(bO, 1), 1)) return false;
    if (!face(t.x * R.m13 + t.y * R.m23 + t.z * R.m33, hb.z + ha.x * aR.m13 + ha.y * aR.m23 + ha.z * aR.m33,
              mat_row(bO, 2), 1)) return false;
    if (feature != 1) minD -= antiB;
    const float antiEdge = .01f;
    minD += antiEdge;
    const float hs[6] = {ha.x, ha.y, ha.z, hb.x, hb.y, hb.z};
    const float tt[3] = {t.x, t.y, t.z};
    for (int e = 0; e < 9; ++e) {
        const EdgeAxis& s = c_edge_axes[e];
        int i = e / 3, j = e % 3;
        float rarb = hs[s.h[0]] * m3_at(aR, s.p[0]) + hs[s.h[1]] * m3_at(aR, s.p[1]) +
                     hs[s.h[2]] * m3_at(aR, s.p[2]) + hs[s.h[3]] * m3_at(aR, s.p[3]);
        float tl = tt[s.tu] * m3_at(R, s.q1) - tt[s.tv] * m3_at(R, s.q2);
        if (m3_at(aR, j * 3 + i) < 1 && (tl > rarb || tl < -rarb)) return false;
        f3 a = mat_row(aO, i), b = mat_row(bO, j);
        f3 ax;
        float td;
        if (tl > 0) {
            ax = f3{b.y * a.z - b.z * a.y, b.z * a.x - b.x * a.z, b.x * a.y - b.y * a.x};
            float inv = 1 / len3(ax);
            td = (tl - rarb) * inv;
            if (td > minD) { feature = 2; minD = td; minAxis = f3{ax.x * inv, ax.y * inv, ax.z * inv}; }
        } else {
            ax = f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
            float inv = 1 / len3(ax);
            td = (-tl - rarb) * inv;
            if (td > minD) { feature = 2; minD = td; minAxis = f3{ax.x * inv, ax.y * inv, ax.z * inv}; }
        }
    }
    if (feature == 2) {
        edge_edge(ha, hb, pA, aO, pB, bO, minAxis, out);
    } else {
        Face fa, fb;
        nearest_face(pA, aO, -minAxis, ha, fa);
        nearest_face(pB, bO, minAxis, hb, fb);
        if (feature == 0) clip_faces(fa, fb, -minAxis, out);
        else clip_faces(fb, fa, minAxis, out);
        if (out.n > 4) prune4(minAxis, out);
    }
    *axis_out = minAxis;
    return true;
}

// SphereTester.AreSpheresColliding (SphereTester.cs:23-56)
__device__ bool sphere_sphere(float ra, float rb, f3 pa, f3 pb, float maxd, f3* pos, f3* nrm, float* depth) {
    float rs = ra + rb;
    f3 d = pb - pa;
    float cd = lensq3(d);
    if (!(cd < (rs + maxd) * (rs + maxd))) return false;
    *pos = (rs > kEps ? d * (ra / rs) : f3{0, 0, 0}) + pa;
    cd = sqrtf(cd);
    *nrm = cd > kBigEps ? divide3(d, cd) : f3{0, 1, 0};
    *depth = rs - cd;
    return true;
}

// BoxSphereTester.AreShapesColliding (BoxSphereTester.cs:23-85)
__device__ bool box_sphere(f3 h, f3 bp, q4 bq, float r, f3 sp, float maxd, f3* pos, f3* nrm, float* depth) {
    f3 lp = qrotate(sp - bp, qconj(bq));
    f3 lc{clampf_ref(lp.x, -h.x, h.x), clampf_ref(lp.y, -h.y, h.y), clampf_ref(lp.z, -h.z, h.z)};
    *pos = qrotate(lc, bq) + bp;
    f3 off = sp - *pos;
    float ol = lensq3(off);
    if (ol > (r + maxd) * (r + maxd)) return false;
    if (ol > kEps) {
        ol = sqrtf(ol);
        *nrm = divide3(off, ol);
        *depth = r - ol;
    } else {
        f3 pd{lc.x < 0 ? lc.x + h.x : h.x - lc.x, lc.y < 0 ? lc.y + h.y : h.y - lc.y,
              lc.z < 0 ? lc.z + h.z : h.z - lc.z};
        f3 n;
        if (pd.x < pd.y && pd.x < pd.z) { n = lc.x > 0 ? f3{1, 0, 0} : f3{-1, 0, 0}; *depth = pd.x; }
        else if (pd.y < pd.z) { n = lc.y > 0 ? f3{0, 1, 0} : f3{0, -1, 0}; *depth = pd.y; }
        else { n = lc.z > 0 ? f3{0, 0, 1} : f3{0, 0, -1}; *depth = pd.z; }
        *depth += r;
        *nrm = qrotate(n, bq);
    }
    return true;
}

__device__ __forceinline__ void update_box_pair(const DeviceWorld& w, PairState& s, LdsList& m) {
    float4 PA = w.pos[s.a], PB = w.pos[s.b], QA = w.quat[s.a], QB = w.quat[s.b];
    float4 SA = w.shape[s.a], SB = w.shape[s.b];
    f3 axis;
    m.n = 0;
    if (box_box(f3_of(SA), f3_of(SB), f3_of(PA), q4{QA.x, QA.y, QA.z, QA.w}, f3_of(PB), q4{QB.x, QB.y, QB.z, QB.w},
                &axis, m)) {
        axis = -axis;
        // Contacts whose id is no longer found, removed afterwards in
        // descending list order (indices fixed, so the state stays in registers).
        uint32_t rem = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (i >= s.ccount) break;
            const int slot = s.clist[i];
            const int id = __float_as_int(get4(s.cnrm, slot).w);
            bool found = false;
            for (int j = m.n - 1; j >= 0; j--) {
                if (m.I(j) == id) {
                    found = true;
                    float4 c = m.P(j);
                    set4(s.cpos, slot, make_float4(c.x, c.y, c.z, -c.w));
                    set4(s.cnrm, slot, f4_of(axis, __int_as_float(id)));
                    m.remove_at(j);
                    break;
                }
            }
            if (!found) rem |= 1u << i;
        }
#pragma unroll
        for (int i = 3; i >= 0; i--)
            if (rem & (1u << i)) remove_contact(s, i);
        for (int i = 0; i < m.n; i++) {
            float4 c = m.P(i);
            add_contact(s, f3_of(c), axis, -c.w, m.I(i));
        }
    } else {
#pragma unroll
        for (int i = 3; i >= 0; i--)
            if (i < s.ccount) remove_contact(s, i);
    }
}

__device__ __forceinline__ void update_single_pair(const DeviceWorld& w, PairState& s) {
    float4 PA = w.pos[s.a], PB = w.pos[s.b];
    float4 SA = w.shape[s.a], SB = w.shape[s.b];
    f3 pos, nrm;
    float depth;
    bool hit;
    if (s.kind == BEPU_PAIR_SPHERE_SPHERE) {
        hit = sphere_sphere(SA.x, SB.x, f3_of(PA), f3_of(PB), w.s.max_contact_distance, &pos, &nrm, &depth);
    } else {
        float4 QA = w.quat[s.a];
        hit = box_sphere(f3_of(SA), f3_of(PA), q4{QA.x, QA.y, QA.z, QA.w}, SB.x, f3_of(PB), w.s.max_contact_distance,
                         &pos, &nrm, &depth);
    }
    bool prev = (s.flags & PF_PREV_COLLIDING) != 0;
    bool colliding = false;
    if (hit) {
        if (!prev && depth >= 0) {
            add_contact(s, pos, nrm, depth, 0);
            colliding = true;
        } else if (prev) {
            const int slot = s.clist[0];
            set4(s.cpos, slot, f4_of(pos, depth));
            set4(s.cnrm, slot, f4_of(nrm, get4(s.cnrm, slot).w));
            colliding = true;
        }
    } else if (prev) {
        remove_contact(s, 0);
    }
    s.flags = colliding ? (s.flags | PF_PREV_COLLIDING) : (s.flags & ~PF_PREV_COLLIDING);
}

__device__ __forceinline__ void init_pair(const DeviceWorld& w, uint64_t key, PairState& s) {
    uint32_t i = (uint32_t)(key >> 32), j = (uint32_t)(key & 0xffffffffu);
    int si = (int)w.shape[i].w, sj = (int)w.shape[j].w;
    s.key = key;
    s.ccount = 0;
    s.npen = 0;
    s.flags = 0;
    for (int k = 0; k < 4; ++k) {
        s.clist[k] = 0; s.pslot[k] = 0; s.penacc[k] = 0; s.fracc[k] = 0;
        s.cpos[k] = make_float4(0, 0, 0, 0); s.cnrm[k] = make_float4(0, 0, 0, 0); s.frlin[k] = make_float4(0, 0, 0, 0);
    }
    s.slacc[0] = s.slacc[1] = 0;
    s.twacc = 0;
    s.color = -1;
    for (int k = 0; k < 6; ++k) s.sllin[k] = 0;
    s.tester = 0;
    s.toi = 1.0f;
    s.a = i;
    s.b = j;
    const bool mi = si == BEPU_SHAPE_STATIC_MESH || si == BEPU_SHAPE_TERRAIN;
    if (mi || sj == BEPU_SHAPE_STATIC_MESH || sj == BEPU_SHAPE_TERRAIN) {
        // StaticMeshPairHandler.Initialize (StaticMeshPairHandler.cs:57-79),
        // TerrainPairHandler.Initialize (TerrainPairHandler.cs:70-93): A is the convex.
        if (mi) {
            s.a = j;
            s.b = i;
        }
        s.kind = (int)w.shape[s.a].w == BEPU_SHAPE_SPHERE ? BEPU_PAIR_MESH_SPHERE : BEPU_PAIR_MESH_CONVEX;
    } else if (si == BEPU_SHAPE_BOX && sj == BEPU_SHAPE_BOX) s.kind = BEPU_PAIR_BOX_BOX;
    else if (si == BEPU_SHAPE_SPHERE && sj == BEPU_SHAPE_SPHERE) s.kind = BEPU_PAIR_SPHERE_SPHERE;
    else if (si == BEPU_SHAPE_BOX && sj == BEPU_SHAPE_SPHERE) s.kind = BEPU_PAIR_BOX_SPHERE;
    else if (si == BEPU_SHAPE_SPHERE && sj == BEPU_SHAPE_BOX) { s.kind = BEPU_PAIR_BOX_SPHERE; s.a = j; s.b = i; }
    else s.kind = BEPU_PAIR_CONVEX;
    // MaterialManager.DefaultMaterialBlender (MaterialManager.cs:77-85)
    float4 ma = w.mat[s.a], mb = w.mat[s.b];
    s.bounce = ma.z * mb.z;
    s.kinetic = ma.y * mb.y;
    s.static_f = ma.x * mb.x;
}


// Counter of general convex work list q = 2 * tester state + (two hulls ? 0 : 1).
__device__ __forceinline__ uint32_t convex_list_counter(uint32_t q) {
    const uint32_t state = q >> 1;
    if (q & 1) return CTR_CONVEX_OTHER0 + state;
    return state == 0 ? CTR_CONVEX : (state == 1 ? CTR_CONVEX_SHALLOW : CTR_CONVEX_DEEP);
}

// Last step's slot of a general convex pair (found in last step's key -> slot
// table; -1 for a new pair) and the pair tester's state it ended that step in
// (0 for a new pair).  A new pair's state is set up in slot t here; k_pair_prep
// copies a carried-over one.  The pair enters this step's table in
// k_narrowphase.
// Last step's slot of the pair with this key, or -1: none, a body inserted
// for this step (possibly into ids a removed world used) has no history, and
// a slot that was no pair (stale or void) is no history either.
__device__ __forceinline__ int prior_pair(const DeviceWorld& w, uint64_t key) {
    uint32_t ka = (uint32_t)(key >> 32), kb = (uint32_t)key;
    const uint32_t step = step_index(w);
    bool fresh = w.born[ka] == step || w.born[kb] == step;
    int old = fresh ? -1 : hash_find(w.hash_keys_old, w.hash_vals_old, w.hash_cap, key);
    if (old >= 0 && (w.hot_old[old].flags & (PF_DEAD | PF_VOID))) old = -1;
    return old;
}

// An overlap of two inactive bodies with no pair to keep: a slot with no state.
__device__ __forceinline__ void void_pair(const DeviceWorld& w, uint64_t key, uint32_t t) {
    PairState s;
    init_pair(w, key, s);
    s.flags |= PF_VOID;
    w.pairs_cur[t] = s;
    w.hot_cur[t] = hot_of(s);
}

__device__ __forceinline__ int begin_pair(const DeviceWorld& w, uint64_t key, uint32_t t, int& tester) {
    int old = prior_pair(w, key);
    tester = 0;
    if (old >= 0) tester = w.pairs_old[old].tester;
    else init_pair(w, key, w.pairs_cur[t]);
    return old;
}

// Every pair of this step goes to a work list.  Pairs with a hull or a
// capsule take the general convex path (NarrowPhaseHelper.cs:360-367); boxes
// and spheres have closed-form testers.  A general convex pair gets its pair
// state and its tester state carried into this step's slot and (slot, last
// slot) appended to
// the list of its pair tester's state (new pairs start Separated), so that a
// wavefront of k_convex mostly runs one branch of the state machine; pairs of
// two hulls fill a state's region from its front, the others from its back.
// A closed-form pair is listed by kind (box-box from the front of np_work, the
// others from its back), so that a wavefront of k_narrowphase runs one
// contact routine.  A pair with a static mesh goes to k_mesh's list with its
// tester table carried over (frozen mesh pairs keep theirs too).  Appends are gathered per workgroup: one global atomic per
// workgroup and list (the list counters are shared by the whole device, so
// per-wavefront atomics would queue on them).
constexpr int kPrepBlock = 512;
__global__ void __launch_bounds__(kPrepBlock) k_pair_prep(DeviceWorld w) {
    __shared__ uint32_t s_cnt[9], s_base[9];
    if (threadIdx.x < 9) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t npairs = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    // List: 2 * state + (two hulls ? 0 : 1) for general convex pairs, 6 for
    // box-box, 7 for the other closed-form pairs, 8 for mesh pairs.
    int my_q = -1;
    uint32_t old_slot = ~0u;
    bool frozen = false, mesh = false;
    uint64_t key = 0;
    if (t < npairs) {
        key = w.pair_keys[t];
        const uint32_t fi = w.bflags[(uint32_t)(key >> 32)], fj = w.bflags[(uint32_t)key];
        mesh = ((fi | fj) & BF_MESH) != 0;
        if (body_inactive(fi) && body_inactive(fj)) {
            // The broad phase reports no overlap of two inactive entries
            // (BroadPhase.cs:97-102): last step's pair is kept as it was (not
            // updated, NarrowPhase.cs:300-333), and none is created.
            frozen = true;
            const int old = prior_pair(w, key);
            if (old >= 0) old_slot = (uint32_t)old;
            else void_pair(w, key, t);
        } else if (mesh) {
            int unused;
            const int old = begin_pair(w, key, t, unused);
            if (old >= 0) old_slot = (uint32_t)old;
            my_q = 8;
        } else if ((fi | fj) & BF_GENERAL) {
            int cstate;
            int old = begin_pair(w, key, t, cstate);
            if (cstate != 1 && cstate != 2) cstate = 0;
            if (old >= 0) old_slot = (uint32_t)old;
            my_q = 2 * cstate + ((fi & fj & BF_HULL) ? 0 : 1);
        } else {
            my_q = ((fi | fj) & BF_SPHERE) ? 7 : 6;
        }
    }
    const int lane = __lane_id();
    {
        // Carried-over states: the wavefront copies them together, three pair
        // states (21 float4 each) per pass with consecutive lanes on
        // consecutive 16-byte chunks, instead of each lane reading its own
        // pair's 336 bytes 16 bytes at a time.
        static_assert(sizeof(PairState) == 21 * 16 && alignof(PairState) == 16, "PairState is copied as 21 float4s");
        constexpr int kFlagsChunk = 1;  // PairState::flags is the w of the second float4
        static_assert(offsetof(PairState, flags) == 16 * kFlagsChunk + 12, "PairState::flags position");
        const int part = lane / 21, chunk = lane % 21;  // lane 63 (part 3) idles
        uint64_t todo = __ballot(old_slot != ~0u);
        while (todo) {
            const int l0 = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const int l1 = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
            if (todo) todo &= todo - 1;
            const int l2 = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
            if (todo) todo &= todo - 1;
            const int src_lane = part == 0 ? l0 : (part == 1 ? l1 : (part == 2 ? l2 : -1));
            const uint32_t src = __shfl(old_slot, src_lane < 0 ? 0 : src_lane);
            const uint32_t dst = __shfl(t, src_lane < 0 ? 0 : src_lane);
            const bool src_frozen = __shfl((int)frozen, src_lane < 0 ? 0 : src_lane) != 0;
            if (src_lane >= 0) {
                float4 v = reinterpret_cast<const float4*>(&w.pairs_old[src])[chunk];
                if (chunk == kFlagsChunk) {
                    // this step's activity bits of the pair
                    uint32_t fl = __float_as_uint(v.w) & ~(uint32_t)(PF_FROZEN | PF_DEAD | PF_VOID);
                    v.w = __uint_as_float(fl | (src_frozen ? (uint32_t)PF_FROZEN : 0u));
                }
                reinterpret_cast<float4*>(&w.pairs_cur[dst])[chunk] = v;
            }
        }
        // Their tester and supplement states the same way, six (10 float4
        // each) per pass.
        static_assert(sizeof(ConvexState) == 10 * 16 && alignof(ConvexState) == 16, "ConvexState is copied as 10 float4s");
        const int cpart = lane / 10, cchunk = lane % 10;  // lanes 60..63 idle
        todo = __ballot(old_slot != ~0u);
        while (todo) {
            int src_lane = -1;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int lk = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
                if (todo) todo &= todo - 1;
                if (cpart == k) src_lane = lk;
            }
            const uint32_t src = __shfl(old_slot, src_lane < 0 ? 0 : src_lane);
            const uint32_t dst = __shfl(t, src_lane < 0 ? 0 : src_lane);
            if (src_lane >= 0)
                reinterpret_cast<float4*>(&w.conv_cur[dst])[cchunk] =
                    reinterpret_cast<const float4*>(&w.conv_old[src])[cchunk];
        }
    }
    if (mesh && old_slot != ~0u) {
        // A mesh pair's tester table (its live entries; the count travels in
        // the pair state).
        const int nt = w.pairs_old[old_slot].tester;
        const MeshTester* src = w.mtest_old + (size_t)old_slot * kMeshTesters;
        MeshTester* dst = w.mtest_cur + (size_t)t * kMeshTesters;
        for (int k = 0; k < nt; ++k) dst[k] = src[k];
    }
    if (frozen && old_slot != ~0u) {
        // A kept pair enters this step's table and, with contacts, its island.
        if (!hash_insert(w.hash_keys_cur, w.hash_vals_cur, w.hash_cap, key, t)) w.counters[ctr(CTR_OVERFLOW)] = 3;
        PairHot o = w.hot_old[old_slot];
        if (o.ccount > 0 && (w.bflags[o.a] & 1u) && (w.bflags[o.b] & 1u)) uf_union(w.uf_parent, o.a, o.b);
        o.flags = (o.flags & ~(uint32_t)(PF_DEAD | PF_VOID)) | PF_FROZEN;
        w.hot_cur[t] = o;
    }
    uint32_t my_off = 0;
    for (int q = 0; q < 9; ++q) {
        const uint64_t mine = __ballot(my_q == q);
        if (!mine) continue;
        const int leader = __ffsll((unsigned long long)mine) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&s_cnt[q], (uint32_t)__popcll(mine));
        base = __shfl(base, leader);
        if (my_q == q) my_off = base + (uint32_t)__popcll(mine & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (threadIdx.x < 9 && s_cnt[threadIdx.x]) {
        const uint32_t q = threadIdx.x;
        uint32_t c;
        if (q == 8) c = CTR_MESH;
        else if (q == 6) c = CTR_NP_BOX;
        else if (q == 7) c = CTR_NP_SINGLE;
        else c = convex_list_counter(q);
        s_base[q] = atomicAdd(&w.counters[ctr(c)], s_cnt[q]);
    }
    __syncthreads();
    if (my_q == 8) {
        w.mesh_work[s_base[8] + my_off] = make_uint2(t, old_slot);
    } else if (my_q >= 0) {
        const uint32_t k = s_base[my_q] + my_off;
        const uint32_t at = (my_q & 1) == 0 ? k : w.pair_cap - 1 - k;
        if (my_q < 6) w.conv_work[(size_t)(my_q >> 1) * w.pair_cap + at] = make_uint2(t, old_slot);
        else w.np_work[at] = t;
    }
}

// One closed-form pair in slot t: last step's state (or a new one) updated by
// its contact routine, stored to this step's slot and entered in the table.
template <int BLOCK>
__device__ __forceinline__ void narrowphase_pair(DeviceWorld& w, uint32_t t, float4* lds_pd, int* lds_id) {
    uint64_t key = w.pair_keys[t];
    int old = prior_pair(w, key);
    const bool frozen = body_inactive(w.bflags[(uint32_t)(key >> 32)]) && body_inactive(w.bflags[(uint32_t)key]);
    if (frozen && old < 0) {
        void_pair(w, key, t);
        return;
    }
    // The state lives in registers from the load to the store (every contact
    // slot index is resolved through selects).
    PairState s;
    if (old >= 0) s = w.pairs_old[old];
    else init_pair(w, key, s);
    s.flags &= ~(uint32_t)(PF_FROZEN | PF_DEAD | PF_VOID);
    if (frozen) {
        s.flags |= PF_FROZEN;  // kept as it was: both bodies inactive
    } else if (s.kind == BEPU_PAIR_BOX_BOX) {
        LdsList m{&lds_pd[threadIdx.x], &lds_id[threadIdx.x], 0, BLOCK};
        update_box_pair(w, s, m);
    } else {
        update_single_pair(w, s);
    }
    w.pairs_cur[t] = s;
    w.hot_cur[t] = hot_of(s);
    if (!hash_insert(w.hash_keys_cur, w.hash_vals_cur, w.hash_cap, key, t)) w.counters[ctr(CTR_OVERFLOW)] = 3;
    // A constraint (a pair with contacts) joins the islands of its two bodies
    // when both are dynamic; the union-find forest was reset by this step's
    // body pass.
    if (s.ccount > 0 && (w.bflags[s.a] & 1u) && (w.bflags[s.b] & 1u)) uf_union(w.uf_parent, s.a, s.b);
}

// Closed-form pairs (box-box, sphere-sphere, box-sphere) from k_pair_prep's
// lists, or in order when every pair is of one kind.  General convex pairs
// are tested by k_convex; the threads past the closed-form pairs enter them
// into this step's key -> slot table, off the path that k_convex waits on.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_narrowphase(DeviceWorld w) {
    __shared__ float4 lds_pd[8 * BLOCK];
    __shared__ int lds_id[8 * BLOCK];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (!w.pair_lists) {
        // Every pair is a closed-form pair of one kind: taken in order.
        if (k >= min(w.counters[ctr(CTR_PAIRS)], w.pair_cap)) return;
        narrowphase_pair<BLOCK>(w, k, lds_pd, lds_id);
        return;
    }
    const uint32_t nbox = w.counters[ctr(CTR_NP_BOX)], nsingle = w.counters[ctr(CTR_NP_SINGLE)];
    if (k >= nbox + nsingle) {
        uint32_t c = k - (nbox + nsingle);
        for (uint32_t q = 0; q < 6; ++q) {
            const uint32_t n = w.counters[ctr(convex_list_counter(q))];
            if (c < n) {
                const uint32_t t = w.conv_work[(size_t)(q >> 1) * w.pair_cap + ((q & 1) == 0 ? c : w.pair_cap - 1 - c)].x;
                if (!hash_insert(w.hash_keys_cur, w.hash_vals_cur, w.hash_cap, w.pair_keys[t], t))
                    w.counters[ctr(CTR_OVERFLOW)] = 3;
                return;
            }
            c -= n;
        }
        return;
    }
    const uint32_t t = w.np_work[k < nbox ? k : w.pair_cap - 1 - (k - nbox)];
    narrowphase_pair<BLOCK>(w, t, lds_pd, lds_id);
}

// Statistics of the last step, computed on request from its pair states
// (kept out of the step: counters shared by every workgroup serialise):
// constraints, contacts, deepest contact and colours used.
__global__ void __launch_bounds__(256) k_stats(DeviceWorld w, uint32_t npairs) {
    __shared__ uint32_t s_red[4][4];
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cons = 0, contacts = 0, colors = 0, testers = 0;
    float mp = 0;
    if (t < npairs) {
        const PairState& s = w.pairs_old[t];
        if (s.kind == BEPU_PAIR_MESH_CONVEX && !(s.flags & (PF_DEAD | PF_VOID))) testers = (uint32_t)s.tester;
        if (s.ccount > 0 && !(s.flags & (PF_DEAD | PF_VOID))) {
            cons = 1;
            contacts = (uint32_t)s.ccount;
            for (int i = 0; i < s.ccount; ++i) mp = fmaxf(mp, s.cpos[s.clist[i]].w);
            colors = (uint32_t)max(s.color + 1, 0);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cons += __shfl_xor(cons, off);
        contacts += __shfl_xor(contacts, off);
        colors = max(colors, (uint32_t)__shfl_xor((int)colors, off));
        mp = fmaxf(mp, __shfl_xor(mp, off));
        testers += __shfl_xor(testers, off);
    }
    __shared__ uint32_t s_testers[4];
    int wv = threadIdx.x >> 6;
    if (__lane_id() == 0) s_testers[wv] = testers;
    if (__lane_id() == 0) {
        s_red[wv][0] = cons;
        s_red[wv][1] = contacts;
        s_red[wv][2] = colors;
        s_red[wv][3] = __float_as_uint(mp);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) {
            cons += s_red[k][0];
            contacts += s_red[k][1];
            colors = max(colors, s_red[k][2]);
            mp = fmaxf(mp, __uint_as_float(s_red[k][3]));
            testers += s_testers[k];
        }
        if (testers) atomicAdd(&w.counters[ctr(CTR_MESH_TESTERS)], testers);
        if (cons) atomicAdd(&w.counters[ctr(CTR_CONS)], cons);
        if (contacts) atomicAdd(&w.counters[ctr(CTR_CONTACTS)], contacts);
        if (colors) atomicMax(&w.counters[ctr(CTR_COLORS)], colors);
        if (mp > 0) atomicMax(&w.counters[ctr(CTR_MAXPEN)], __float_as_uint(mp));
    }
}

void launch_stats(const DeviceWorld& w, uint32_t npairs, hipStream_t st) {
    launch_fill_u32(w.counters + ctr(CTR_CONS), 0u, 1, st);
    launch_fill_u32(w.counters + ctr(CTR_CONTACTS), 0u, 1, st);
    launch_fill_u32(w.counters + ctr(CTR_COLORS), 0u, 1, st);
    launch_fill_u32(w.counters + ctr(CTR_MAXPEN), 0u, 1, st);
    launch_fill_u32(w.counters + ctr(CTR_MESH_TESTERS), 0u, 1, st);
    if (npairs) hipLaunchKernelGGL(k_stats, dim3((npairs + 255) / 256), dim3(256), 0, st, w, npairs);
}

void launch_pair_prep(const DeviceWorld& w, hipStream_t st) {
    if (w.pair_cap == 0) return;
    hipLaunchKernelGGL(k_pair_prep, dim3((w.pair_cap + kPrepBlock - 1) / kPrepBlock), dim3(kPrepBlock), 0, st, w);
}

void launch_narrowphase(const DeviceWorld& w, uint32_t block, hipStream_t st) {
    if (w.pair_cap == 0) return;
    // The grid covers the pair capacity; threads past this step's pair count exit.
    if (block == 64) hipLaunchKernelGGL(k_narrowphase<64>, dim3((w.pair_cap + 63) / 64), dim3(64), 0, st, w);
    else if (block == 256) hipLaunchKernelGGL(k_narrowphase<256>, dim3((w.pair_cap + 255) / 256), dim3(256), 0, st, w);
    else hipLaunchKernelGGL(k_narrowphase<128>, dim3((w.pair_cap + 127) / 128), dim3(128), 0, st, w);
}

}  // namespace bepu
