// refcpu_collide.cpp -- oracle restatement of the convex collision tests on
// the step path: box-box SAT + clipping, sphere-sphere, box-sphere.
//
// TEST INFRASTRUCTURE ONLY (see refcpu.h).  Every expression keeps the
// operand order of the reference so that FP32 results are bit-identical to
// what the C# code computes on an IEEE binary32 machine.
#include <cfloat>
#include "refcpu.h"

namespace ref {

namespace {

const float kEpsilon = 1e-7f;   // Toolbox.Epsilon (BEPUutilities/Toolbox.cs:25)
const float kBigEpsilon = 1e-5f;  // Toolbox.BigEpsilon (Toolbox.cs:20)

inline V3 row(const M3& m, int i) {
    if (i == 0) return V3{m.m11, m.m12, m.m13};
    if (i == 1) return V3{m.m21, m.m22, m.m23};
    return V3{m.m31, m.m32, m.m33};
}
inline float at(const M3& m, int i, int j) {
    const float* p = &m.m11;
    return p[i * 3 + j];
}

// ---------------------------------------------------------------------------
// Edge-edge contact (BoxBoxCollider.cs:2392-2698)

// FindHighestIndices (BoxBoxCollider.cs:2866-2890)
void find_highest_indices(const float dots[4], int* hi, int* second) {
    int h = 0;
    float hv = dots[0];
    for (int i = 1; i < 4; i++)
        if (dots[i] > hv) { h = i; hv = dots[i]; }
    int s = 0;
    float sv = -FLT_MAX;
    for (int i = 0; i < 4; i++)
        if (i != h && dots[i] > sv) { s = i; sv = dots[i]; }
    *hi = h;
    *second = s;
}

// GetEdgeData (BoxBoxCollider.cs:2700-2864): edge `index` (0:--,1:-+,2:+-,3:++)
// of the four edges parallel to local axis `axis`.
void get_edge_data(int index, int axis, float x, float y, float z, V3* s, V3* e, int* sid, int* eid) {
    switch (index + axis * 4) {
    case 0: *s = V3{-x, -y, -z}; *sid = 0; *e = V3{x, -y, -z}; *eid = 4; break;
    case 1: *s = V3{-x, -y, z}; *sid = 1; *e = V3{x, -y, z}; *eid = 5; break;
    case 2: *s = V3{-x, y, -z}; *sid = 2; *e = V3{x, y, -z}; *eid = 6; break;
    case 3: *s = V3{-x, y, z}; *sid = 3; *e = V3{x, y, z}; *eid = 7; break;
    case 4: *s = V3{-x, -y, -z}; *sid = 0; *e = V3{-x, y, -z}; *eid = 2; break;
    case 5: *s = V3{-x, -y, z}; *sid = 1; *e = V3{-x, y, z}; *eid = 3; break;
    case 6: *s = V3{x, -y, -z}; *sid = 4; *e = V3{x, y, -z}; *eid = 6; break;
    case 7: *s = V3{x, -y, z}; *sid = 5; *e = V3{x, y, z}; *eid = 7; break;
    case 8: *s = V3{-x, -y, -z}; *sid = 0; *e = V3{-x, -y, z}; *eid = 1; break;
    case 9: *s = V3{-x, y, -z}; *sid = 2; *e = V3{-x, y, z}; *eid = 3; break;
    case 10: *s = V3{x, -y, -z}; *sid = 4; *e = V3{x, -y, z}; *eid = 5; break;
    default: *s = V3{x, y, -z}; *sid = 6; *e = V3{x, y, z}; *eid = 7; break;
    }
}

// GetEdgeId / GetContactId (BoxBoxCollider.cs:6275-6289)
inline int edge_id(int a, int b) { return (a + 1) * 571 + (b + 1) * 577; }
inline int contact_id4(int a1, int a2, int b1, int b2) {
    return edge_id(a1, a2) * 2549 + edge_id(b1, b2) * 2857;
}

// GetClosestPointsBetweenSegments (BoxBoxCollider.cs:2902-2990)
bool closest_points_segments(const V3& p1, const V3& q1, const V3& p2, const V3& q2, V3* c1, V3* c2) {
    V3 d1 = sub(q1, p1);
    V3 d2 = sub(q2, p2);
    V3 r = sub(p1, p2);
    float a = length_sq(d1);
    float e = length_sq(d2);
    float f = dot(d2, r);
    float s, t;
    if (a <= kEpsilon && e <= kEpsilon) {
        *c1 = p1;
        *c2 = p2;
        return false;
    }
    if (a <= kEpsilon) {
        s = 0.0f;
        t = f / e;
        if (t < 0 || t > 1) return false;
    } else {
        float c = dot(d1, r);
        if (e <= kEpsilon) {
            t = 0.0f;
            s = clampf(-c / a, 0.0f, 1.0f);
        } else {
            float b = dot(d1, d2);
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
    *c1 = add(mul(d1, s), p1);
    *c2 = add(mul(d2, t), p2);
    return true;
}

void pick_edges(const V3& m, float hx, float hy, float hz, V3* s1, V3* e1, int* s1id, int* e1id, V3* s2,
                V3* e2, int* s2id, int* e2id) {
    float dots[4];
    int axis;
    if (std::fabs(m.x) < kEpsilon) {
        dots[0] = -hy * m.y - hz * m.z;
        dots[1] = -hy * m.y + hz * m.z;
        dots[2] = hy * m.y - hz * m.z;
        dots[3] = hy * m.y + hz * m.z;
        axis = 0;
    } else if (std::fabs(m.y) < kEpsilon) {
        dots[0] = -hx * m.x - hz * m.z;
        dots[1] = -hx * m.x + hz * m.z;
        dots[2] = hx * m.x - hz * m.z;
        dots[3] = hx * m.x + hz * m.z;
        axis = 1;
    } else {
        dots[0] = -hx * m.x - hy * m.y;
        dots[1] = -hx * m.x + hy * m.y;
        dots[2] = hx * m.x - hy * m.y;
        dots[3] = hx * m.x + hy * m.y;
        axis = 2;
    }
    int h, s;
    find_highest_indices(dots, &h, &s);
    get_edge_data(h, axis, hx, hy, hz, s1, e1, s1id, e1id);
    get_edge_data(s, axis, hx, hy, hz, s2, e2, s2id, e2id);
}

void edge_edge_contact(const V3& ha, const V3& hb, const V3& posA, const M3& oA, const V3& posB,
                       const M3& oB, const V3& mtd, BoxContactList* out) {
    V3 mtdA = m3_transform_transpose(neg(mtd), oA);
    V3 mtdB = m3_transform_transpose(mtd, oB);
    V3 aS1, aE1, aS2, aE2, bS1, bE1, bS2, bE2;
    int aS1id, aE1id, aS2id, aE2id, bS1id, bE1id, bS2id, bE2id;
    pick_edges(mtdA, ha.x, ha.y, ha.z, &aS1, &aE1, &aS1id, &aE1id, &aS2, &aE2, &aS2id, &aE2id);
    pick_edges(mtdB, hb.x, hb.y, hb.z, &bS1, &bE1, &bS1id, &bE1id, &bS2, &bE2, &bS2id, &bE2id);
    aS1 = m3_transform(aS1, oA); aE1 = m3_transform(aE1, oA);
    bS1 = m3_transform(bS1, oB); bE1 = m3_transform(bE1, oB);
    aS2 = m3_transform(aS2, oA); aE2 = m3_transform(aE2, oA);
    bS2 = m3_transform(bS2, oB); bE2 = m3_transform(bE2, oB);
    aS1 = add(aS1, posA); aE1 = add(aE1, posA); bS1 = add(bS1, posB); bE1 = add(bE1, posB);
    aS2 = add(aS2, posA); aE2 = add(aE2, posA); bS2 = add(bS2, posB); bE2 = add(bE2, posB);
    out->n = 0;
    struct Seg { const V3 *s, *e; int sid, eid; };
    Seg as[2] = {{&aS1, &aE1, aS1id, aE1id}, {&aS2, &aE2, aS2id, aE2id}};
    Seg bs[2] = {{&bS1, &bE1, bS1id, bE1id}, {&bS2, &bE2, bS2id, bE2id}};
    // Order: (A1,B1), (A1,B2), (A2,B1), (A2,B2) (BoxBoxCollider.cs:2600-2690)
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            V3 onA, onB;
            if (closest_points_segments(*as[i].s, *as[i].e, *bs[j].s, *bs[j].e, &onA, &onB)) {
                V3 offset = sub(onA, onB);
                float d = dot(offset, mtd);
                if (d < 0) {
                    BoxContact c;
                    c.pos = onA;
                    c.depth = d;
                    c.id = contact_id4(as[i].sid, as[i].eid, bs[j].sid, bs[j].eid);
                    out->add(c);
                }
            }
        }
}

// ---------------------------------------------------------------------------
// Face contacts (BoxBoxCollider.cs:3459-3487, 4517-5278, 6007-6289)

struct BoxFace {
    int id[4];
    V3 v[4];
    V3 normal;
    float width, height;
};
struct FaceEdge {
    V3 a, b;
    float edge_distance;
    int id;
    V3 perp;
};

// Matrix.Transform(Vector3, Matrix) with a rotation + translation (Matrix.cs:717-725)
inline V3 xform4(const V3& c, const M3& o, const V3& p) {
    return V3{c.x * o.m11 + c.y * o.m21 + c.z * o.m31 + p.x, c.x * o.m12 + c.y * o.m22 + c.z * o.m32 + p.y,
              c.x * o.m13 + c.y * o.m23 + c.z * o.m33 + p.z};
}

// GetNearestFace (BoxBoxCollider.cs:6042-6177).  When no axis strictly dominates,
// the face stays zero-initialised as in the reference.
void nearest_face(const V3& position, const M3& o, const V3& mtd, float hw, float hh, float hl, BoxFace* f) {
    *f = BoxFace{};
    f->id[0] = f->id[1] = f->id[2] = f->id[3] = 0;
    f->v[0] = f->v[1] = f->v[2] = f->v[3] = V3{0, 0, 0};
    f->normal = V3{0, 0, 0};
    f->width = f->height = 0;
    float xDot = o.m11 * mtd.x + o.m12 * mtd.y + o.m13 * mtd.z;
    float yDot = o.m21 * mtd.x + o.m22 * mtd.y + o.m23 * mtd.z;
    float zDot = o.m31 * mtd.x + o.m32 * mtd.y + o.m33 * mtd.z;
    float absX = std::fabs(xDot), absY = std::fabs(yDot), absZ = std::fabs(zDot);
    int bit;
    if (absX > absY && absX > absZ) {
        if (xDot < 0) { hw = -hw; bit = 0; } else bit = 1;
        f->v[0] = xform4(V3{hw, hh, hl}, o, position);
        f->v[1] = xform4(V3{hw, -hh, hl}, o, position);
        f->v[2] = xform4(V3{hw, -hh, -hl}, o, position);
        f->v[3] = xform4(V3{hw, hh, -hl}, o, position);
        f->normal = xDot < 0 ? V3{-o.m11, -o.m12, -o.m13} : V3{o.m11, o.m12, o.m13};
        f->width = hh * 2;
        f->height = hl * 2;
        f->id[0] = bit + 2 + 4; f->id[1] = bit + 4; f->id[2] = bit + 2; f->id[3] = bit;
    } else if (absY > absX && absY > absZ) {
        if (yDot < 0) { hh = -hh; bit = 0; } else bit = 2;
        f->v[0] = xform4(V3{hw, hh, hl}, o, position);
        f->v[1] = xform4(V3{-hw, hh, hl}, o, position);
        f->v[2] = xform4(V3{-hw, hh, -hl}, o, position);
        f->v[3] = xform4(V3{hw, hh, -hl}, o, position);
        f->normal = yDot < 0 ? V3{-o.m21, -o.m22, -o.m23} : V3{o.m21, o.m22, o.m23};
        f->width = hw * 2;
        f->height = hl * 2;
        f->id[0] = 1 + bit + 4; f->id[1] = bit + 4; f->id[2] = 1 + bit; f->id[3] = bit;
    } else if (absZ > absX && absZ > absY) {
        if (zDot < 0) { hl = -hl; bit = 0; } else bit = 4;
        f->v[0] = xform4(V3{hw, hh, hl}, o, position);
        f->v[1] = xform4(V3{-hw, hh, hl}, o, position);
        f->v[2] = xform4(V3{-hw, -hh, hl}, o, position);
        f->v[3] = xform4(V3{hw, -hh, hl}, o, position);
        f->normal = zDot < 0 ? V3{-o.m31, -o.m32, -o.m33} : V3{o.m31, o.m32, o.m33};
        f->width = hw * 2;
        f->height = hh * 2;
        f->id[0] = 1 + 2 + bit; f->id[1] = 2 + bit; f->id[2] = 1 + bit; f->id[3] = bit;
    }
}

// BoxFace.GetEdge (BoxBoxCollider.cs:6220-6268)
FaceEdge face_edge(const BoxFace& f, int i) {
    FaceEdge e;
    V3 inside;
    int i0 = i, i1 = (i + 1) & 3, i2 = (i + 2) & 3;
    e.a = f.v[i0];
    e.b = f.v[i1];
    inside = f.v[i2];
    e.id = edge_id(f.id[i0], f.id[i1]);
    V3 dir = normalize(sub(e.b, e.a));
    e.perp = cross(dir, f.normal);
    V3 off = sub(inside, e.a);
    float d = dot(e.perp, off);
    if (d > 0) e.perp = V3{-e.perp.x, -e.perp.y, -e.perp.z};
    e.edge_distance = dot(e.a, e.perp);
    return e;
}

// ComputeIntersection (BoxBoxCollider.cs:6007-6040)
bool compute_intersection(const V3& a1, const V3& a2, const FaceEdge& ce, V3* out) {
    V3 offset = sub(ce.a, a1);
    V3 dir = sub(a2, a1);
    float distanceToPlane = dot(offset, ce.perp);
    float dirLen = dot(dir, ce.perp);
    float t = distanceToPlane / dirLen;
    if (t < 0 || t > 1) return false;
    offset = mul(dir, t);
    *out = add(offset, a1);
    offset = sub(*out, ce.a);
    dir = sub(ce.b, ce.a);
    t = dot(dir, offset);
    if (t < 0 || t > length_sq(dir)) return false;
    return true;
}

// ClipFacesDirect, non-ALLOWUNSAFE (BoxBoxCollider.cs:4517-5278)
void clip_faces(const BoxFace& cf, const BoxFace& face, const V3& mtd, BoxContactList* out) {
    out->n = 0;
    V3 clipX = sub(cf.v[3], cf.v[2]);
    V3 clipY = sub(cf.v[1], cf.v[2]);
    float invCW = 1 / cf.width;
    float invCH = 1 / cf.height;
    float invCW2 = invCW * invCW;
    clipX = V3{clipX.x * invCW2, clipX.y * invCW2, clipX.z * invCW2};
    float invCH2 = invCH * invCH;
    clipY = V3{clipY.x * invCH2, clipY.y * invCH2, clipY.z * invCH2};

    V3 faceX = sub(face.v[3], face.v[2]);
    V3 faceY = sub(face.v[1], face.v[2]);
    float invFW = 1 / face.width;
    float invFH = 1 / face.height;
    float invFW2 = invFW * invFW;
    faceX = V3{faceX.x * invFW2, faceX.y * invFW2, faceX.z * invFW2};
    float invFH2 = invFH * invFH;
    faceY = V3{faceY.x * invFH2, faceY.y * invFH2, faceY.z * invFH2};

    V3 clipCenter = add(cf.v[0], cf.v[2]);
    float clipCenterX = dot(clipCenter, clipX);
    float clipCenterY = dot(clipCenter, clipY);
    clipCenterX *= .5f;
    clipCenterY *= .5f;
    V3 faceCenter = add(face.v[0], face.v[2]);
    float faceCenterX = dot(faceCenter, faceX);
    float faceCenterY = dot(faceCenter, faceY);
    faceCenterX *= .5f;
    faceCenterY *= .5f;

    float extentX = .5f + .01f * invCW;
    float extentY = .5f + .01f * invCH;
    float clipMaxX = clipCenterX + extentX, clipMaxY = clipCenterY + extentY;
    float clipMinX = clipCenterX - extentX, clipMinY = clipCenterY - extentY;
    extentX = .5f + .01f * invFW;
    extentY = .5f + .01f * invFH;
    float faceMaxX = faceCenterX + extentX, faceMaxY = faceCenterY + extentY;
    float faceMinX = faceCenterX - extentX, faceMinY = faceCenterY - extentY;

    bool maxX[4], minX[4], maxY[4], minY[4];  // face vertices against the clip face
    for (int i = 0; i < 4; ++i) {
        float dx = dot(clipX, face.v[i]);
        maxX[i] = dx < clipMaxX;
        minX[i] = dx > clipMinX;
        float dy = dot(clipY, face.v[i]);
        maxY[i] = dy < clipMaxY;
        minY[i] = dy > clipMinY;
    }
    bool cmaxX[4], cminX[4], cmaxY[4], cminY[4];  // clip face vertices against the face
    for (int i = 0; i < 4; ++i) {
        float dx = dot(faceX, cf.v[i]);
        cmaxX[i] = dx < faceMaxX;
        cminX[i] = dx > faceMinX;
        float dy = dot(faceY, cf.v[i]);
        cmaxY[i] = dy < faceMaxY;
        cminY[i] = dy > faceMinY;
    }

    BoxContact item{V3{0, 0, 0}, 0, 0};
    for (int i = 0; i < 4; ++i)
        if (minX[i] && maxX[i] && minY[i] && maxY[i]) {
            item.pos = face.v[i];
            item.id = face.id[i];
            out->add(item);
        }
    // Depths of the face vertices.
    BoxContactList temp = *out;
    out->n = 0;
    float clipFaceDot = dot(cf.v[0], mtd);
    for (int i = 0; i < temp.n; ++i) {
        item = temp.e[i];
        float faceDot = dot(item.pos, mtd);
        item.depth = faceDot - clipFaceDot;
        if (item.depth <= 0) out->add(item);
    }
    int previousCount = out->n;
    if (previousCount >= 4) return;

    // Clip face vertices projected onto the face plane.
    float b = dot(face.v[0], face.normal);
    for (int i = 0; i < 4; ++i)
        if (cminX[i] && cmaxX[i] && cminY[i] && cmaxY[i]) {
            float a = dot(cf.v[i], face.normal);
            V3 v = mul(face.normal, a - b);
            v = sub(cf.v[i], v);
            item.pos = v;
            item.id = cf.id[i] + 8;
            out->add(item);
        }
    {
        temp = *out;
        int post = out->n;
        for (int i = post - 1; i >= previousCount; i--) out->remove_at(i);
        for (int i = previousCount; i < temp.n; i++) {
            item = temp.e[i];
            float faceDot = dot(item.pos, mtd);
            item.depth = faceDot - clipFaceDot;
            if (item.depth <= 0) out->add(item);
        }
    }
    previousCount = out->n;
    if (previousCount >= 4) return;

    // Face edges intersected with the four clip edges.  Edge k of the clip face
    // is bounded by flag set k: 0 -> maxY, 1 -> minX, 2 -> minY, 3 -> maxX.
    // For an outside face vertex the two neighbouring face edges are tried in the
    // reference's order; vertex 4 tries (v1, v3) except on clip edge 2 where it
    // tries (v3, v1) (BoxBoxCollider.cs:4785-5262).
    const bool* flags[4] = {maxY, minX, minY, maxX};
    // Face edge (from, to) by neighbour: edge between vertex i and neighbour j.
    auto seg = [&](int i, int j, int* s, int* e) {
        // edges V1V2 (0,1), V2V3 (1,2), V3V4 (2,3), V4V1 (3,0)
        int lo = i < j ? i : j, hi = i < j ? j : i;
        if (lo == 0 && hi == 3) { *s = 3; *e = 0; }
        else { *s = lo; *e = hi; }
    };
    for (int k = 0; k < 4; ++k) {
        FaceEdge ce = face_edge(cf, k);
        const bool* F = flags[k];
        for (int i = 0; i < 4; ++i) {
            if (F[i]) continue;
            int n1, n2;
            if (i == 0) { n1 = 1; n2 = 3; }
            else if (i == 1) { n1 = 0; n2 = 2; }
            else if (i == 2) { n1 = 1; n2 = 3; }
            else if (k == 2) { n1 = 2; n2 = 0; }
            else { n1 = 0; n2 = 2; }
            int nb[2] = {n1, n2};
            for (int q = 0; q < 2; ++q) {
                int j = nb[q];
                if (!F[j] || out->n >= 8) continue;
                int s, e;
                seg(i, j, &s, &e);
                V3 v;
                if (compute_intersection(face.v[s], face.v[e], ce, &v)) {
                    item.pos = v;
                    item.id = edge_id(face.id[s], face.id[e]) * 2549 + ce.id * 2857;
                    out->add(item);
                }
            }
        }
    }
    {
        temp = *out;
        int post = out->n;
        for (int i = post - 1; i >= previousCount; i--) out->remove_at(i);
        for (int i = previousCount; i < temp.n; i++) {
            item = temp.e[i];
            float faceDot = dot(item.pos, mtd);
            item.depth = faceDot - clipFaceDot;
            if (item.depth <= 0) out->add(item);
        }
    }
}

// PruneContactsMaxDistance, non-ALLOWUNSAFE (BoxBoxCollider.cs:3612-3741)
void prune_max_distance(const V3& mtd, const BoxContactList& in, BoxContactList* out) {
    int count = in.n;
    BoxContact deepest = in.e[0];
    for (int i = 1; i < count; i++)
        if (in.e[i].depth > deepest.depth) deepest = in.e[i];
    BoxContact furthest = in.e[0];
    float furthestDistance = distance_sq(deepest.pos, furthest.pos);
    for (int i = 1; i < count; i++) {
        float d = distance_sq(deepest.pos, in.e[i].pos);
        if (d > furthestDistance) { furthestDistance = d; furthest = in.e[i]; }
    }
    V3 xAxis = sub(furthest.pos, deepest.pos);
    V3 yAxis = cross(mtd, xAxis);
    BoxContact minD = in.e[0], maxD = in.e[0];
    float minY = dot(minD.pos, yAxis);
    float maxY = minY;
    for (int i = 1; i < count; i++) {
        float d = dot(yAxis, in.e[i].pos);
        if (d < minY) { minY = d; minD = in.e[i]; }
        else if (d > maxY) { maxY = d; maxD = in.e[i]; }
    }
    BoxContactList o;
    o.add(deepest);
    o.add(furthest);
    o.add(minD);
    o.add(maxD);
    *out = o;
}

void face_contacts(const V3& ha, const V3& hb, const V3& posA, const M3& oA, const V3& posB, const M3& oB,
                   bool aIsFaceOwner, const V3& mtd, BoxContactList* out) {
    BoxFace fa, fb;
    V3 nm = neg(mtd);
    nearest_face(posA, oA, nm, ha.x, ha.y, ha.z, &fa);
    nearest_face(posB, oB, mtd, hb.x, hb.y, hb.z, &fb);
    if (aIsFaceOwner) clip_faces(fa, fb, nm, out);
    else clip_faces(fb, fa, mtd, out);
    if (out->n > 4) {
        BoxContactList pruned;
        prune_max_distance(mtd, *out, &pruned);
        *out = pruned;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// BoxBoxCollider.AreBoxesColliding (BoxBoxCollider.cs:1494-2386)
bool box_box_colliding(const V3& ha, const V3& hb, const V3& posA, const Q4& qA, const V3& posB,
                       const Q4& qB, float* distance, V3* axis, BoxContactList* contacts) {
    const float aX = ha.x, aY = ha.y, aZ = ha.z;
    const float bX = hb.x, bY = hb.y, bZ = hb.z;
    contacts->n = 0;
    M3 aO = m3_from_quat(qA);
    M3 bO = m3_from_quat(qB);
    V3 t = sub(posB, posA);
    float tempDistance;
    float minimumDistance = -FLT_MAX;
    V3 minimumAxis{0, 0, 0};
    int minimumFeature = 2;
    M3 bR, absBR;

    // A face normals (BoxBoxCollider.cs:1528-1658)
    bR.m11 = aO.m11 * bO.m11 + aO.m12 * bO.m12 + aO.m13 * bO.m13;
    bR.m12 = aO.m11 * bO.m21 + aO.m12 * bO.m22 + aO.m13 * bO.m23;
    bR.m13 = aO.m11 * bO.m31 + aO.m12 * bO.m32 + aO.m13 * bO.m33;
    absBR.m11 = std::fabs(bR.m11) + kEpsilon;
    absBR.m12 = std::fabs(bR.m12) + kEpsilon;
    absBR.m13 = std::fabs(bR.m13) + kEpsilon;
    float tX = t.x;
    t.x = t.x * aO.m11 + t.y * aO.m12 + t.z * aO.m13;

    auto faceTest = [&](float tv, float rarb, const V3& r, int feature) -> bool {
        if (tv > rarb) { *distance = tv - rarb; *axis = neg(r); return false; }
        if (tv < -rarb) { *distance = -tv - rarb; *axis = r; return false; }
        if (tv > 0) {
            tempDistance = tv - rarb;
            if (tempDistance > minimumDistance) {
                minimumDistance = tempDistance;
                minimumAxis = neg(r);
                minimumFeature = feature;
            }
        } else {
            tempDistance = -tv - rarb;
            if (tempDistance > minimumDistance) {
                minimumDistance = tempDistance;
                minimumAxis = r;
                minimumFeature = feature;
            }
        }
        return true;
    };

    float rarb = aX + bX * absBR.m11 + bY * absBR.m12 + bZ * absBR.m13;
    if (!faceTest(t.x, rarb, row(aO, 0), 0)) return false;

    bR.m21 = aO.m21 * bO.m11 + aO.m22 * bO.m12 + aO.m23 * bO.m13;
    bR.m22 = aO.m21 * bO.m21 + aO.m22 * bO.m22 + aO.m23 * bO.m23;
    bR.m23 = aO.m21 * bO.m31 + aO.m22 * bO.m32 + aO.m23 * bO.m33;
    absBR.m21 = std::fabs(bR.m21) + kEpsilon;
    absBR.m22 = std::fabs(bR.m22) + kEpsilon;
    absBR.m23 = std::fabs(bR.m23) + kEpsilon;
    float tY = t.y;
    t.y = tX * aO.m21 + t.y * aO.m22 + t.z * aO.m23;
    rarb = aY + bX * absBR.m21 + bY * absBR.m22 + bZ * absBR.m23;
    if (!faceTest(t.y, rarb, row(aO, 1), 0)) return false;

    bR.m31 = aO.m31 * bO.m11 + aO.m32 * bO.m12 + aO.m33 * bO.m13;
    bR.m32 = aO.m31 * bO.m21 + aO.m32 * bO.m22 + aO.m33 * bO.m23;
    bR.m33 = aO.m31 * bO.m31 + aO.m32 * bO.m32 + aO.m33 * bO.m33;
    absBR.m31 = std::fabs(bR.m31) + kEpsilon;
    absBR.m32 = std::fabs(bR.m32) + kEpsilon;
    absBR.m33 = std::fabs(bR.m33) + kEpsilon;
    t.z = tX * aO.m31 + tY * aO.m32 + t.z * aO.m33;
    rarb = aZ + bX * absBR.m31 + bY * absBR.m32 + bZ * absBR.m33;
    if (!faceTest(t.z, rarb, row(aO, 2), 0)) return false;

    const float antiBBias = .01f;
    minimumDistance += antiBBias;

    // B face normals (BoxBoxCollider.cs:1662-1784)
    rarb = bX + aX * absBR.m11 + aY * absBR.m21 + aZ * absBR.m31;
    float tl = t.x * bR.m11 + t.y * bR.m21 + t.z * bR.m31;
    if (!faceTest(tl, rarb, row(bO, 0), 1)) return false;
    rarb = bY + aX * absBR.m12 + aY * absBR.m22 + aZ * absBR.m32;
    tl = t.x * bR.m12 + t.y * bR.m22 + t.z * bR.m32;
    if (!faceTest(tl, rarb, row(bO, 1), 1)) return false;
    rarb = bZ + aX * absBR.m13 + aY * absBR.m23 + aZ * absBR.m33;
    tl = t.x * bR.m13 + t.y * bR.m23 + t.z * bR.m33;
    if (!faceTest(tl, rarb, row(bO, 2), 1)) return false;

    if (minimumFeature != 1) minimumDistance -= antiBBias;
    float antiEdgeBias = .01f;
    minimumDistance += antiEdgeBias;

    // Edge axes A.i x B.j (BoxBoxCollider.cs:1795-2352).  For each (i, j):
    // rarb = h1*|R[p1]| + h2*|R[p2]| + h3*|R[p3]| + h4*|R[p4]| (left to right),
    // tl = t[u]*R[q1] - t[v]*R[q2], separation is only tested when
    // |R[j][i]| < 1 (the reference indexes the transposed entry), and the
    // candidate axis is cross(b_j, a_i) when tl > 0, else cross(a_i, b_j).
    struct EdgeSpec {
        int h[4];     // 0..2 -> aX..aZ, 3..5 -> bX..bZ
        int p[4][2];  // absBR indices
        int tu, q1[2], tv, q2[2];
    };
    static const EdgeSpec specs[9] = {
        // A.X x B.X
        {{1, 2, 4, 5}, {{2, 0}, {1, 0}, {0, 2}, {0, 1}}, 2, {1, 0}, 1, {2, 0}},
        // A.X x B.Y
        {{1, 2, 3, 5}, {{2, 1}, {1, 1}, {0, 2}, {0, 0}}, 2, {1, 1}, 1, {2, 1}},
        // A.X x B.Z
        {{1, 2, 3, 4}, {{2, 2}, {1, 2}, {0, 1}, {0, 0}}, 2, {1, 2}, 1, {2, 2}},
        // A.Y x B.X
        {{0, 2, 4, 5}, {{2, 0}, {0, 0}, {1, 2}, {1, 1}}, 0, {2, 0}, 2, {0, 0}},
        // A.Y x B.Y
        {{0, 2, 3, 5}, {{2, 1}, {0, 1}, {1, 2}, {1, 0}}, 0, {2, 1}, 2, {0, 1}},
        // A.Y x B.Z
        {{0, 2, 3, 4}, {{2, 2}, {0, 2}, {1, 1}, {1, 0}}, 0, {2, 2}, 2, {0, 2}},
        // A.Z x B.X
        {{0, 1, 4, 5}, {{1, 0}, {0, 0}, {2, 2}, {2, 1}}, 1, {0, 0}, 0, {1, 0}},
        // A.Z x B.Y
        {{0, 1, 3, 5}, {{1, 1}, {0, 1}, {2, 2}, {2, 0}}, 1, {0, 1}, 0, {1, 1}},
        // A.Z x B.Z
        {{0, 1, 3, 4}, {{1, 2}, {0, 2}, {2, 1}, {2, 0}}, 1, {0, 2}, 0, {1, 2}},
    };
    const float hs[6] = {aX, aY, aZ, bX, bY, bZ};
    const float tt[3] = {t.x, t.y, t.z};
    for (int e = 0; e < 9; ++e) {
        const EdgeSpec& s = specs[e];
        int i = e / 3, j = e % 3;
        rarb = hs[s.h[0]] * at(absBR, s.p[0][0], s.p[0][1]) + hs[s.h[1]] * at(absBR, s.p[1][0], s.p[1][1]) +
               hs[s.h[2]] * at(absBR, s.p[2][0], s.p[2][1]) + hs[s.h[3]] * at(absBR, s.p[3][0], s.p[3][1]);
        tl = tt[s.tu] * at(bR, s.q1[0], s.q1[1]) - tt[s.tv] * at(bR, s.q2[0], s.q2[1]);
        V3 a = row(aO, i), b = row(bO, j);
        if (at(absBR, j, i) < 1) {
            if (tl > rarb) {
                *distance = tl - rarb;
                *axis = V3{b.y * a.z - b.z * a.y, b.z * a.x - b.x * a.z, b.x * a.y - b.y * a.x};
                return false;
            }
            if (tl < -rarb) {
                *distance = -tl - rarb;
                *axis = V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
                return false;
            }
        }
        V3 tempAxis;
        float inv;
        if (tl > 0) {
            tempAxis = V3{b.y * a.z - b.z * a.y, b.z * a.x - b.x * a.z, b.x * a.y - b.y * a.x};
            inv = 1 / length(tempAxis);
            tempDistance = (tl - rarb) * inv;
        } else {
            tempAxis = V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
            inv = 1 / length(tempAxis);
            tempDistance = (-tl - rarb) * inv;
        }
        if (tempDistance > minimumDistance) {
            minimumFeature = 2;
            minimumDistance = tempDistance;
            tempAxis.x *= inv;
            tempAxis.y *= inv;
            tempAxis.z *= inv;
            minimumAxis = tempAxis;
        }
    }

    if (minimumFeature == 2) {
        edge_edge_contact(ha, hb, posA, aO, posB, bO, minimumAxis, contacts);
    } else {
        minimumDistance -= antiEdgeBias;
        face_contacts(ha, hb, posA, aO, posB, bO, minimumFeature == 0, minimumAxis, contacts);
    }
    *distance = minimumDistance;
    *axis = minimumAxis;
    return true;
}

// SphereTester.AreSpheresColliding (SphereTester.cs:23-56)
bool sphere_sphere_colliding(float ra, float rb, const V3& pa, const V3& pb, float maxd, ContactData* c) {
    float radiusSum = ra + rb;
    V3 diff = sub(pb, pa);
    float centerDistance = length_sq(diff);
    if (centerDistance < (radiusSum + maxd) * (radiusSum + maxd)) {
        if (radiusSum > kEpsilon) c->pos = mul(diff, ra / radiusSum);
        else c->pos = V3{0, 0, 0};
        c->pos = add(c->pos, pa);
        centerDistance = std::sqrt(centerDistance);
        if (centerDistance > kBigEpsilon) c->normal = divide(diff, centerDistance);
        else c->normal = V3{0, 1, 0};
        c->depth = radiusSum - centerDistance;
        return true;
    }
    return false;
}

// BoxSphereTester.AreShapesColliding (BoxSphereTester.cs:23-85)
bool box_sphere_colliding(const V3& h, const V3& boxPos, const Q4& boxQ, float radius, const V3& spherePos,
                          float maxd, ContactData* c) {
    // RigidTransform.TransformByInverse (RigidTransform.cs:149-156)
    V3 local = qtransform(sub(spherePos, boxPos), qconj(boxQ));
    V3 lc;
    lc.x = clampf(local.x, -h.x, h.x);
    lc.y = clampf(local.y, -h.y, h.y);
    lc.z = clampf(local.z, -h.z, h.z);
    // RigidTransform.Transform (RigidTransform.cs:135-140)
    c->pos = add(qtransform(lc, boxQ), boxPos);
    V3 offset = sub(spherePos, c->pos);
    float offsetLength = length_sq(offset);
    if (offsetLength > (radius + maxd) * (radius + maxd)) return false;
    if (offsetLength > kEpsilon) {
        offsetLength = std::sqrt(offsetLength);
        c->normal = divide(offset, offsetLength);
        c->depth = radius - offsetLength;
    } else {
        V3 pd;
        pd.x = lc.x < 0 ? lc.x + h.x : h.x - lc.x;
        pd.y = lc.y < 0 ? lc.y + h.y : h.y - lc.y;
        pd.z = lc.z < 0 ? lc.z + h.z : h.z - lc.z;
        if (pd.x < pd.y && pd.x < pd.z) {
            c->normal = lc.x > 0 ? V3{1, 0, 0} : V3{-1, 0, 0};
            c->depth = pd.x;
        } else if (pd.y < pd.z) {
            c->normal = lc.y > 0 ? V3{0, 1, 0} : V3{0, -1, 0};
            c->depth = pd.y;
        } else {
            c->normal = lc.z > 0 ? V3{0, 0, 1} : V3{0, 0, -1};
            c->depth = pd.z;
        }
        c->depth += radius;
        c->normal = qtransform(c->normal, boxQ);
    }
    return true;
}

}  // namespace ref
