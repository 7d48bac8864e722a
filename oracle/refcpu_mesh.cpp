// refcpu_mesh.cpp -- oracle restatement of the static mesh's data and its
// bounding box tree: TransformableMeshData (world-space triangles through an
// affine transform) and MeshBoundingBoxTree (insertion build over a permuted
// triangle order, depth-first box queries).
//
// TEST INFRASTRUCTURE ONLY (see refcpu.h).  Each expression keeps the operand
// order of the reference.
#include <algorithm>
#include <cfloat>
#include <cstdint>
#include "refcpu.h"

namespace ref {

// MeshBoundingBoxTree.LeafMargin (MeshBoundingBoxTree.cs:385)
static const float kLeafMargin = .001f;

// AffineTransform.Transform (BEPUutilities/AffineTransform.cs:148-152)
V3 affine_transform(const V3& p, const Affine& a) { return add(m3_transform(p, a.m), a.t); }

// AffineTransform.Multiply (AffineTransform.cs:186-194)
Affine affine_multiply(const Affine& a, const Affine& b) {
    Affine r;
    r.m = m3_mul(a.m, b.m);
    r.t = add(m3_transform(a.t, b.m), b.t);
    return r;
}

// Toolbox.GetTriangleBoundingBox (BEPUutilities/Toolbox.cs:1516-1580), one axis:
// the largest of the three, and the smaller of the other two.
static void tri_axis(float a, float b, float c, float* mn, float* mx) {
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
void triangle_bounding_box(const V3& a, const V3& b, const V3& c, V3* mn, V3* mx) {
    tri_axis(a.x, b.x, c.x, &mn->x, &mx->x);
    tri_axis(a.y, b.y, c.y, &mn->y, &mx->y);
    tri_axis(a.z, b.z, c.z, &mn->z, &mx->z);
}

void mesh_world_triangle(const MeshShape& m, int tri, V3 out[3]) {
    for (int k = 0; k < 3; ++k) out[k] = affine_transform(m.verts[m.idx[tri + k]], m.xf);
}

namespace {

struct Box {
    V3 mn, mx;
};

// BoundingBox.CreateMerged (BEPUutilities/BoundingBox.cs:187-214)
Box merged(const Box& a, const Box& b) {
    Box r;
    r.mn.x = a.mn.x < b.mn.x ? a.mn.x : b.mn.x;
    r.mn.y = a.mn.y < b.mn.y ? a.mn.y : b.mn.y;
    r.mn.z = a.mn.z < b.mn.z ? a.mn.z : b.mn.z;
    r.mx.x = a.mx.x > b.mx.x ? a.mx.x : b.mx.x;
    r.mx.y = a.mx.y > b.mx.y ? a.mx.y : b.mx.y;
    r.mx.z = a.mx.z > b.mx.z ? a.mx.z : b.mx.z;
    return r;
}

float volume(const Box& b) {
    V3 o = sub(b.mx, b.mn);
    return o.x * o.y * o.z;
}

// LeafNode constructor (MeshBoundingBoxTree.cs:397-409) with
// MeshBoundingBoxTreeData.GetBoundingBox (MeshBoundingBoxTreeData.cs:48-56).
Box leaf_box(const MeshShape& m, int tri) {
    V3 v[3];
    mesh_world_triangle(m, tri, v);
    Box b;
    b.mn.x = v[0].x < v[1].x ? v[0].x : v[1].x;
    b.mn.y = v[0].y < v[1].y ? v[0].y : v[1].y;
    b.mn.z = v[0].z < v[1].z ? v[0].z : v[1].z;
    b.mn.x = b.mn.x < v[2].x ? b.mn.x : v[2].x;
    b.mn.y = b.mn.y < v[2].y ? b.mn.y : v[2].y;
    b.mn.z = b.mn.z < v[2].z ? b.mn.z : v[2].z;
    b.mx.x = v[0].x > v[1].x ? v[0].x : v[1].x;
    b.mx.y = v[0].y > v[1].y ? v[0].y : v[1].y;
    b.mx.z = v[0].z > v[1].z ? v[0].z : v[1].z;
    b.mx.x = b.mx.x > v[2].x ? b.mx.x : v[2].x;
    b.mx.y = b.mx.y > v[2].y ? b.mx.y : v[2].y;
    b.mx.z = b.mx.z > v[2].z ? b.mx.z : v[2].z;
    b.mx.x += kLeafMargin;
    b.mx.y += kLeafMargin;
    b.mx.z += kLeafMargin;
    b.mn.x -= kLeafMargin;
    b.mn.y -= kLeafMargin;
    b.mn.z -= kLeafMargin;
    return b;
}

Box box_of(const MeshNode& n) { return Box{n.bmin, n.bmax}; }
void set_box(MeshNode& n, const Box& b) {
    n.bmin = b.mn;
    n.bmax = b.mx;
}

// BoundingBox.Intersects(ref BoundingBox, out bool) (BoundingBox.cs:73-86), `n` being `this`.
bool intersects(const MeshNode& n, const V3& mn, const V3& mx) {
    if (mn.x > n.bmax.x || mn.y > n.bmax.y || mn.z > n.bmax.z) return false;
    if (n.bmin.x > mx.x || n.bmin.y > mx.y || n.bmin.z > mx.z) return false;
    return true;
}

}  // namespace

// Insert (MeshBoundingBoxTree.cs:95-118), LeafNode.TryToInsert (:424-432) and
// InternalNode.TryToInsert (:252-338): a new leaf walks down from the root,
// each step taking the child whose volume grows least (B on ties) and
// widening its box, until that child is a leaf, which it pairs with.
void mesh_build_tree(MeshShape* m) {
    m->nodes.clear();
    const int64_t ntri = (int64_t)(m->idx.size() / 3);
    if (ntri == 0) return;
    m->nodes.reserve((size_t)(2 * ntri));
    int root = -1;
    for (int64_t i = 0; i < (int64_t)m->idx.size(); i += 3) {
        const int tri = (int)(((982451653LL * (i / 3)) % ntri) * 3);
        MeshNode leaf;
        set_box(leaf, leaf_box(*m, tri));
        leaf.a = -1;
        leaf.b = tri;
        const int leaf_id = (int)m->nodes.size();
        m->nodes.push_back(leaf);
        if (root < 0) {
            root = leaf_id;
            continue;
        }
        if (m->nodes[root].a < 0) {
            MeshNode in;
            set_box(in, merged(box_of(m->nodes[root]), box_of(m->nodes[leaf_id])));
            in.a = root;
            in.b = leaf_id;
            root = (int)m->nodes.size();
            m->nodes.push_back(in);
            continue;
        }
        set_box(m->nodes[root], merged(box_of(m->nodes[leaf_id]), box_of(m->nodes[root])));
        int node = root;
        const Box nb = box_of(m->nodes[leaf_id]);
        while (true) {
            const int ca = m->nodes[node].a, cb = m->nodes[node].b;
            const Box ba = box_of(m->nodes[ca]), bb = box_of(m->nodes[cb]);
            const Box mA = merged(ba, nb), mB = merged(bb, nb);
            const float oA = volume(ba), oB = volume(bb), vA = volume(mA), vB = volume(mB);
            if (vA - oA < vB - oB) {
                if (m->nodes[ca].a < 0) {
                    MeshNode in;
                    set_box(in, mA);
                    in.a = ca;
                    in.b = leaf_id;
                    const int id = (int)m->nodes.size();
                    m->nodes.push_back(in);
                    m->nodes[node].a = id;
                    break;
                }
                set_box(m->nodes[ca], mA);
                node = ca;
            } else {
                if (m->nodes[cb].a < 0) {
                    MeshNode in;
                    set_box(in, mB);
                    in.a = leaf_id;
                    in.b = cb;
                    const int id = (int)m->nodes.size();
                    m->nodes.push_back(in);
                    m->nodes[node].b = id;
                    break;
                }
                set_box(m->nodes[cb], mB);
                node = cb;
            }
        }
    }
    // Renumber so that the root is node 0 (children keep their ids otherwise).
    if (root != 0) {
        std::vector<int> id(m->nodes.size());
        for (size_t k = 0; k < id.size(); ++k) id[k] = (int)k;
        id[root] = 0;
        id[0] = root;
        std::swap(m->nodes[0], m->nodes[root]);
        for (MeshNode& n : m->nodes)
            if (n.a >= 0) {
                n.a = id[n.a];
                n.b = id[n.b];
            }
    }
}

// GetOverlaps(BoundingBox) (MeshBoundingBoxTree.cs:112-122) with
// InternalNode.GetOverlaps (:214-226): the root's box is tested first, then
// at each internal node child A's subtree before child B's.
void mesh_overlaps(const MeshShape& m, const V3& bmin, const V3& bmax, std::vector<int>* out) {
    out->clear();
    if (m.nodes.empty() || !intersects(m.nodes[0], bmin, bmax)) return;
    std::vector<int> stack;
    stack.push_back(0);
    while (!stack.empty()) {
        const MeshNode& n = m.nodes[stack.back()];
        stack.pop_back();
        if (n.a < 0) {
            out->push_back(n.b);
            continue;
        }
        if (intersects(m.nodes[n.b], bmin, bmax)) stack.push_back(n.b);
        if (intersects(m.nodes[n.a], bmin, bmax)) stack.push_back(n.a);
    }
}

// ---------------------------------------------------------------------------
// Terrains

// Matrix3x3.Determinant (Matrix3x3.cs:1231-1235) and Matrix3x3.Invert (:505-531)
M3 m3_invert(const M3& m) {
    float det = m.m11 * m.m22 * m.m33 + m.m12 * m.m23 * m.m31 + m.m13 * m.m21 * m.m32 - m.m31 * m.m22 * m.m13 -
                m.m32 * m.m23 * m.m11 - m.m33 * m.m21 * m.m12;
    float di = 1 / det;
    M3 r;
    r.m11 = (m.m22 * m.m33 - m.m23 * m.m32) * di;
    r.m12 = (m.m13 * m.m32 - m.m33 * m.m12) * di;
    r.m13 = (m.m12 * m.m23 - m.m22 * m.m13) * di;
    r.m21 = (m.m23 * m.m31 - m.m21 * m.m33) * di;
    r.m22 = (m.m11 * m.m33 - m.m13 * m.m31) * di;
    r.m23 = (m.m13 * m.m21 - m.m11 * m.m23) * di;
    r.m31 = (m.m21 * m.m32 - m.m22 * m.m31) * di;
    r.m32 = (m.m12 * m.m31 - m.m11 * m.m32) * di;
    r.m33 = (m.m11 * m.m22 - m.m12 * m.m21) * di;
    return r;
}

Affine affine_invert(const Affine& a) {
    Affine r;
    r.m = m3_invert(a.m);
    r.t = neg(m3_transform(a.t, r.m));
    return r;
}

int terrain_sidedness(const Affine& xf) {
    // Matrix3x3.AdjugateTranspose (Matrix3x3.cs:673-705): its Up row is
    // (m21, m22, m23) of the transposed adjugate, i.e. (m12, m22, m32) of the adjugate.
    const M3& m = xf.m;
    V3 up_n{m.m13 * m.m32 - m.m33 * m.m12, m.m11 * m.m33 - m.m13 * m.m31, m.m12 * m.m31 - m.m11 * m.m32};
    V3 up{m.m21, m.m22, m.m23};
    return dot(up_n, up) < 0 ? SIDED_CLOCKWISE : SIDED_COUNTERCLOCKWISE;
}

void terrain_bounding_box(const TerrainShapeRef& t, V3* mn, V3* mx) {
    float minX = FLT_MAX, maxX = -FLT_MAX, minY = FLT_MAX, maxY = -FLT_MAX, minZ = FLT_MAX, maxZ = -FLT_MAX;
    V3 minXv{0, 0, 0}, maxXv{0, 0, 0}, minYv{0, 0, 0}, maxYv{0, 0, 0}, minZv{0, 0, 0}, maxZv{0, 0, 0};
    for (int i = 0; i < t.width; i++)
        for (int j = 0; j < t.length; j++) {
            V3 v = m3_transform(V3{(float)i, t.h(i, j), (float)j}, t.xf.m);
            if (v.x < minX) { minX = v.x; minXv = v; } else if (v.x > maxX) { maxX = v.x; maxXv = v; }
            if (v.y < minY) { minY = v.y; minYv = v; } else if (v.y > maxY) { maxY = v.y; maxYv = v; }
            if (v.z < minZ) { minZ = v.z; minZv = v; } else if (v.z > maxZ) { maxZ = v.z; maxZv = v; }
        }
    mn->x = minXv.x + t.xf.t.x;
    mn->y = minYv.y + t.xf.t.y;
    mn->z = minZv.z + t.xf.t.z;
    mx->x = maxXv.x + t.xf.t.x;
    mx->y = maxYv.y + t.xf.t.y;
    mx->z = maxZv.z + t.xf.t.z;
}

void terrain_overlaps(const TerrainShapeRef& t, const V3& lmin, const V3& lmax, std::vector<int>* out) {
    out->clear();
    const int width = t.width;
    const int minX = std::max((int)lmin.x, 0), minY = std::max((int)lmin.z, 0);
    const int maxX = std::min((int)lmax.x, width - 2), maxY = std::min((int)lmax.z, t.length - 2);
    for (int i = minX; i <= maxX; i++)
        for (int j = minY; j <= maxY; j++) {
            float y1 = t.h(i, j), y2 = t.h(i + 1, j), y3 = t.h(i, j + 1), y4 = t.h(i + 1, j + 1);
            float highest = y1, lowest = y1;
            if (y2 > highest) highest = y2; else if (y2 < lowest) lowest = y2;
            if (y3 > highest) highest = y3; else if (y3 < lowest) lowest = y3;
            if (y4 > highest) highest = y4; else if (y4 < lowest) lowest = y4;
            if (lmax.y < lowest || lmin.y > highest) continue;
            int quad = (i + j * width) << 1;
            out->push_back(quad);
            out->push_back(quad | 1);
        }
}

void terrain_local_indices(const TerrainShapeRef& t, int tri, int col[3], int row[3]) {
    int quad = tri / 2;
    int r = quad / t.width;
    int c = quad - r * t.width;
    const bool blur = t.organization == 0;
    if ((tri & 1) == 0) {
        col[0] = c; row[0] = r;
        col[1] = c + 1; row[1] = r;
        if (blur) { col[2] = c; row[2] = r + 1; } else { col[2] = c + 1; row[2] = r + 1; }
    } else if (blur) {
        col[0] = c; row[0] = r + 1;
        col[2] = c + 1; row[2] = r + 1;
        col[1] = c + 1; row[1] = r;
    } else {
        col[0] = c; row[0] = r;
        col[1] = c + 1; row[1] = r + 1;
        col[2] = c; row[2] = r + 1;
    }
}

// TerrainShape.GetPosition (TerrainShape.cs:432-452): clamped indices, world space.
static V3 terrain_position(const TerrainShapeRef& t, int c, int r) {
    if (c <= 0) c = 0;
    else if (c >= t.width) c = t.width - 1;
    if (r <= 0) r = 0;
    else if (r >= t.length) r = t.length - 1;
    return affine_transform(V3{(float)c, t.h(c, r), (float)r}, t.xf);
}

void terrain_world_triangle(const TerrainShapeRef& t, int tri, V3 out[3]) {
    int quad = tri / 2;
    int r = quad / t.width;
    int c = quad - r * t.width;
    const bool blur = t.organization == 0;
    if ((tri & 1) == 0) {
        out[0] = terrain_position(t, c, r);
        out[1] = terrain_position(t, c + 1, r);
        out[2] = blur ? terrain_position(t, c, r + 1) : terrain_position(t, c + 1, r + 1);
    } else if (blur) {
        out[0] = terrain_position(t, c, r + 1);
        out[1] = terrain_position(t, c + 1, r + 1);
        out[2] = terrain_position(t, c + 1, r);
    } else {
        out[0] = terrain_position(t, c, r);
        out[1] = terrain_position(t, c, r + 1);
        out[2] = terrain_position(t, c + 1, r + 1);
    }
}

}  // namespace ref
