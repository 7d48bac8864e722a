// refcpu.h -- internal declarations of the oracle (CPU restatement of the
// BEPUphysics v1 per-step rigid-body path).
//
// TEST INFRASTRUCTURE ONLY: the oracle is the checker the tests, the smoke
// entry and the benchmark's CPU-baseline leg compare the GPU library against.
// Nothing in the product library includes or links it.
#pragma once
#include <cstdint>
#include <vector>
#include "refmath.h"

namespace ref {

// BoxContactData (CollisionTests/CollisionAlgorithms/BoxBoxCollider.cs:10-27)
struct BoxContact {
    V3 pos;
    float depth;
    int id;
};
// TinyStructList<BoxContactData>: 8 inline entries; Add beyond 8 is ignored,
// RemoveAt shifts (BEPUutilities/DataStructures/TinyStructList.cs:50-83,190-240).
struct BoxContactList {
    BoxContact e[8];
    int n = 0;
    void add(const BoxContact& c) {
        if (n < 8) e[n++] = c;
    }
    void remove_at(int i) {
        for (int k = i; k < n - 1; ++k) e[k] = e[k + 1];
        --n;
    }
};

// ContactData (CollisionTests/ContactData.cs)
struct ContactData {
    V3 pos;
    V3 normal;
    float depth;
    int id;
};

// BoxBoxCollider.AreBoxesColliding, non-ALLOWUNSAFE overload
// (BoxBoxCollider.cs:1494-2386).  ha/hb are half extents (halfWidth, halfHeight, halfLength).
bool box_box_colliding(const V3& ha, const V3& hb, const V3& posA, const Q4& qA, const V3& posB,
                       const Q4& qB, float* distance, V3* axis, BoxContactList* contacts);

// SphereTester.AreSpheresColliding (CollisionAlgorithms/SphereTester.cs:23-56)
bool sphere_sphere_colliding(float radiusA, float radiusB, const V3& posA, const V3& posB,
                             float max_contact_distance, ContactData* contact);

// BoxSphereTester.AreShapesColliding (CollisionAlgorithms/BoxSphereTester.cs:23-85)
bool box_sphere_colliding(const V3& half, const V3& boxPos, const Q4& boxQ, float radius,
                          const V3& spherePos, float max_contact_distance, ContactData* contact);

// ---------------------------------------------------------------------------
// General convex path (refcpu_convex.cpp)

// RigidTransform (BEPUutilities/RigidTransform.cs)
struct RT {
    V3 pos;
    Q4 q;
};

// What the support mapping of a ConvexShape needs: BEPU_SHAPE_* kind, box
// half extents (margin included, as BoxShape stores them), capsule half
// length, hull vertices (recentred surface points) and the collision margin.
struct ConvexShapeRef {
    int kind;
    V3 half;
    float half_length;
    const V3* verts;
    int nverts;
    float margin;
};

// GeneralConvexPairTester state kept across steps (GeneralConvexPairTester.cs:15-50):
// the state machine, the boolean GJK's warm-start axis and the deep-contact
// search direction.  The cached simplex is always the initial one (a Point
// simplex at the shapes' origins) because simplex caching is off by default.
enum { CT_SEPARATED = 0, CT_SHALLOW = 1, CT_DEEP = 2 };
struct ConvexTesterState {
    int state;
    V3 separating_axis;
    V3 local_direction;
};

V3 support_local(const ConvexShapeRef& s, const V3& d);
Q4 qconcat(const Q4& a, const Q4& b);
V3 rt_transform(const V3& p, const RT& t);
V3 rt_transform_by_inverse(const V3& p, const RT& t);
RT local_transform(const RT& a, const RT& b);
bool gjk_intersecting(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, V3* axis);
bool gjk_closest_points(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, V3* ca, V3* cb);
// The same from the cached Point simplex whose B-side local point is `simplex_b`.
bool gjk_closest_points_from(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb,
                             const V3& simplex_b, V3* ca, V3* cb);
bool mpr_contact(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& axis,
                 ContactData* c);
// va / vb: the entities' linear velocities (used on entering deep contact from separation).
bool convex_candidate(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& va,
                      const V3& vb, ConvexTesterState* st, ContactData* c);
// Casts of continuous collision detection (GJKToolbox.cs:195-495); *t is the
// hit parameter in units of `dir`.
// `normal` (optional) receives RayHit.Normal: the search direction of the
// last advance, in world space (not normalised).
bool sphere_cast(V3 origin, V3 dir, float radius, const ConvexShapeRef& shape, const RT& xf, float max_length, float* t,
                 V3* normal = nullptr);
bool ray_cast(V3 origin, V3 dir, const ConvexShapeRef& shape, const RT& xf, float max_length, float* t,
              V3* normal = nullptr);
bool ccd_sphere_cast(const V3& origin, const V3& dir, float radius, const ConvexShapeRef& shape, const RT& xf,
                     float max_length, float scaling, float* t);
bool ccd_sphere_cast_normal(const V3& origin, const V3& dir, float radius, const ConvexShapeRef& shape, const RT& xf,
                            float max_length, float scaling, float* t, V3* normal);

// ---------------------------------------------------------------------------
// Triangles of static meshes (refcpu_convex.cpp, refcpu_mesh.cpp)

// A TriangleShape (CollisionShapes/ConvexShapes/TriangleShape.cs) as a
// ConvexShapeRef: kind kShapeTriangle, verts -> its three vertices.
constexpr int kShapeTriangle = 16;
// TriangleSidedness (BEPUutilities/TriangleSidedness.cs)
enum { SIDED_DOUBLE = 0, SIDED_CLOCKWISE = 1, SIDED_COUNTERCLOCKWISE = 2 };
// VoronoiRegion (BEPUutilities/VoronoiRegion.cs)
enum { VR_A = 0, VR_B, VR_C, VR_AB, VR_AC, VR_BC, VR_ABC };
// A mesh triangle in the local space of the convex it is tested against.
struct TriangleRef {
    V3 v[3];
    int sidedness;
    float margin;
};
// TriangleConvexPairTester's state kept across steps (TriangleConvexPairTester.cs:14-20):
// the state machine, the escape counter and the boolean GJK's warm-start axis.
enum { TT_PLANE = 0, TT_SEPARATED = 1, TT_NEAR = 2, TT_DEEP = 3 };
struct TriTesterState {
    int state;
    int escapes;
    V3 axis;
};
ConvexShapeRef triangle_shape(const TriangleRef& t);
// ConvexShape.GetLocalExtremePoint (ConvexShape.cs:104-113): with the margin.
V3 local_extreme_point(const ConvexShapeRef& s, const V3& d);
// Toolbox.GetClosestPointOnTriangleToPoint (BEPUutilities/Toolbox.cs:338-419).
int closest_point_on_triangle(const V3& a, const V3& b, const V3& c, const V3& p, V3* out);
// TriangleConvexPairTester.GenerateContactCandidates (TriangleConvexPairTester.cs:23-46):
// up to two candidates in the convex's local space; `min_radius` is the
// convex shape's MinimumRadius, `core_scaling` MotionSettings.CoreShapeScaling.
int triangle_convex_candidates(const ConvexShapeRef& convex, float min_radius, float core_scaling,
                               const TriangleRef& t, TriTesterState* st, ContactData out[2]);
// TriangleConvexPairTester.GetRegion (TriangleConvexPairTester.cs:708-897).
int triangle_convex_region(const TriangleRef& t, const ContactData& c);
// TriangleSpherePairTester.GenerateContactCandidates (TriangleSpherePairTester.cs:20-94):
// at most one candidate; *region is the tester's lastRegion.
int triangle_sphere_candidates(float radius, const TriangleRef& t, ContactData* out, int* region);

// AffineTransform (BEPUutilities/AffineTransform.cs): linear part and translation.
struct Affine {
    M3 m;
    V3 t;
};
// A static mesh shape (StaticMeshShape.cs, TransformableMeshData.cs): local
// vertices, index triples, the world transform and its MeshBoundingBoxTree
// (DataStructures/MeshBoundingBoxTree.cs) flattened: node k has a box and,
// for an internal node, children a and b; a leaf has a = -1 and b = its
// triangle (the position of its first index).
struct MeshNode {
    V3 bmin, bmax;
    int a, b;
};
struct MeshShape {
    std::vector<V3> verts;
    std::vector<int> idx;
    Affine xf;
    std::vector<MeshNode> nodes;  // nodes[0] is the root (empty mesh: no nodes)
};
// A terrain (CollisionShapes/TerrainShape.cs, BroadPhaseEntries/Terrain.cs):
// heights[column * length + row] for `width` columns (x) and `length` rows
// (z), the quads' triangle organization (QuadTriangleOrganization:
// 0 BottomLeftUpperRight, 1 BottomRightUpperLeft) and the world transform.
struct TerrainShapeRef {
    int width, length, organization;
    std::vector<float> heights;
    Affine xf;
    float h(int column, int row) const { return heights[(size_t)column * length + row]; }
};
// Matrix3x3.Invert (BEPUutilities/Matrix3x3.cs:505-531), AffineTransform.Invert (AffineTransform.cs:173-178)
M3 m3_invert(const M3& m);
Affine affine_invert(const Affine& a);
// TerrainShape.GetBoundingBox (TerrainShape.cs:86-140)
void terrain_bounding_box(const TerrainShapeRef& t, V3* mn, V3* mx);
// TerrainShape.GetOverlaps (TerrainShape.cs:497-543): triangle ids (2 per quad) over a local box.
void terrain_overlaps(const TerrainShapeRef& t, const V3& lmin, const V3& lmax, std::vector<int>* out);
// TerrainShape.GetLocalIndices (TerrainShape.cs:626-663): (column, row) of the three vertices.
void terrain_local_indices(const TerrainShapeRef& t, int tri, int col[3], int row[3]);
// TerrainShape.GetTriangle (TerrainShape.cs:607-622): the world-space triangle.
void terrain_world_triangle(const TerrainShapeRef& t, int tri, V3 out[3]);
// Terrain.WorldTransform setter (Terrain.cs:41-75): the sidedness its winding gets.
int terrain_sidedness(const Affine& xf);

// MeshBoundingBoxTree.Reconstruct (MeshBoundingBoxTree.cs:60-69).
void mesh_build_tree(MeshShape* m);
// MeshBoundingBoxTree.GetOverlaps(BoundingBox) (MeshBoundingBoxTree.cs:112-122):
// triangles (first-index positions) in the tree's depth-first order.
void mesh_overlaps(const MeshShape& m, const V3& bmin, const V3& bmax, std::vector<int>* out);
// TransformableMeshData.GetTriangle (TransformableMeshData.cs:67-72): the world-space triangle.
void mesh_world_triangle(const MeshShape& m, int tri, V3 out[3]);
V3 affine_transform(const V3& p, const Affine& a);
// Toolbox.GetTriangleBoundingBox (BEPUutilities/Toolbox.cs:1516-1580)
void triangle_bounding_box(const V3& a, const V3& b, const V3& c, V3* mn, V3* mx);
Affine affine_multiply(const Affine& a, const Affine& b);

}  // namespace ref
