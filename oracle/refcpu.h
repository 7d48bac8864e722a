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
bool mpr_contact(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& axis,
                 ContactData* c);
// va / vb: the entities' linear velocities (used on entering deep contact from separation).
bool convex_candidate(const ConvexShapeRef& a, const ConvexShapeRef& b, const RT& ta, const RT& tb, const V3& va,
                      const V3& vb, ConvexTesterState* st, ContactData* c);
// Casts of continuous collision detection (GJKToolbox.cs:195-495); *t is the
// hit parameter in units of `dir`.
bool sphere_cast(V3 origin, V3 dir, float radius, const ConvexShapeRef& shape, const RT& xf, float max_length, float* t);
bool ray_cast(V3 origin, V3 dir, const ConvexShapeRef& shape, const RT& xf, float max_length, float* t);
bool ccd_sphere_cast(const V3& origin, const V3& dir, float radius, const ConvexShapeRef& shape, const RT& xf,
                     float max_length, float scaling, float* t);

}  // namespace ref
