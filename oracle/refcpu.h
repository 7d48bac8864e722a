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

}  // namespace ref
