// refcpu_world.cpp -- oracle restatement of one BEPUphysics v1 time step
// (Space.DoTimeStep, BEPUphysics/Space.cs:496-521) for convex bodies, run
// single-threaded, plus its C API for the tests (ctypes).
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
// against; it is never linked into the product.
//
// Deliberate, documented deviations from the reference (all parity tests
// state them):
//  * Solve order.  The reference visits solver updateables in the permuted
//    order of PermutationMapper over the Solver.Add history (Solver.cs:239-263).
//    Both the oracle and the GPU library use the deterministic colour order
//    defined in colour_constraints() below; impulses are compared once the
//    reference's constraint order is permuted to that colour order.
//  * Pair orientation.  The reference takes A/B from DynamicHierarchy traversal
//    order; here A is the lower body id, except box-sphere where A is the box
//    as BoxSpherePairHandler.Initialize enforces (BoxSpherePairHandler.cs:78-97).
//  * Deactivation.  Sleeping is out of scope: every dynamic body behaves as
//    IsAlwaysActive; the per-body stabilization state is kept
//    (SimulationIslandMember.cs:53-70, EntityBase.cs:927-936).
#include <algorithm>
#include <cfloat>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/bepu_gpu.h"
#include "refcpu.h"

namespace ref {

struct Body {
    int shape;       // BEPU_SHAPE_*
    V3 half;         // box half extents
    float radius;    // sphere / capsule radius (= collision margin)
    float half_length;  // capsule
    float margin;    // collision margin (box and hull: the default margin)
    int hull_off, hull_n;  // hull vertices in refcpu_world::hull_verts
    bool dynamic;
    float mass, inv_mass;
    M3 local_inertia_inv;
    M3 inertia_inv;  // world space, refreshed by the force update
    V3 pos;
    Q4 q;
    M3 R;            // orientationMatrix
    V3 v, w;
    float lin_damp, ang_damp, lin_boost, ang_boost;
    float prev_velocity;
    bool is_slowing;
    float time_below;
    bool allow_stabilization;
    bool always_active;  // SimulationIslandMember.IsAlwaysActive
    bool candidate;      // SimulationIslandMember.IsDeactivationCandidate
    // Activity (DeactivationManagement/SimulationIsland.cs, SimulationIslandMember.cs).
    // A dynamic body sleeps with its island; a kinematic body's activity is
    // its time below the limit used as a flag (SimulationIslandMember.cs:
    // 227-250): <= 0 active, 1 inactive.
    bool asleep;         // its island is inactive
    bool wake;           // Activate() since the last step: the island is active at the next one
    bool nodeact;        // its island may not deactivate at its next check (SimulationIsland.allowDeactivation)
    uint32_t label;      // its island at the end of the last step (kNoLabel: its own)
    bepu_material mat;
    V3 bmin, bmax;
    // Space.Remove(entity) (Space.cs:364-482): the body left every stage; its
    // broadphase entry is gone, so none of its pairs survive.
    bool removed;
    // PositionUpdateMode (PositionUpdating/ICCDPositionUpdateable.cs:6-21):
    // 0 Discrete, 1 Passive, 2 Continuous; the shape's MinimumRadius; and
    // the orientation of the collidable's world transform while the linear
    // motion waits for the times of impact (PreUpdatePosition turns the
    // entity but refreshes that transform only for discrete bodies).
    int pmode;
    float min_radius;
    Q4 q_world;
    // A static mesh (StaticMesh.cs): its shape (refcpu_world::mesh_shapes),
    // TriangleSidedness and ImproveBoundaryBehavior; -1 for an entity.
    int mesh;
    int sidedness;
    bool improve_boundary;
    // A terrain (Terrain.cs): its entry in refcpu_world::terrains; -1 otherwise.
    int terrain;
};

// A static collidable (a static mesh or a terrain): not an entity.
inline bool is_static(const Body& b) { return b.mesh >= 0 || b.terrain >= 0; }

struct ContactObj {
    V3 pos, normal;
    float depth;
    int id;
};

struct PenC {  // ContactPenetrationConstraint
    int slot;
    float acc;
    float lin[3], angA[3], angB[3];
    float softness, bias, v2i;
    V3 ra, rb;
    bool aDyn, bDyn;
    bool ais;  // isActiveInSolver
    int cur, zero;
};
struct FricC {  // ContactFrictionConstraint
    float acc;
    float lin[3], angA[3], angB[3];
    float friction, v2i;
    bool aDyn, bDyn, ais;
    int cur, zero;
};
struct Sliding {  // SlidingFrictionTwoAxis
    bool active, ais;
    int cur, zero;
    float acc[2];
    float lin[6], angA[6], angB[6];  // Matrix2x3 M11 M12 M13 M21 M22 M23
    float v2i[4];                    // Matrix2x2 M11 M12 M21 M22
    float friction;
    int count;
    V3 center, relvel, ra, rb;
    bool aDyn, bDyn;
};
struct Twist {  // TwistFrictionConstraint
    bool active, ais;
    int cur, zero;
    float acc;
    float ang[3];
    float v2i, friction;
    float lever[4];
    int count;
    bool aDyn, bDyn;
};

struct Pair {
    uint32_t a, b;
    int kind;
    float kinetic, static_f, bounce;
    ContactObj slot[4];
    bool used[4];
    int clist[4];
    int ccount;
    PenC pen[4];
    int npen;
    FricC fric[4];
    int nfric;
    Sliding sl;
    Twist tw;
    bool prev_colliding;
    // GeneralConvexContactManifold: the pair tester and the per-contact
    // supplement data (local offsets, base depth) indexed by slot.
    ConvexTesterState gc;
    V3 off_a[4], off_b[4];
    float base_depth[4];
    bool group_ais;
    int group_cur, group_zero;
    int color;
    // Not updated this step: both bodies were inactive when the broad phase
    // ran (BroadPhase.cs:97-102; NarrowPhase.cs:300-333 keeps such pairs).
    bool frozen;
    // CollidablePairHandler.timeOfImpact (CollidablePairHandler.cs:59): 1 for
    // a new pair, recomputed by ConvexPairHandler.UpdateTimeOfImpact in the
    // steps it has a continuous body and an active one, kept otherwise.
    float toi;
    // Mesh pairs with a convex (not a sphere): the manifold's table of
    // TriangleConvexPairTesters, refcpu_world::mesh_testers[mesh_slot].
    int mesh_slot;
    // Terrain pairs: the local box of this step's triangle query (the time of
    // impact sweeps the triangles the manifold found).
    V3 tq_min, tq_max;  // a terrain pair's last triangle query box (TerrainContactManifold)
    bool tq_valid;
};

// A Terrain (BroadPhaseEntries/Terrain.cs): its shape's heights, its world
// transform, the sidedness that transform gives its triangles.
struct TerrainInst {
    TerrainShapeRef shape;
    int sidedness;
};

// One entry of TriangleMeshConvexContactManifold.activePairTesters
// (TriangleMeshConvexContactManifold.cs:19-20): the triangle's vertex
// indices (TriangleIndices) and its tester.
struct MeshTester {
    int ia, ib, ic;
    TriTesterState st;
    bool updated;
};

// BallSocketJoint (Constraints/TwoEntity/Joints/BallSocketJoint.cs) and its
// SolverUpdateable bookkeeping (SolverSettings: currentIterations,
// iterationsAtZeroImpulse; isActiveInSolver).
struct JointC {
    uint32_t a, b;       // ConnectionA / ConnectionB
    uint32_t local;      // its ordinal among the joints of its world (the colouring key)
    bool dead;           // a connection was removed: the joint no longer acts
    V3 localA, localB;   // localAnchorA / localAnchorB
    float stiffness, damping;        // SpringSettings (600000, 90000)
    float max_corr, max_corr_sq;     // Joint.MaxCorrectiveVelocity (float.MaxValue)
    V3 acc;              // accumulatedImpulse
    // prestep
    V3 offA, offB, bias;
    M3 mass;
    float softness;
    bool ais;
    int cur, zero;
    int color;
};

// The joints' colouring keys lie above every pair key.
constexpr uint64_t kJointKey = 1ull << 63;

struct Settings {
    bepu_settings s;
};

}  // namespace ref

using namespace ref;

namespace ref {

// A persistent pool for the multithreaded step: run(n, fn) calls fn(begin, end)
// over [0, n) in chunks on every worker and the caller, and returns when all
// chunks are done.
class Pool {
public:
    explicit Pool(int nthreads) {
        for (int i = 1; i < nthreads; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    int size() const { return (int)workers_.size() + 1; }
    void run(size_t n, size_t chunk, const std::function<void(size_t, size_t)>& fn) {
        if (n == 0) return;
        if (workers_.empty() || n <= chunk) {
            fn(0, n);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            fn_ = &fn;
            n_ = n;
            chunk_ = chunk;
            next_.store(0);
            pending_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> l(m_);
        done_cv_.wait(l, [this] { return pending_ == 0; });
        fn_ = nullptr;
    }

private:
    void work() {
        for (;;) {
            size_t b = next_.fetch_add(chunk_);
            if (b >= n_) break;
            (*fn_)(b, std::min(n_, b + chunk_));
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
            }
            work();
            std::lock_guard<std::mutex> g(m_);
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    uint64_t gen_ = 0;
    bool stop_ = false;
    int pending_ = 0;
    const std::function<void(size_t, size_t)>* fn_ = nullptr;
    size_t n_ = 0, chunk_ = 1;
    std::atomic<size_t> next_{0};
};

// The pair states of the last step in key order.
struct PairTable {
    std::vector<uint64_t> keys;
    std::vector<Pair> vals;
    const Pair* find(uint64_t k) const {
        auto it = std::lower_bound(keys.begin(), keys.end(), k);
        return (it != keys.end() && *it == k) ? &vals[it - keys.begin()] : nullptr;
    }
};

}  // namespace ref

struct refcpu_world {
    bepu_settings s;
    std::vector<Body> bodies;
    std::vector<V3> hull_verts;
    PairTable pairs;
    std::vector<uint64_t> overlaps;
    std::vector<uint64_t> order;  // solve order of the last step (pair keys)
    int ncolors = 0;
    uint32_t steps = 0;
    float max_pen = 0;
    int threads = 1;
    std::unique_ptr<Pool> pool;  // threads > 1
    int broadphase_mode = 0;  // 0: uniform grid, 1: brute force, 2: sort and sweep
    uint32_t deact_cursor = 0xffffffffu;  // last island looked at by the deactivation rotation (none yet)
    // Solve order: 0 colour by colour (the GPU's); 1 the reference's
    // single-threaded order (prestep and warm start in list order, then every
    // iteration through one permutation of the list per step); 2 the same
    // with the list in descending key order (another order of addition).
    int solve_order = 0;
    int64_t perm_index = 0;
    bool ccd = false;  // a body was ever passive or continuous
    std::vector<JointC> joints;
    std::vector<MeshShape> mesh_shapes;
    std::vector<TerrainShapeRef> terrain_shapes;  // heights and organization (transform per terrain)
    std::vector<TerrainInst> terrains;
    // Tester tables of the mesh pairs (Pair::mesh_slot), allocated as pairs
    // are created (under the lock: pairs are updated on the pool) and freed
    // when they end.
    std::deque<std::vector<MeshTester>> mesh_testers;
    std::vector<int> mesh_free;
    std::mutex mesh_mu;
    // Largest per-pair list sizes any mesh manifold update met (tests): contact
    // candidates, edge contacts, vertex contacts, blocked vertices, blocked
    // edges, pair testers.
    int mesh_peak[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

// Runs fn(begin, end) over [0, n): on the pool when the world is multithreaded.
void par_for(refcpu_world* w, size_t n, size_t chunk, const std::function<void(size_t, size_t)>& fn) {
    if (w->pool) w->pool->run(n, chunk, fn);
    else fn(0, n);
}

// ---------------------------------------------------------------------------
// Entity construction (host side, once per body)

void init_body(refcpu_world* w, Body* b, const bepu_motion_state* st, float mass, const bepu_material* m,
               uint32_t flags, const M3& volume_distribution) {
    // Entity.Initialize(collidable, mass) -> BecomeDynamic(mass, VolumeDistribution * (mass * 2.5))
    // (EntityBase.cs:632-645, 867-905; InertiaHelper.InertiaTensorScale = 2.5)
    if (mass > 0) {
        b->dynamic = true;
        M3 inertia = m3_scale(volume_distribution, mass * 2.5f);
        b->local_inertia_inv = m3_adaptive_invert(inertia);
        b->mass = mass;
        b->inv_mass = 1 / mass;
    } else {
        // BecomeKinematic: infinite mass, zero inverse inertia.
        b->dynamic = false;
        b->local_inertia_inv = M3{0, 0, 0, 0, 0, 0, 0, 0, 0};
        b->mass = 0;
        b->inv_mass = 0;
    }
    b->pos = V3{st->position[0], st->position[1], st->position[2]};
    // Orientation setter normalises (EntityBase.cs:68-84)
    b->q = qnormalize(Q4{st->orientation[0], st->orientation[1], st->orientation[2], st->orientation[3]});
    b->R = m3_from_quat(b->q);
    b->inertia_inv = m3_mul(m3_mul_transposed(b->R, b->local_inertia_inv), b->R);
    b->v = V3{st->linear_velocity[0], st->linear_velocity[1], st->linear_velocity[2]};
    b->w = V3{st->angular_velocity[0], st->angular_velocity[1], st->angular_velocity[2]};
    b->lin_damp = w->s.linear_damping;
    b->ang_damp = w->s.angular_damping;
    b->lin_boost = b->ang_boost = 0;
    b->prev_velocity = 0;
    b->is_slowing = false;
    b->time_below = 0;
    b->allow_stabilization = (flags & BEPU_FLAG_NO_STABILIZATION) == 0;
    b->always_active = (flags & BEPU_FLAG_ALWAYS_ACTIVE) != 0;
    b->candidate = false;
    // DeactivationManager.Add activates the new member (DeactivationManager.cs:164):
    // a kinematic body's flag becomes -1 (active until two quiet steps pass).
    if (!b->dynamic) b->time_below = -1;
    b->asleep = b->wake = b->nodeact = false;
    b->label = 0xffffffffu;
    b->mat = m ? *m : bepu_material{1.0f, 0.8f, 0.0f};
    b->pmode = (flags & BEPU_FLAG_CCD_CONTINUOUS) ? 2 : ((flags & BEPU_FLAG_CCD_PASSIVE) ? 1 : 0);
    if (b->pmode) w->ccd = true;
    b->q_world = b->q;
    b->mesh = -1;
    b->terrain = -1;
    b->sidedness = SIDED_DOUBLE;
    b->improve_boundary = false;
}

// ---------------------------------------------------------------------------
// Body stages

// SimulationIslandMember.UpdateDeactivationCandidacy, dynamic and active branch
// (SimulationIslandMember.cs:53-95).  Clearing the candidate flag goes
// through the IsDeactivationCandidate setter, which also zeroes the time
// below the limit (SimulationIslandMember.cs:171-190): an always-active body
// clears it every step, so its time never accumulates.
void candidacy(refcpu_world* w, Body& b) {
    if (!b.dynamic || b.removed) return;
    float velocity = length_sq(b.v) + length_sq(b.w);
    b.is_slowing = velocity <= b.prev_velocity;
    float lim = w->s.velocity_lower_limit;
    if (velocity < lim * lim) b.time_below += w->s.dt;
    else b.time_below = 0;
    if (!b.always_active) {
        if (!b.candidate) {
            if (b.time_below > w->s.low_velocity_time_minimum && b.is_slowing) b.candidate = true;
        } else if (b.time_below <= w->s.low_velocity_time_minimum) {
            b.candidate = false;
            b.time_below = 0;
        }
    } else {
        b.candidate = false;
        b.time_below = 0;
    }
    b.prev_velocity = velocity;
}

constexpr uint32_t kNoLabel = 0xffffffffu;

uint32_t label_of(const refcpu_world* w, uint32_t i) {
    const uint32_t l = w->bodies[i].label;
    return l == kNoLabel ? i : l;
}

// SimulationIslandMember.UpdateDeactivationCandidacy for a kinematic member
// (SimulationIslandMember.cs:96-139): an inactive one (flag 1) stays so until
// activated; an active one becomes a candidate at zero velocity (flag -1 ->
// 0 -> 1: inactive after a full quiet step) and is active again (-1) when it
// moves.  Returns whether it is active after the update.
bool kinematic_candidacy(refcpu_world* w, Body& b) {
    (void)w;
    if (b.time_below > 0) return false;
    float velocity = length_sq(b.v) + length_sq(b.w);
    if (velocity == 0 && !b.always_active) {
        if (b.time_below == 0) b.time_below = 1;
        else if (b.time_below < 0) b.time_below = 0;
    } else {
        b.time_below = -1;
    }
    return b.time_below <= 0;
}

inline bool kinematic_active(const Body& b) { return !b.dynamic && b.time_below <= 0; }

// Whether a body takes part in this step's stages (IsActive): a dynamic body
// when its island is active, a kinematic one by its flag.
inline bool body_active(const Body& b) { return !b.removed && (b.dynamic ? !b.asleep : b.time_below <= 0); }

// SimulationIslandMember.Activate (SimulationIslandMember.cs:256-277): the
// candidate flag is cleared (which zeroes the time below the limit) and the
// island becomes active; a kinematic member's flag becomes -1.
void activate(Body& b) {
    if (b.removed || is_static(b)) return;  // a static collidable has no activity
    if (b.dynamic) {
        b.candidate = false;
        b.time_below = 0;
        b.wake = true;
    } else {
        b.time_below = -1;
    }
}

struct UnionFind {
    std::vector<uint32_t> p;
    explicit UnionFind(size_t n) : p(n) {
        for (size_t i = 0; i < n; ++i) p[i] = (uint32_t)i;
    }
    uint32_t find(uint32_t x) {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    void unite(uint32_t a, uint32_t b) {
        a = find(a);
        b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

// DeactivationManager.Update (DeactivationManager.cs:196-205) at the start of
// the step, over the islands the last step ended with (the connected
// components of its constraints; split attempts are flushed at once):
//   * kinematic members update their flag; an active one keeps the islands
//     it touches through constraints active and forbids their next
//     deactivation (SimulationIslandMember.cs:105-139);
//   * an island with a member activated since the last step is active;
//   * members of active islands update their candidacy (the dynamic branch,
//     SimulationIslandMember.cs:53-95), the others only remember their speed;
//   * DeactivateObjects (DeactivationManager.cs:291-320): islands are looked
//     at in turn, continuing after the last one looked at in the previous
//     step, until the members looked at reach MaximumDeactivationAttempts-
//     PerFrame or every island was looked at once; an active island whose
//     members are all candidates goes to sleep (SimulationIsland.TryToDeactivate,
//     SimulationIsland.cs:92-114).  The turn order is by smallest member id
//     (the reference's follows its island list).  The rotation runs in a step
//     only if the world holds a dynamic body that is not always active.
void deactivation_stage(refcpu_world* w) {
    const uint32_t n = (uint32_t)w->bodies.size();
    for (Body& b : w->bodies)
        if (!b.dynamic && !b.removed) kinematic_candidacy(w, b);
    std::vector<uint8_t> act(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const Body& b = w->bodies[i];
        if (b.dynamic && !b.removed && b.wake) act[label_of(w, i)] = 1;
    }
    // A moving kinematic member keeps its connected islands awake: its
    // connections are its contact constraints and its joints.
    auto kinematic_rule = [&](uint32_t a, uint32_t b) {
        Body& A = w->bodies[a];
        Body& B = w->bodies[b];
        if (A.removed || B.removed) return;
        if (kinematic_active(A) && B.dynamic) {
            act[label_of(w, b)] = 1;
            B.nodeact = true;
        }
        if (kinematic_active(B) && A.dynamic) {
            act[label_of(w, a)] = 1;
            A.nodeact = true;
        }
    };
    for (const Pair& p : w->pairs.vals)
        if (p.ccount > 0) kinematic_rule(p.a, p.b);
    for (const JointC& j : w->joints)
        if (!j.dead) kinematic_rule(j.a, j.b);
    struct Agg {
        uint32_t members = 0, cands = 0, minid = kNoLabel;
        bool active = false, nodeact = false;
    };
    std::vector<Agg> agg(n);
    bool sleepy = false;
    for (uint32_t i = 0; i < n; ++i) {
        Body& b = w->bodies[i];
        if (!b.dynamic || b.removed) continue;
        const uint32_t L = label_of(w, i);
        const bool active = !b.asleep || act[L];
        if (active) {
            candidacy(w, b);
            b.asleep = false;
        } else {
            b.prev_velocity = length_sq(b.v) + length_sq(b.w);
        }
        b.wake = false;
        Agg& g = agg[L];
        g.members++;
        g.cands += b.candidate ? 1u : 0u;
        g.active = g.active || active;
        g.nodeact = g.nodeact || b.nodeact;
        g.minid = std::min(g.minid, i);
        if (!b.always_active) sleepy = true;
    }
    if (!sleepy) return;
    std::vector<uint32_t> reps;
    for (uint32_t i = 0; i < n; ++i) {
        const Body& b = w->bodies[i];
        if (b.dynamic && !b.removed && agg[label_of(w, i)].minid == i) reps.push_back(i);
    }
    if (reps.empty()) return;
    size_t start = 0;
    if (w->deact_cursor != kNoLabel)
        start = (size_t)(std::upper_bound(reps.begin(), reps.end(), w->deact_cursor) - reps.begin());
    std::vector<uint8_t> checked(n, 0), sleep(n, 0);
    int64_t budget = 0;
    for (size_t k = 0; k < reps.size(); ++k) {
        if (budget >= (int64_t)w->s.max_deactivation_attempts) break;
        const uint32_t r = reps[(start + k) % reps.size()];
        const uint32_t L = label_of(w, r);
        const Agg& g = agg[L];
        checked[L] = 1;
        if (!g.nodeact && g.active && g.cands == g.members) sleep[L] = 1;
        budget += g.members;
        w->deact_cursor = r;
    }
    for (uint32_t i = 0; i < n; ++i) {
        Body& b = w->bodies[i];
        if (!b.dynamic || b.removed) continue;
        const uint32_t L = label_of(w, i);
        if (checked[L]) b.nodeact = false;
        if (sleep[L]) b.asleep = true;
    }
}

// Entity.UpdateForForces (EntityBase.cs:911-977); ForceUpdater.gravityDt (ForceUpdater.cs:85-92)
void force_update(refcpu_world* w, Body& b) {
    if (!b.dynamic || b.removed || b.asleep) return;  // ForceUpdater.cs:75: active entities only
    const bepu_settings& s = w->s;
    V3 gravityDt = mul(V3{s.gravity[0], s.gravity[1], s.gravity[2]}, s.dt);
    b.v = add(gravityDt, b.v);
    float lim = s.velocity_lower_limit;
    if (s.use_stabilization && b.allow_stabilization &&
        (b.is_slowing || b.time_below > s.low_velocity_time_minimum)) {
        float energy = length_sq(b.v) + length_sq(b.w);
        if (energy < lim * lim) {
            float boost = 1 - (float)(std::sqrt((double)energy) / (double)(2.0f * lim));
            // ModifyAngularDamping / ModifyLinearDamping (EntityBase.cs:1205-1221)
            float total = b.ang_damp + b.ang_boost;
            float rem = 1 - total;
            b.ang_boost += boost * rem;
            total = b.lin_damp + b.lin_boost;
            rem = 1 - total;
            b.lin_boost += boost * rem;
        }
    }
    float linear = b.lin_damp + b.lin_boost;
    if (linear > 0) b.v = mul(b.v, (float)std::pow((double)clampf(1 - linear, 0, 1), (double)s.dt));
    float angular = b.ang_damp + b.ang_boost;
    if (angular > 0) b.w = mul(b.w, (float)std::pow((double)clampf(1 - angular, 0, 1), (double)s.dt));
    b.lin_boost = 0;
    b.ang_boost = 0;
    b.inertia_inv = m3_mul(m3_mul_transposed(b.R, b.local_inertia_inv), b.R);
}

// Shape.GetBoundingBox at a pose (BoxShape.cs:149-174; SphereShape.cs:75-86;
// CapsuleShape.cs:95-137; ConvexHullShape.cs:206-286).
void shape_bounding_box(const refcpu_world* w, const Body& b, const V3& bpos, const Q4& bq, V3* bmin, V3* bmax) {
    if (b.shape == BEPU_SHAPE_BOX) {
        M3 o = m3_from_quat(bq);
        V3 right{signf_int(o.m11) * b.half.x, signf_int(o.m21) * b.half.y, signf_int(o.m31) * b.half.z};
        V3 up{signf_int(o.m12) * b.half.x, signf_int(o.m22) * b.half.y, signf_int(o.m32) * b.half.z};
        V3 back{signf_int(o.m13) * b.half.x, signf_int(o.m23) * b.half.y, signf_int(o.m33) * b.half.z};
        V3 off;
        off.x = right.x * o.m11 + right.y * o.m21 + right.z * o.m31;
        off.y = up.x * o.m12 + up.y * o.m22 + up.z * o.m32;
        off.z = back.x * o.m13 + back.y * o.m23 + back.z * o.m33;
        *bmax = add(bpos, off);
        *bmin = sub(bpos, off);
    } else if (b.shape == BEPU_SHAPE_CAPSULE) {
        // Quaternion.TransformY (Quaternion.cs:586-608)
        const Q4& r = bq;
        float x2 = r.x + r.x, y2 = r.y + r.y, z2 = r.z + r.z;
        float xx2 = r.x * x2, xy2 = r.x * y2, yz2 = r.y * z2, zz2 = r.z * z2;
        float wx2 = r.w * x2, wz2 = r.w * z2;
        float hl = b.half_length, m = b.radius;
        V3 up{hl * (xy2 - wz2), hl * (1.f - xx2 - zz2), hl * (yz2 + wx2)};
        V3 mn, mx;
        if (up.x > 0) { mx.x = up.x + m; mn.x = -up.x - m; } else { mx.x = -up.x + m; mn.x = up.x - m; }
        if (up.y > 0) { mx.y = up.y + m; mn.y = -up.y - m; } else { mx.y = -up.y + m; mn.y = up.y - m; }
        if (up.z > 0) { mx.z = up.z + m; mn.z = -up.z - m; } else { mx.z = -up.z + m; mn.z = up.z - m; }
        *bmin = add(bpos, mn);
        *bmax = add(bpos, mx);
    } else if (b.shape == BEPU_SHAPE_CONVEX_HULL) {
        M3 o = m3_from_quat(bq);
        const V3* v = &w->hull_verts[b.hull_off];
        V3 right{o.m11, o.m21, o.m31}, up{o.m12, o.m22, o.m32}, back{o.m13, o.m23, o.m33};
        float maxX = dot(v[0], right), minX = maxX;
        float maxY = dot(v[0], up), minY = maxY;
        float maxZ = dot(v[0], back), minZ = maxZ;
        int iminX = 0, imaxX = 0, iminY = 0, imaxY = 0, iminZ = 0, imaxZ = 0;
        for (int i = 1; i < b.hull_n; ++i) {
            float d = dot(v[i], right);
            if (d < minX) { minX = d; iminX = i; } else if (d > maxX) { maxX = d; imaxX = i; }
            d = dot(v[i], up);
            if (d < minY) { minY = d; iminY = i; } else if (d > maxY) { maxY = d; imaxY = i; }
            d = dot(v[i], back);
            if (d < minZ) { minZ = d; iminZ = i; } else if (d > maxZ) { maxZ = d; imaxZ = i; }
        }
        // EntityShape.TransformLocalExtremePoints (EntityShape.cs:58-63)
        V3 pos{v[imaxX].x * o.m11 + v[imaxX].y * o.m21 + v[imaxX].z * o.m31,
               v[imaxY].x * o.m12 + v[imaxY].y * o.m22 + v[imaxY].z * o.m32,
               v[imaxZ].x * o.m13 + v[imaxZ].y * o.m23 + v[imaxZ].z * o.m33};
        V3 neg_{v[iminX].x * o.m11 + v[iminX].y * o.m21 + v[iminX].z * o.m31,
                v[iminY].x * o.m12 + v[iminY].y * o.m22 + v[iminY].z * o.m32,
                v[iminZ].x * o.m13 + v[iminZ].y * o.m23 + v[iminZ].z * o.m33};
        float m = b.margin;
        *bmax = V3{bpos.x + pos.x + m, bpos.y + pos.y + m, bpos.z + pos.z + m};
        *bmin = V3{bpos.x + neg_.x - m, bpos.y + neg_.y - m, bpos.z + neg_.z - m};
    } else {
        float r = b.radius;
        *bmin = V3{bpos.x - r, bpos.y - r, bpos.z - r};
        *bmax = V3{bpos.x + r, bpos.y + r, bpos.z + r};
    }
}

// EntityCollidable.UpdateBoundingBox -> Shape.GetBoundingBox -> ExpandBoundingBox
// (EntityCollidable.cs:140-148,195-257)
void bounding_box(refcpu_world* w, Body& b) {
    // BoundingBoxUpdater.cs:48: active entries only (a sleeping body keeps its
    // box; recomputing an inactive kinematic body's gives the same box).
    if (b.removed || (b.dynamic && b.asleep)) return;
    if (is_static(b)) return;  // a static mesh's or terrain's box is fixed
    shape_bounding_box(w, b, b.pos, b.q, &b.bmin, &b.bmax);
    float dt = w->s.dt;
    if (dt > 0) {
        if (b.v.x > 0) b.bmax.x += b.v.x * dt * 1.0f; else b.bmin.x += b.v.x * dt * 1.0f;
        if (b.v.y > 0) b.bmax.y += b.v.y * dt * 1.0f; else b.bmin.y += b.v.y * dt * 1.0f;
        if (b.v.z > 0) b.bmax.z += b.v.z * dt * 1.0f; else b.bmin.z += b.v.z * dt * 1.0f;
    }
}

// Entity.PreUpdatePosition (EntityBase.cs:1124-1155): the orientation of
// every active body; the position of the discrete ones.
void integrate(refcpu_world* w, Body& b) {
    if (!body_active(b)) return;  // ContinuousPositionUpdater.cs:71: active entities only
    float dt = w->s.dt;
    b.q_world = b.q;
    V3 inc = mul(b.w, dt * .5f);
    Q4 m{inc.x, inc.y, inc.z, 0};
    m = qmul(m, b.q);
    b.q = qadd(b.q, m);
    b.q = qnormalize(b.q);
    b.R = m3_from_quat(b.q);
    if (b.pmode == 0) {
        inc = mul(b.v, dt);
        b.pos = add(b.pos, inc);
    }
}

ConvexShapeRef shape_ref(const refcpu_world* w, const Body& b);

// The collidable's world transform while the times of impact are computed:
// a moving passive or continuous body still has the pose of the step's
// start (its orientation was turned, its transform not refreshed), the
// others their current pose.
RT ccd_transform(const Body& b) {
    return RT{b.pos, (b.pmode != 0 && body_active(b)) ? b.q_world : b.q};
}

void mesh_time_of_impact(refcpu_world* w, Pair& p);

// ConvexPairHandler.UpdateTimeOfImpact (ConvexPairHandler.cs:37-111), as the
// pair's continuous bodies request it (EntityBase.cs:1071-1082: every
// continuous body asks each of its pairs, active or not; with two continuous
// bodies only one request computes).  A pair that no request computes keeps
// its time of impact.
void pair_time_of_impact(refcpu_world* w, Pair& p) {
    if (p.kind == BEPU_PAIR_MESH_CONVEX || p.kind == BEPU_PAIR_MESH_SPHERE) {
        mesh_time_of_impact(w, p);
        return;
    }
    const Body& A = w->bodies[p.a];
    const Body& B = w->bodies[p.b];
    if (A.pmode != 2 && B.pmode != 2) return;
    if (!body_active(A) && !body_active(B)) return;
    // Discrete bodies already moved: the sweep is the other body's motion.
    V3 vel;
    if (A.pmode == 0) vel = B.v;
    else if (B.pmode == 0) vel = neg(A.v);
    else vel = sub(B.v, A.v);
    vel = mul(vel, w->s.dt);
    const float vsq = length_sq(vel);
    const float scaling = w->s.core_shape_scaling;
    const RT ta = ccd_transform(A), tb = ccd_transform(B);
    float t;
    const float ra = A.min_radius * scaling;
    p.toi = 1;
    if (ra * ra < vsq && ccd_sphere_cast(ta.pos, neg(vel), ra, shape_ref(w, B), tb, p.toi, scaling, &t)) p.toi = t;
    const float rb = B.min_radius * scaling;
    if (rb * rb < vsq && ccd_sphere_cast(tb.pos, vel, rb, shape_ref(w, A), ta, p.toi, scaling, &t)) p.toi = t;
    // Embedded from the start: no clamping (the contacts handle it).
    if (p.toi == 0) p.toi = 1;
}

// ContinuousPositionUpdater.UpdateSingleThreaded after the pre-update
// (ContinuousPositionUpdater.cs:140-154): times of impact of the pairs with a
// continuous body, then Entity.UpdatePositionContinuously (EntityBase.cs:
// 1093-1121) for every active passive or continuous body: the linear motion
// scaled by the smallest time of impact of its pairs.
void continuous_update(refcpu_world* w) {
    std::vector<float> least(w->bodies.size(), 1.0f);
    for (Pair& p : w->pairs.vals) {
        pair_time_of_impact(w, p);
        if (p.toi < least[p.a]) least[p.a] = p.toi;
        if (p.toi < least[p.b]) least[p.b] = p.toi;
    }
    const float dt = w->s.dt;
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        Body& b = w->bodies[i];
        if (b.pmode == 0 || !body_active(b)) continue;
        V3 inc = mul(b.v, dt * least[i]);
        b.pos = add(b.pos, inc);
    }
}

// ---------------------------------------------------------------------------
// Broadphase: the pair set of DynamicHierarchy equals the closed-interval
// AABB overlap set (BoundingBox.Intersects, BoundingBox.cs:58-69) filtered
// by the default collision rules: kinematic-kinematic pairs are NoBroadPhase
// (CollisionRules.cs:175-178) and one entry must be active (BroadPhase.cs:97-102).

inline bool intersects(const Body& a, const Body& b) {
    if (a.removed || b.removed) return false;
    if (b.bmin.x > a.bmax.x || b.bmin.y > a.bmax.y || b.bmin.z > a.bmax.z) return false;
    if (a.bmin.x > b.bmax.x || a.bmin.y > b.bmax.y || a.bmin.z > b.bmax.z) return false;
    return true;
}
inline uint64_t pair_key(uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return ((uint64_t)lo << 32) | hi;
}

void broadphase_bruteforce(refcpu_world* w, std::vector<uint64_t>* out) {
    out->clear();
    size_t n = w->bodies.size();
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) {
            const Body& a = w->bodies[i];
            const Body& b = w->bodies[j];
            if (!a.dynamic && !b.dynamic) continue;
            if (intersects(a, b)) out->push_back(pair_key((uint32_t)i, (uint32_t)j));
        }
    std::sort(out->begin(), out->end());
}

// SortAndSweep1D-style sweep along x (SortAndSweep/SortAndSweep1D.cs:140-172); same set.
void broadphase_sweep(refcpu_world* w, std::vector<uint64_t>* out) {
    out->clear();
    size_t n = w->bodies.size();
    std::vector<uint32_t> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        float fa = w->bodies[a].bmin.x, fb = w->bodies[b].bmin.x;
        return fa < fb || (fa == fb && a < b);
    });
    // The sweep is split over ranges of the sorted order; the pair set is
    // sorted afterwards, so the split does not change the result.
    const size_t chunk = 2048;
    std::vector<std::vector<uint64_t>> parts((n + chunk - 1) / chunk);
    par_for(w, parts.size(), 1, [&](size_t p0, size_t p1) {
        for (size_t p = p0; p < p1; ++p) {
            std::vector<uint64_t>& o = parts[p];
            for (size_t i = p * chunk; i < std::min(n, (p + 1) * chunk); ++i) {
                const Body& a = w->bodies[idx[i]];
                for (size_t j = i + 1; j < n; ++j) {
                    const Body& b = w->bodies[idx[j]];
                    if (b.bmin.x > a.bmax.x) break;
                    if (!a.dynamic && !b.dynamic) continue;
                    if (intersects(a, b)) o.push_back(pair_key(idx[i], idx[j]));
                }
            }
        }
    });
    for (auto& o : parts) out->insert(out->end(), o.begin(), o.end());
    std::sort(out->begin(), out->end());
}

// The same pair set from a uniform grid: cell edge twice the median largest
// AABB side; a body spanning more than 8 cells per axis (or with a non-finite
// box) is tested against every body instead.  A pair is emitted only from
// the cell that holds the maximum of the two boxes' minimum corners, a point
// of their intersection, so each overlapping pair is found exactly once.
void broadphase_grid(refcpu_world* w, std::vector<uint64_t>* out) {
    out->clear();
    size_t n = w->bodies.size();
    if (n < 2) return;
    std::vector<float> side;
    side.reserve(n);
    for (const Body& b : w->bodies) {
        if (b.removed) continue;
        float e = std::max(std::max(b.bmax.x - b.bmin.x, b.bmax.y - b.bmin.y), b.bmax.z - b.bmin.z);
        if (std::isfinite(e)) side.push_back(e);
    }
    if (side.empty()) return;
    std::nth_element(side.begin(), side.begin() + side.size() / 2, side.end());
    double h = std::max(2.0 * (double)side[side.size() / 2], 1e-3);
    auto cell = [&](float v) { return (int64_t)std::floor((double)v / h); };
    auto pack = [](int64_t x, int64_t y, int64_t z) {
        return ((uint64_t)(x & 0x1fffff) << 42) | ((uint64_t)(y & 0x1fffff) << 21) | (uint64_t)(z & 0x1fffff);
    };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    std::vector<uint32_t> large;
    for (uint32_t i = 0; i < n; ++i) {
        const Body& b = w->bodies[i];
        if (b.removed) continue;
        bool finite = std::isfinite(b.bmin.x) && std::isfinite(b.bmin.y) && std::isfinite(b.bmin.z) &&
                      std::isfinite(b.bmax.x) && std::isfinite(b.bmax.y) && std::isfinite(b.bmax.z);
        if (!finite) { large.push_back(i); continue; }
        int64_t x0 = cell(b.bmin.x), x1 = cell(b.bmax.x), y0 = cell(b.bmin.y), y1 = cell(b.bmax.y);
        int64_t z0 = cell(b.bmin.z), z1 = cell(b.bmax.z);
        if (x1 - x0 > 8 || y1 - y0 > 8 || z1 - z0 > 8 || std::llabs(x0) > (1 << 19) || std::llabs(y0) > (1 << 19) ||
            std::llabs(z0) > (1 << 19)) {
            large.push_back(i);
            continue;
        }
        for (int64_t x = x0; x <= x1; ++x)
            for (int64_t y = y0; y <= y1; ++y)
                for (int64_t z = z0; z <= z1; ++z) grid[pack(x, y, z)].push_back(i);
    }
    std::vector<std::pair<uint64_t, const std::vector<uint32_t>*>> cells;
    cells.reserve(grid.size());
    for (auto& kv : grid) cells.push_back({kv.first, &kv.second});
    std::vector<std::vector<uint64_t>> parts(cells.size() / 256 + 1);
    par_for(w, parts.size(), 1, [&](size_t p0, size_t p1) {
        for (size_t p = p0; p < p1; ++p) {
            for (size_t c = p * 256; c < std::min(cells.size(), (p + 1) * 256); ++c) {
                const std::vector<uint32_t>& l = *cells[c].second;
                for (size_t i = 0; i < l.size(); ++i)
                    for (size_t j = i + 1; j < l.size(); ++j) {
                        const Body& a = w->bodies[l[i]];
                        const Body& b = w->bodies[l[j]];
                        if (!a.dynamic && !b.dynamic) continue;
                        if (!intersects(a, b)) continue;
                        uint64_t own = pack(cell(std::max(a.bmin.x, b.bmin.x)), cell(std::max(a.bmin.y, b.bmin.y)),
                                            cell(std::max(a.bmin.z, b.bmin.z)));
                        if (own == cells[c].first) parts[p].push_back(pair_key(l[i], l[j]));
                    }
            }
        }
    });
    for (auto& o : parts) out->insert(out->end(), o.begin(), o.end());
    std::vector<char> is_large(n, 0);
    for (uint32_t i : large) is_large[i] = 1;
    for (uint32_t L : large)
        for (uint32_t j = 0; j < n; ++j) {
            if (j == L || (is_large[j] && j < L)) continue;
            const Body& a = w->bodies[L];
            const Body& b = w->bodies[j];
            if (!a.dynamic && !b.dynamic) continue;
            if (intersects(a, b)) out->push_back(pair_key(L, j));
        }
    std::sort(out->begin(), out->end());
}

// ---------------------------------------------------------------------------
// Contact constraint bookkeeping (ConvexContactManifoldConstraint.cs:117-155,
// NonConvexContactManifoldConstraint.cs:98-139, ContactManifold.cs)

inline bool convex_kind(int kind) { return kind == BEPU_PAIR_BOX_BOX || kind == BEPU_PAIR_CONVEX; }

void constraint_add_contact(Pair& p, int slot) {
    PenC& pc = p.pen[p.npen++];
    std::memset(&pc, 0, sizeof(pc));
    pc.slot = slot;
    pc.acc = 0;
    if (convex_kind(p.kind)) {
        if (!p.tw.active) {
            // TwistFrictionConstraint.Setup / SlidingFrictionTwoAxis.Setup
            p.tw.active = true;
            p.sl.active = true;
            for (int k = 0; k < 6; ++k) p.sl.lin[k] = 0;
        }
    } else {
        FricC& f = p.fric[p.nfric++];
        std::memset(&f, 0, sizeof(f));
        f.acc = 0;
        f.lin[0] = f.lin[1] = f.lin[2] = 0;
    }
}

void constraint_remove_contact(Pair& p, int slot) {
    int idx = -1;
    for (int i = 0; i < p.npen; ++i)
        if (p.pen[i].slot == slot) { idx = i; break; }
    if (idx >= 0) {
        for (int k = idx; k < p.npen - 1; ++k) p.pen[k] = p.pen[k + 1];
        p.npen--;
    }
    if (convex_kind(p.kind)) {
        if (p.npen == 0) {
            p.tw.active = false;
            p.tw.acc = 0;
            p.sl.active = false;
            p.sl.acc[0] = p.sl.acc[1] = 0;
        }
    } else if (idx >= 0) {
        // The friction constraint added with the removed penetration constraint
        // sits at the same list position.
        for (int k = idx; k < p.nfric - 1; ++k) p.fric[k] = p.fric[k + 1];
        p.nfric--;
    }
}

// ContactManifold.Remove (FastRemoveAt on the contact list, ContactManifold.cs)
void manifold_remove(Pair& p, int index) {
    int slot = p.clist[index];
    p.ccount--;
    if (index < p.ccount) p.clist[index] = p.clist[p.ccount];
    p.used[slot] = false;
    constraint_remove_contact(p, slot);
}
// ContactManifold.Add
void manifold_add(Pair& p, const ContactData& c) {
    int slot = 0;
    while (p.used[slot]) ++slot;
    p.used[slot] = true;
    p.slot[slot].pos = c.pos;
    p.slot[slot].normal = c.normal;
    p.slot[slot].depth = c.depth;
    p.slot[slot].id = c.id;
    p.clist[p.ccount++] = slot;
    constraint_add_contact(p, slot);
}

// BoxContactManifold.Update, non-ALLOWUNSAFE (BoxContactManifold.cs:125-206)
void update_box_box(refcpu_world* w, Pair& p) {
    const Body& A = w->bodies[p.a];
    const Body& B = w->bodies[p.b];
    float distance;
    V3 axis;
    BoxContactList m;
    if (box_box_colliding(A.half, B.half, A.pos, A.q, B.pos, B.q, &distance, &axis, &m)) {
        axis = neg(axis);
        int toRemove[4];
        int nrem = 0;
        for (int i = 0; i < p.ccount; i++) {
            bool found = false;
            ContactObj& c = p.slot[p.clist[i]];
            for (int j = m.n - 1; j >= 0; j--) {
                if (c.id == m.e[j].id) {
                    found = true;
                    c.pos = m.e[j].pos;
                    c.depth = -m.e[j].depth;
                    c.normal = axis;
                    m.remove_at(j);
                    break;
                }
            }
            if (!found) toRemove[nrem++] = i;
        }
        for (int i = nrem - 1; i >= 0; i--) manifold_remove(p, toRemove[i]);
        for (int i = 0; i < m.n; i++) {
            ContactData nc{m.e[i].pos, axis, -m.e[i].depth, m.e[i].id};
            manifold_add(p, nc);
        }
    } else {
        for (int i = p.ccount - 1; i >= 0; i--) manifold_remove(p, i);
    }
}

// SphereContactManifold.Update / BoxSphereContactManifold.Update
// (SphereContactManifold.cs:55-81, BoxSphereContactManifold.cs:55-81)
void update_single_contact(refcpu_world* w, Pair& p) {
    const Body& A = w->bodies[p.a];
    const Body& B = w->bodies[p.b];
    ContactData cd{};
    bool hit;
    if (p.kind == BEPU_PAIR_SPHERE_SPHERE)
        hit = sphere_sphere_colliding(A.radius, B.radius, A.pos, B.pos, w->s.max_contact_distance, &cd);
    else
        hit = box_sphere_colliding(A.half, A.pos, A.q, B.radius, B.pos, w->s.max_contact_distance, &cd);
    bool colliding = false;
    if (hit) {
        if (!p.prev_colliding && cd.depth >= 0) {
            cd.id = 0;
            manifold_add(p, cd);
            colliding = true;
        } else if (p.prev_colliding) {
            ContactObj& c = p.slot[p.clist[0]];
            c.normal = cd.normal;
            c.depth = cd.depth;
            c.pos = cd.pos;
            colliding = true;
        }
    } else if (p.prev_colliding) {
        // SphereContactManifold.Remove uses an ordered RemoveAt on a 1-element list.
        manifold_remove(p, 0);
    }
    p.prev_colliding = colliding;
}

ConvexShapeRef shape_ref(const refcpu_world* w, const Body& b) {
    ConvexShapeRef s;
    std::memset(&s, 0, sizeof(s));
    s.kind = b.shape;
    s.half = b.half;
    s.half_length = b.half_length;
    s.margin = b.margin;
    if (b.shape == BEPU_SHAPE_CONVEX_HULL) {
        s.verts = &w->hull_verts[b.hull_off];
        s.nverts = b.hull_n;
    }
    return s;
}

// GeneralConvexContactManifold.Add: the contact plus its supplement data
// (GeneralConvexContactManifold.cs:128-137).
void manifold_add_general(Pair& p, const ContactData& c, const RT& ta, const RT& tb) {
    int slot = 0;
    while (p.used[slot]) ++slot;
    p.base_depth[slot] = c.depth;
    p.off_a[slot] = rt_transform_by_inverse(c.pos, ta);
    p.off_b[slot] = rt_transform_by_inverse(c.pos, tb);
    manifold_add(p, c);
}

// ContactReducer.ReduceContacts for four contacts and one candidate
// (CollisionTests/ContactReducer.cs:197-318).  Returns the list index to
// remove (or -1) and whether the candidate is added.
int reduce_contacts(const Pair& p, const ContactData& cand, bool* add_candidate) {
    const ContactObj* c[4];
    for (int i = 0; i < 4; ++i) c[i] = &p.slot[p.clist[i]];
    float maxDepth = -FLT_MAX;
    int deepest = -1;
    for (int i = 0; i < 4; i++)
        if (c[i]->depth > maxDepth) { deepest = i; maxDepth = c[i]->depth; }
    if (cand.depth > maxDepth) deepest = 4;
    V3 deepPos = deepest < 4 ? c[deepest]->pos : cand.pos;
    float furthestDist = 0;
    int furthest = -1;
    for (int i = 0; i < 4; i++) {
        float d = distance_sq(c[i]->pos, deepPos);
        if (d > furthestDist) { furthestDist = d; furthest = i; }
    }
    if (distance_sq(cand.pos, deepPos) > furthestDist) furthest = 4;
    V3 farPos = furthest < 4 ? c[furthest]->pos : cand.pos;
    V3 xAxis = sub(deepPos, farPos);
    V3 yAxis = cross(xAxis, c[0]->normal);
    float minY = FLT_MAX, maxY = -FLT_MAX;
    int iminY = -1, imaxY = -1;
    for (int i = 0; i < 4; i++) {
        float d = dot(c[i]->pos, yAxis);
        if (d < minY) { iminY = i; minY = d; }
        if (d > maxY) { imaxY = i; maxY = d; }
    }
    float d = dot(cand.pos, yAxis);
    if (d < minY) iminY = 4;
    if (d > maxY) imaxY = 4;
    if (deepest == 4 || furthest == 4 || iminY == 4 || imaxY == 4) {
        *add_candidate = true;
        for (int i = 0; i < 4; i++)
            if (!(i == deepest || i == furthest || i == iminY || i == imaxY)) return i;
        return -1;
    }
    *add_candidate = false;
    return -1;
}

// GeneralConvexContactManifold.Update (GeneralConvexContactManifold.cs:71-126) with
// ContactRefresher.ContactRefresh (CollisionTests/ContactRefresher.cs:18-69) and
// IsContactUnique (GeneralConvexContactManifold.cs:145-166).
void update_general_convex(refcpu_world* w, Pair& p) {
    const Body& A = w->bodies[p.a];
    const Body& B = w->bodies[p.b];
    RT ta{A.pos, A.q}, tb{B.pos, B.q};
    int rem[4], nrem = 0;
    for (int k = 0; k < p.ccount; k++) {
        int slot = p.clist[k];
        ContactObj& c = p.slot[slot];
        V3 na = rt_transform(p.off_a[slot], ta);
        V3 nb = rt_transform(p.off_b[slot], tb);
        V3 ab = sub(nb, na);
        float d = dot(ab, c.normal);
        V3 t = mul(c.normal, d);
        t = sub(ab, t);
        d = length_sq(t);
        if (d > w->s.contact_invalidation_length_sq) {
            rem[nrem++] = k;
        } else {
            d = dot(ab, c.normal);
            c.depth = p.base_depth[slot] - d;
            if (c.depth < -w->s.max_contact_distance) rem[nrem++] = k;
            else c.pos = mul(add(nb, na), .5f);
        }
    }
    for (int i = nrem - 1; i >= 0; i--) manifold_remove(p, rem[i]);
    ContactData cand;
    if (convex_candidate(shape_ref(w, A), shape_ref(w, B), ta, tb, A.v, B.v, &p.gc, &cand)) {
        for (int i = 0; i < p.ccount; ++i) {
            if (dot(p.slot[p.clist[i]].normal, cand.normal) < 0) {
                manifold_remove(p, i);
                break;
            }
        }
        bool unique = true;
        for (int i = 0; i < p.ccount; i++) {
            int slot = p.clist[i];
            ContactObj& c = p.slot[slot];
            if (distance_sq(c.pos, cand.pos) < w->s.contact_min_separation_sq) {
                c.normal = cand.normal;
                c.pos = cand.pos;
                c.depth = cand.depth;
                p.base_depth[slot] = cand.depth;
                p.off_a[slot] = rt_transform_by_inverse(cand.pos, ta);
                p.off_b[slot] = rt_transform_by_inverse(cand.pos, tb);
                unique = false;
                break;
            }
        }
        if (unique) {
            if (p.ccount == 4) {
                bool add_cand;
                int r = reduce_contacts(p, cand, &add_cand);
                if (r >= 0) manifold_remove(p, r);
                if (add_cand) manifold_add_general(p, cand, ta, tb);
            } else {
                manifold_add_general(p, cand, ta, tb);
            }
        }
    } else {
        for (int i = p.ccount - 1; i >= 0; i--) manifold_remove(p, i);
    }
}

// ---------------------------------------------------------------------------
// Static meshes: StaticMeshContactManifold over TriangleMeshConvexContactManifold
// (CollisionTests/Manifolds/TriangleMeshConvexContactManifold.cs:101-434,
// StaticMeshContactManifold.cs:33-71).  The mesh transform of the manifold is
// the identity (MeshTransform, TriangleMeshConvexContactManifold.cs:76-82):
// the mesh's vertices reach world space through its data's affine transform.

const RT kMeshTransform{V3{0, 0, 0}, Q4{0, 0, 0, 1}};
// CollisionDetectionSettings.nonconvexNormalDotMinimum (CollisionDetectionSettings.cs:52)
const float kNonconvexNormalDotMinimum = .99f;

// ContactReducer.ReduceContacts, general form (CollisionTests/ContactReducer.cs:19-195):
// the manifold's contacts (list order) and the candidates; fills the contact
// list indices to remove (ascending) and the candidates kept.
void reduce_contacts_general(const Pair& p, const std::vector<ContactData>& cand, std::vector<int>* remove,
                             std::vector<ContactData>* keep) {
    const int nc = p.ccount;
    auto cpos = [&](int i) { return p.slot[p.clist[i]].pos; };
    float maxDepth = -FLT_MAX;
    int deepest = -1;
    V3 normal{0, 0, 0};
    for (int i = 0; i < nc; i++) {
        const ContactObj& c = p.slot[p.clist[i]];
        normal = add(normal, c.normal);
        if (c.depth > maxDepth) { deepest = i; maxDepth = c.depth; }
    }
    for (size_t i = 0; i < cand.size(); i++) {
        normal = add(normal, cand[i].normal);
        if (cand[i].depth > maxDepth) { deepest = nc + (int)i; maxDepth = cand[i].depth; }
    }
    if (length_sq(normal) < 1e-7f) {
        if (nc > 0) normal = p.slot[p.clist[0]].normal;
        else normal = cand[0].normal;
    }
    V3 deepPos = deepest < nc ? cpos(deepest) : cand[deepest - nc].pos;
    float furthestDist = 0;
    int furthest = -1;
    for (int i = 0; i < nc; i++) {
        float d = distance_sq(cpos(i), deepPos);
        if (d > furthestDist) { furthestDist = d; furthest = i; }
    }
    for (size_t i = 0; i < cand.size(); i++) {
        float d = distance_sq(cand[i].pos, deepPos);
        if (d > furthestDist) { furthestDist = d; furthest = nc + (int)i; }
    }
    if (furthest == -1) {
        if (nc > 0) {
            for (int i = 1; i < nc; i++) remove->push_back(i);
            return;
        }
        keep->push_back(cand[0]);
        return;
    }
    V3 farPos = furthest < nc ? cpos(furthest) : cand[furthest - nc].pos;
    V3 xAxis = sub(deepPos, farPos);
    V3 yAxis = cross(xAxis, normal);
    float minY = FLT_MAX, maxY = -FLT_MAX;
    int iminY = -1, imaxY = -1;
    for (int i = 0; i < nc; i++) {
        float d = dot(cpos(i), yAxis);
        if (d < minY) { iminY = i; minY = d; }
        if (d > maxY) { imaxY = i; maxY = d; }
    }
    for (size_t i = 0; i < cand.size(); i++) {
        float d = dot(cand[i].pos, yAxis);
        if (d < minY) { iminY = nc + (int)i; minY = d; }
        if (d > maxY) { imaxY = nc + (int)i; maxY = d; }
    }
    for (size_t i = 0; i < cand.size(); i++) {
        int t = (int)i + nc;
        if (t == deepest || t == furthest || t == iminY || t == imaxY) keep->push_back(cand[i]);
    }
    for (int i = 0; i < nc; i++)
        if (!(i == deepest || i == furthest || i == iminY || i == imaxY)) remove->push_back(i);
}

// The boundary sets of the improved boundary handling
// (TriangleMeshConvexContactManifold.cs:27-49, 773-787): edges are unordered
// vertex index pairs.
struct BoundaryContact {
    bool should_correct;
    V3 corrected;
    int e0, e1;  // edge (vertex contacts: e0 is the vertex)
    ContactData c;
};
struct BoundarySets {
    std::vector<int> blocked_vertices;
    std::vector<std::pair<int, int>> blocked_edges;
    std::vector<BoundaryContact> edges, vertices;
    void block_vertex(int v) {
        if (!has_vertex(v)) blocked_vertices.push_back(v);
    }
    void block_edge(int a, int b) {
        if (!has_edge(a, b)) blocked_edges.push_back({a, b});
    }
    bool has_vertex(int v) const {
        for (int x : blocked_vertices)
            if (x == v) return true;
        return false;
    }
    bool has_edge(int a, int b) const {
        for (const auto& e : blocked_edges)
            if ((e.first == a && e.second == b) || (e.first == b && e.second == a)) return true;
        return false;
    }
};

// GetNormal (TriangleMeshConvexContactManifold.cs:450-481)
V3 mesh_corrected_normal(const V3& uncorrected, const TriangleRef& t) {
    V3 ab = sub(t.v[1], t.v[0]), ac = sub(t.v[2], t.v[0]);
    V3 n;
    if (t.sidedness == SIDED_DOUBLE) {
        n = cross(ab, ac);
        if (dot(n, uncorrected) < 0) n = neg(n);
    } else if (t.sidedness == SIDED_CLOCKWISE) {
        n = cross(ac, ab);
    } else {
        n = cross(ab, ac);
    }
    if (length_sq(n) < 1e-7f) n = uncorrected;
    return n;
}

// AnalyzeCandidate (TriangleMeshConvexContactManifold.cs:483-596): true for a
// face contact (always kept); edge and vertex contacts are queued and every
// other region of the triangle is blocked.
bool mesh_analyze(int region, const int ix[3], const TriangleRef& t, bool should_correct, const ContactData& c,
                  BoundarySets& sets) {
    const int A = ix[0], B = ix[1], C = ix[2];
    BoundaryContact bc;
    bc.should_correct = should_correct;
    bc.c = c;
    switch (region) {
    case VR_A:
    case VR_B:
    case VR_C: {
        const int v = region == VR_A ? A : (region == VR_B ? B : C);
        bc.corrected = mesh_corrected_normal(c.normal, t);
        bc.e0 = v;
        bc.e1 = -1;
        sets.vertices.push_back(bc);
        sets.block_edge(A, B);
        sets.block_edge(B, C);
        sets.block_edge(A, C);
        if (region != VR_A) sets.block_vertex(A);
        if (region != VR_B) sets.block_vertex(B);
        if (region != VR_C) sets.block_vertex(C);
        return false;
    }
    case VR_AB:
    case VR_AC:
    case VR_BC: {
        bc.corrected = mesh_corrected_normal(c.normal, t);
        bc.e0 = region == VR_BC ? B : A;
        bc.e1 = region == VR_AB ? B : C;
        sets.edges.push_back(bc);
        if (region != VR_AB) sets.block_edge(A, B);
        if (region != VR_BC) sets.block_edge(B, C);
        if (region != VR_AC) sets.block_edge(A, C);
        sets.block_vertex(A);
        sets.block_vertex(B);
        sets.block_vertex(C);
        return false;
    }
    default:
        sets.block_edge(A, B);
        sets.block_edge(B, C);
        sets.block_edge(A, C);
        sets.block_vertex(A);
        sets.block_vertex(B);
        sets.block_vertex(C);
        return true;
    }
}

// TriangleMeshConvexContactManifold.Update (TriangleMeshConvexContactManifold.cs:101-434)
// ConvexShape.GetLocalBoundingBox (ConvexShapes/ConvexShape.cs:225-281): the
// shape at `shape_xf` in the frame of the affine `space`, from the extreme
// points (with margin) along that frame's axes.
void convex_local_bounding_box(const ConvexShapeRef& cvx, const RT& shape_xf, const Affine& space, V3* mn, V3* mx) {
    const Affine inv = affine_invert(space);
    // AffineTransform.Multiply(RigidTransform, AffineTransform) (AffineTransform.cs:202-211)
    Affine t;
    t.m = m3_mul(m3_from_quat(shape_xf.q), inv.m);
    t.t = add(m3_transform(shape_xf.pos, inv.m), inv.t);
    const M3& m = t.m;
    V3 right = local_extreme_point(cvx, V3{m.m11, m.m21, m.m31});
    V3 left = local_extreme_point(cvx, V3{-m.m11, -m.m21, -m.m31});
    V3 up = local_extreme_point(cvx, V3{m.m12, m.m22, m.m32});
    V3 down = local_extreme_point(cvx, V3{-m.m12, -m.m22, -m.m32});
    V3 backward = local_extreme_point(cvx, V3{m.m13, m.m23, m.m33});
    V3 forward = local_extreme_point(cvx, V3{-m.m13, -m.m23, -m.m33});
    // EntityShape.TransformLocalExtremePoints (EntityShape.cs:58-63)
    V3 pos{right.x * m.m11 + right.y * m.m21 + right.z * m.m31, up.x * m.m12 + up.y * m.m22 + up.z * m.m32,
           backward.x * m.m13 + backward.y * m.m23 + backward.z * m.m33};
    V3 neg_{left.x * m.m11 + left.y * m.m21 + left.z * m.m31, down.x * m.m12 + down.y * m.m22 + down.z * m.m32,
            forward.x * m.m13 + forward.y * m.m23 + forward.z * m.m33};
    *mx = V3{t.t.x + pos.x, t.t.y + pos.y, t.t.z + pos.z};
    *mn = V3{t.t.x + neg_.x, t.t.y + neg_.y, t.t.z + neg_.z};
}

// TerrainContactManifold.FindOverlappingTriangles (TerrainContactManifold.cs:30-64):
// the convex's box in the terrain's local frame, grown by its motion over the step.
void terrain_query_box(refcpu_world* w, const Body& A, const Affine& xf, V3* mn, V3* mx) {
    convex_local_bounding_box(shape_ref(w, A), RT{A.pos, A.q}, xf, mn, mx);
    M3 inverse = m3_invert(xf.m);
    V3 tv = m3_transform(A.v, inverse);
    tv = mul(tv, w->s.dt);
    if (tv.x > 0) mx->x += tv.x; else mn->x += tv.x;
    if (tv.y > 0) mx->y += tv.y; else mn->y += tv.y;
    if (tv.z > 0) mx->z += tv.z; else mn->z += tv.z;
}

void update_mesh(refcpu_world* w, Pair& p) {
    const Body& A = w->bodies[p.a];
    const Body& M = w->bodies[p.b];
    const bool terrain = M.terrain >= 0;
    const TerrainShapeRef* ts = terrain ? &w->terrains[M.terrain].shape : nullptr;
    const MeshShape* shape = terrain ? nullptr : &w->mesh_shapes[M.mesh];
    const RT ta{A.pos, A.q};
    // ContactRefresher.ContactRefresh (ContactRefresher.cs:18-69)
    int rem[4], nrem = 0;
    for (int k = 0; k < p.ccount; k++) {
        int slot = p.clist[k];
        ContactObj& c = p.slot[slot];
        V3 na = rt_transform(p.off_a[slot], ta);
        V3 nb = rt_transform(p.off_b[slot], kMeshTransform);
        V3 ab = sub(nb, na);
        float d = dot(ab, c.normal);
        V3 t = mul(c.normal, d);
        t = sub(ab, t);
        d = length_sq(t);
        if (d > w->s.contact_invalidation_length_sq) {
            rem[nrem++] = k;
        } else {
            d = dot(ab, c.normal);
            c.depth = p.base_depth[slot] - d;
            if (c.depth < -w->s.max_contact_distance) rem[nrem++] = k;
            else c.pos = mul(add(nb, na), .5f);
        }
    }
    for (int i = nrem - 1; i >= 0; i--) manifold_remove(p, rem[i]);

    std::vector<int> tris;
    if (terrain) {
        terrain_query_box(w, A, ts->xf, &p.tq_min, &p.tq_max);
        p.tq_valid = true;
        terrain_overlaps(*ts, p.tq_min, p.tq_max, &tris);
    } else {
        mesh_overlaps(*shape, A.bmin, A.bmax, &tris);
    }
    const bool improve = M.improve_boundary;
    BoundarySets sets;
    std::vector<ContactData> cand;

    // RigidTransform.Invert (RigidTransform.cs:95-100), AffineTransform.CreateFromRigidTransform
    // (AffineTransform.cs:232-236), PrecomputeTriangleTransform (StaticMeshContactManifold.cs:44-49).
    RT inv;
    inv.q = qconj(A.q);
    inv.pos = neg(qtransform(A.pos, inv.q));
    Affine convex_inverse{m3_from_quat(inv.q), inv.pos};
    const Affine to_local = affine_multiply(terrain ? ts->xf : shape->xf, convex_inverse);
    V3 lmin, lmax;
    shape_bounding_box(w, A, V3{0, 0, 0}, Q4{0, 0, 0, 1}, &lmin, &lmax);
    const M3 orientation = m3_from_quat(A.q);
    const ConvexShapeRef cvx = shape_ref(w, A);
    const bool sphere = p.kind == BEPU_PAIR_MESH_SPHERE;
    std::vector<MeshTester>* testers = sphere ? nullptr : &w->mesh_testers[p.mesh_slot];
    int guaranteed = 0;

    // AddLocalContact and IsContactUnique (TriangleMeshConvexContactManifold.cs:436-447, 617-674)
    auto add_local = [&](ContactData c) {
        c.pos = add(m3_transform(c.pos, orientation), A.pos);
        c.normal = m3_transform(c.normal, orientation);
        for (int i = 0; i < p.ccount; i++) {
            const int slot = p.clist[i];
            ContactObj& e = p.slot[slot];
            if (distance_sq(e.pos, c.pos) < w->s.contact_min_separation_sq) {
                float d = dot(e.normal, c.normal);
                if (std::fabs(d) >= kNonconvexNormalDotMinimum) {
                    e.normal = c.normal;
                    e.pos = c.pos;
                    e.depth = c.depth;
                    p.base_depth[slot] = c.depth;
                    p.off_a[slot] = rt_transform_by_inverse(c.pos, ta);
                    p.off_b[slot] = rt_transform_by_inverse(c.pos, kMeshTransform);
                    return;
                }
            }
        }
        for (const ContactData& e : cand) {
            if (distance_sq(e.pos, c.pos) < w->s.contact_min_separation_sq) {
                float d = dot(e.normal, c.normal);
                if (std::fabs(d) >= kNonconvexNormalDotMinimum) return;
            }
        }
        cand.push_back(c);
    };

    for (int tri : tris) {
        // ConfigureLocalTriangle (StaticMeshContactManifold.cs:51-66, TerrainContactManifold.cs:76-93)
        int ix[3];
        TriangleRef t;
        if (terrain) {
            int col[3], row[3];
            terrain_local_indices(*ts, tri, col, row);
            for (int k = 0; k < 3; ++k) {
                ix[k] = row[k] * ts->width + col[k];  // TerrainVertexIndices.ToSequentialIndex
                t.v[k] = affine_transform(V3{(float)col[k], ts->h(col[k], row[k]), (float)row[k]}, to_local);
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                ix[k] = shape->idx[tri + k];
                t.v[k] = affine_transform(shape->verts[ix[k]], to_local);
            }
        }
        t.sidedness = M.sidedness;
        t.margin = 0;
        // Toolbox.GetTriangleBoundingBox (BEPUutilities/Toolbox.cs:1516-1580) against the local box.
        V3 tmin, tmax;
        triangle_bounding_box(t.v[0], t.v[1], t.v[2], &tmin, &tmax);
        if (tmin.x > lmax.x || tmin.y > lmax.y || tmin.z > lmax.z) continue;
        if (lmin.x > tmax.x || lmin.y > tmax.y || lmin.z > tmax.z) continue;
        ContactData out[2];
        int n = 0, region_sphere = VR_ABC;
        bool should_correct = false;
        MeshTester* tester = nullptr;
        if (sphere) {
            n = triangle_sphere_candidates(cvx.margin, t, &out[0], &region_sphere);
        } else {
            for (MeshTester& m : *testers)
                if (m.ia == ix[0] && m.ib == ix[1] && m.ic == ix[2]) { tester = &m; break; }
            if (!tester) {
                MeshTester m;
                m.ia = ix[0];
                m.ib = ix[1];
                m.ic = ix[2];
                m.st.state = TT_PLANE;
                m.st.escapes = 0;
                m.st.axis = V3{0, 0, 0};
                testers->push_back(m);
                tester = &testers->back();
            }
            tester->updated = true;
            n = triangle_convex_candidates(cvx, A.min_radius, w->s.core_shape_scaling, t, &tester->st, out);
            should_correct = tester->st.state == TT_DEEP;
        }
        for (int j = 0; j < n; j++) {
            if (improve) {
                const int region = sphere ? region_sphere : triangle_convex_region(t, out[j]);
                if (mesh_analyze(region, ix, t, should_correct, out[j], sets)) {
                    guaranteed++;
                    add_local(out[j]);
                }
            } else {
                add_local(out[j]);
            }
        }
    }

    if (improve) {
        // Edge contacts fighting each other in one plane get the triangle normal
        // (TriangleMeshConvexContactManifold.cs:241-316).
        if (sets.vertices.empty() && guaranteed == 0 && sets.edges.size() > 1) {
            bool same_plane = true, against = false;
            V3 first = sets.edges[0].c.normal;
            sets.edges[0].corrected = normalize(sets.edges[0].corrected);
            float d = dot(first, sets.edges[0].corrected);
            if (std::fabs(d) > .01f) {
                same_plane = false;
            } else {
                for (size_t i = 1; i < sets.edges.size(); i++) {
                    d = dot(sets.edges[i].c.normal, first);
                    if (d < 0) against = true;
                    d = dot(sets.edges[i].c.normal, sets.edges[0].corrected);
                    if (std::fabs(d) > .01f) {
                        same_plane = false;
                        break;
                    }
                }
            }
            if (same_plane && against) {
                sets.edges[0].c.normal = sets.edges[0].corrected;
                sets.edges[0].should_correct = true;
                for (size_t i = 1; i < sets.edges.size(); i++) {
                    sets.edges[i].corrected = normalize(sets.edges[i].corrected);
                    d = dot(sets.edges[i].corrected, sets.edges[i].c.normal);
                    if (d < .01) {
                        sets.edges[i].c.normal = sets.edges[i].corrected;
                        sets.edges[i].should_correct = true;
                    }
                }
            }
        }
        // Unblocked boundary contacts as they are; blocked ones with their
        // normal corrected when allowed (TriangleMeshConvexContactManifold.cs:322-376).
        auto emit = [&](BoundaryContact& bc, bool blocked) {
            if (!blocked) {
                add_local(bc.c);
            } else if (bc.should_correct || guaranteed == 0) {
                bc.corrected = normalize(bc.corrected);
                float d = dot(bc.corrected, bc.c.normal);
                bc.c.normal = bc.corrected;
                bc.c.depth *= maxf(0, d);
                add_local(bc.c);
            }
        };
        for (BoundaryContact& bc : sets.edges) emit(bc, sets.has_edge(bc.e0, bc.e1));
        for (BoundaryContact& bc : sets.vertices) emit(bc, sets.has_vertex(bc.e0));
    }

    // Stale pair testers (TriangleMeshConvexContactManifold.cs:387-401).
    int stale = 0;
    if (testers) {
        for (int i = (int)testers->size() - 1; i >= 0; --i) {
            if (!(*testers)[i].updated) {
                ++stale;
                (*testers)[i] = testers->back();
                testers->pop_back();
            } else {
                (*testers)[i].updated = false;
            }
        }
    }

    {
        const int sizes[6] = {(int)cand.size(), (int)sets.edges.size(), (int)sets.vertices.size(),
                              (int)sets.blocked_vertices.size(), (int)sets.blocked_edges.size(),
                              testers ? (int)testers->size() + stale : 0};
        std::lock_guard<std::mutex> g(w->mesh_mu);
        for (int k = 0; k < 6; ++k) w->mesh_peak[k] = std::max(w->mesh_peak[k], sizes[k]);
    }
    if (p.ccount + (int)cand.size() > 4) {
        std::vector<int> remove;
        std::vector<ContactData> keep;
        reduce_contacts_general(p, cand, &remove, &keep);
        for (int i = (int)remove.size() - 1; i >= 0; i--) manifold_remove(p, remove[i]);
        for (int i = (int)keep.size() - 1; i >= 0; i--) manifold_add_general(p, keep[i], ta, kMeshTransform);
    } else {
        for (const ContactData& c : cand) manifold_add_general(p, c, ta, kMeshTransform);
    }
}

// StaticMeshPairHandler.UpdateTimeOfImpact (StaticMeshPairHandler.cs:95-159):
// for an active continuous convex, its core sphere swept along its motion
// against each triangle the manifold overlapped this step; a hit further than
// Toolbox.BigEpsilon (on a one-sided mesh, only from the side the mesh keeps)
// shortens the time of impact.
void mesh_time_of_impact(refcpu_world* w, Pair& p) {
    const Body& A = w->bodies[p.a];
    const Body& M = w->bodies[p.b];
    if (!(body_active(A) && A.pmode == 2)) return;
    V3 velocity = mul(A.v, w->s.dt);
    float vsq = length_sq(velocity);
    const float minimumRadius = A.min_radius * w->s.core_shape_scaling;
    p.toi = 1;
    if (!(minimumRadius * minimumRadius < vsq)) return;
    const bool terrain = M.terrain >= 0;
    const TerrainShapeRef* ts = terrain ? &w->terrains[M.terrain].shape : nullptr;
    std::vector<int> tris;
    if (terrain) {
        // The overlaps the manifold's last update found (none before its first).
        if (p.tq_valid) terrain_overlaps(*ts, p.tq_min, p.tq_max, &tris);
    }
    else mesh_overlaps(w->mesh_shapes[M.mesh], A.bmin, A.bmax, &tris);
    const RT ta = ccd_transform(A);
    // TerrainPairHandler.UpdateTimeOfImpact (TerrainPairHandler.cs:117-176):
    // the terrain's up direction decides which side of a triangle counts.
    const V3 terrain_up = terrain ? V3{ts->xf.m.m21, ts->xf.m.m22, ts->xf.m.m23} : V3{0, 0, 0};
    for (int tri : tris) {
        TriangleRef t;
        if (terrain) terrain_world_triangle(*ts, tri, t.v);
        else mesh_world_triangle(w->mesh_shapes[M.mesh], tri, t.v);
        for (int k = 0; k < 3; ++k) t.v[k] = sub(t.v[k], ta.pos);
        t.sidedness = M.sidedness;
        t.margin = 0;
        float hitT;
        V3 hitN;
        if (ccd_sphere_cast_normal(V3{0, 0, 0}, velocity, minimumRadius, triangle_shape(t), kMeshTransform, p.toi,
                                   w->s.core_shape_scaling, &hitT, &hitN) &&
            hitT > 1e-5f) {
            if (terrain) {
                V3 n = cross(sub(t.v[2], t.v[0]), sub(t.v[1], t.v[0]));
                float d = dot(n, terrain_up);
                if (d < 0) d = dot(n, hitN);
                else d = -dot(n, hitN);
                if (d < 0) p.toi = hitT;
            } else if (M.sidedness != SIDED_DOUBLE) {
                V3 n = cross(sub(t.v[1], t.v[0]), sub(t.v[2], t.v[0]));
                float d = dot(n, hitN);
                if ((M.sidedness == SIDED_COUNTERCLOCKWISE && d < 0) || (M.sidedness == SIDED_CLOCKWISE && d > 0))
                    p.toi = hitT;
            } else {
                p.toi = hitT;
            }
        }
    }
}

Pair make_pair_state(refcpu_world* w, uint32_t i, uint32_t j) {
    Pair p;
    std::memset(&p, 0, sizeof(p));
    const Body& bi = w->bodies[i];
    const Body& bj = w->bodies[j];
    int si = bi.shape, sj = bj.shape;
    p.mesh_slot = -1;
    if (is_static(bi) || is_static(bj)) {
        // StaticMeshPairHandler.Initialize (StaticMeshPairHandler.cs:57-79),
        // TerrainPairHandler.Initialize (TerrainPairHandler.cs:70-93): A is the convex.
        p.a = is_static(bi) ? j : i;
        p.b = is_static(bi) ? i : j;
        if (w->bodies[p.a].shape == BEPU_SHAPE_SPHERE) {
            p.kind = BEPU_PAIR_MESH_SPHERE;
        } else {
            p.kind = BEPU_PAIR_MESH_CONVEX;
            std::lock_guard<std::mutex> g(w->mesh_mu);
            if (!w->mesh_free.empty()) {
                p.mesh_slot = w->mesh_free.back();
                w->mesh_free.pop_back();
            } else {
                p.mesh_slot = (int)w->mesh_testers.size();
                w->mesh_testers.emplace_back();
            }
            w->mesh_testers[p.mesh_slot].clear();
        }
    } else if (si == BEPU_SHAPE_BOX && sj == BEPU_SHAPE_BOX) { p.kind = BEPU_PAIR_BOX_BOX; p.a = i; p.b = j; }
    else if (si == BEPU_SHAPE_SPHERE && sj == BEPU_SHAPE_SPHERE) { p.kind = BEPU_PAIR_SPHERE_SPHERE; p.a = i; p.b = j; }
    else if (si == BEPU_SHAPE_BOX && sj == BEPU_SHAPE_SPHERE) { p.kind = BEPU_PAIR_BOX_SPHERE; p.a = i; p.b = j; }
    else if (si == BEPU_SHAPE_SPHERE && sj == BEPU_SHAPE_BOX) { p.kind = BEPU_PAIR_BOX_SPHERE; p.a = j; p.b = i; }
    else { p.kind = BEPU_PAIR_CONVEX; p.a = i; p.b = j; }
    // MaterialManager.DefaultMaterialBlender (MaterialManager.cs:77-85)
    const bepu_material& ma = w->bodies[p.a].mat;
    const bepu_material& mb = w->bodies[p.b].mat;
    p.bounce = ma.bounciness * mb.bounciness;
    p.kinetic = ma.kinetic_friction * mb.kinetic_friction;
    p.static_f = ma.static_friction * mb.static_friction;
    p.toi = 1;
    return p;
}

// ---------------------------------------------------------------------------
// Solve order.  Constraints conflict when they share a dynamic body.  The
// colour of a constraint is the smallest colour not used by its conflicting
// neighbours of higher priority, processing constraints in decreasing
// priority; priority(key) = mix64(key), a bijection.  Within a colour no two
// constraints share a dynamic body, so any order gives identical results; the
// canonical order is ascending key.  The solve order is (colour, key).

uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// A constraint of the solver: a contact manifold constraint (a pair with
// contacts) or a joint.
struct CRef {
    Pair* p;
    JointC* j;
    uint32_t a() const { return p ? p->a : j->a; }
    uint32_t b() const { return p ? p->b : j->b; }
    uint64_t key() const { return p ? pair_key(p->a, p->b) : (kJointKey | j->local); }
    int& color() const { return p ? p->color : j->color; }
};

void colour_constraints(refcpu_world* w, std::vector<CRef>& cons) {
    std::vector<std::pair<uint64_t, CRef>> pri;
    pri.reserve(cons.size());
    for (const CRef& c : cons) pri.push_back({mix64(c.key()), c});
    std::sort(pri.begin(), pri.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    // Per dynamic body: the colours already taken, as a bitmask for the first
    // 64 and a list beyond.
    std::vector<uint64_t> mask(w->bodies.size(), 0);
    std::map<uint32_t, std::vector<int>> more;
    int ncol = 0;
    for (auto& e : pri) {
        const CRef& cr = e.second;
        const uint32_t ia = cr.a(), ib = cr.b();
        bool da = w->bodies[ia].dynamic, db = w->bodies[ib].dynamic;
        uint64_t m = (da ? mask[ia] : 0) | (db ? mask[ib] : 0);
        int c;
        if (~m) {
            c = __builtin_ctzll(~m);
        } else {
            std::vector<char> taken;
            for (uint32_t bi : {ia, ib}) {
                if (!w->bodies[bi].dynamic) continue;
                auto it = more.find(bi);
                if (it == more.end()) continue;
                for (int k : it->second) {
                    if ((size_t)(k - 64) >= taken.size()) taken.resize(k - 63, 0);
                    taken[k - 64] = 1;
                }
            }
            c = 64;
            while ((size_t)(c - 64) < taken.size() && taken[c - 64]) ++c;
        }
        cr.color() = c;
        if (c + 1 > ncol) ncol = c + 1;
        for (uint32_t bi : {ia, ib}) {
            if (!w->bodies[bi].dynamic) continue;
            if (c < 64) mask[bi] |= 1ull << c;
            else more[bi].push_back(c);
        }
    }
    std::sort(cons.begin(), cons.end(), [](const CRef& x, const CRef& y) {
        if (x.color() != y.color()) return x.color() < y.color();
        return x.key() < y.key();
    });
    w->ncolors = ncol;
}

// ---------------------------------------------------------------------------
// Solver rows

inline void apply_linear(Body& b, const float l[3]) {  // Entity.ApplyLinearImpulse (EntityBase.cs:754-760)
    b.v.x += l[0] * b.inv_mass;
    b.v.y += l[1] * b.inv_mass;
    b.v.z += l[2] * b.inv_mass;
}
inline void apply_angular(Body& b, const float i[3]) {  // Entity.ApplyAngularImpulse (EntityBase.cs:769-788)
    const M3& t = b.inertia_inv;
    b.w.x += i[0] * t.m11 + i[1] * t.m21 + i[2] * t.m31;
    b.w.y += i[0] * t.m12 + i[1] * t.m22 + i[2] * t.m32;
    b.w.z += i[0] * t.m13 + i[1] * t.m23 + i[2] * t.m33;
}

// ContactPenetrationConstraint.RelativeVelocity (ContactPenetrationConstraint.cs:86-104)
inline float pen_relvel(const PenC& c, const Body& A, const Body& B) {
    float lambda = A.v.x * c.lin[0] + A.v.y * c.lin[1] + A.v.z * c.lin[2] + A.w.x * c.angA[0] + A.w.y * c.angA[1] +
                   A.w.z * c.angA[2];
    lambda += -B.v.x * c.lin[0] - B.v.y * c.lin[1] - B.v.z * c.lin[2] + B.w.x * c.angB[0] + B.w.y * c.angB[1] +
              B.w.z * c.angB[2];
    return lambda;
}

// ContactPenetrationConstraint.Update (ContactPenetrationConstraint.cs:114-233)
void pen_update(refcpu_world* w, Pair& p, PenC& c, Body& A, Body& B) {
    const bepu_settings& s = w->s;
    const ContactObj& ct = p.slot[c.slot];
    c.aDyn = A.dynamic;
    c.bDyn = B.dynamic;
    c.lin[0] = -ct.normal.x;
    c.lin[1] = -ct.normal.y;
    c.lin[2] = -ct.normal.z;
    c.ra = sub(ct.pos, A.pos);
    c.angA[0] = (c.ra.y * c.lin[2]) - (c.ra.z * c.lin[1]);
    c.angA[1] = (c.ra.z * c.lin[0]) - (c.ra.x * c.lin[2]);
    c.angA[2] = (c.ra.x * c.lin[1]) - (c.ra.y * c.lin[0]);
    c.rb = sub(ct.pos, B.pos);
    c.angB[0] = (c.lin[1] * c.rb.z) - (c.lin[2] * c.rb.y);
    c.angB[1] = (c.lin[2] * c.rb.x) - (c.lin[0] * c.rb.z);
    c.angB[2] = (c.lin[0] * c.rb.y) - (c.lin[1] * c.rb.x);
    float entryA, entryB, tX, tY, tZ;
    if (c.aDyn) {
        const M3& I = A.inertia_inv;
        tX = c.angA[0] * I.m11 + c.angA[1] * I.m21 + c.angA[2] * I.m31;
        tY = c.angA[0] * I.m12 + c.angA[1] * I.m22 + c.angA[2] * I.m32;
        tZ = c.angA[0] * I.m13 + c.angA[1] * I.m23 + c.angA[2] * I.m33;
        entryA = tX * c.angA[0] + tY * c.angA[1] + tZ * c.angA[2] + A.inv_mass;
    } else entryA = 0;
    if (c.bDyn) {
        const M3& I = B.inertia_inv;
        tX = c.angB[0] * I.m11 + c.angB[1] * I.m21 + c.angB[2] * I.m31;
        tY = c.angB[0] * I.m12 + c.angB[1] * I.m22 + c.angB[2] * I.m32;
        tZ = c.angB[0] * I.m13 + c.angB[1] * I.m23 + c.angB[2] * I.m33;
        entryB = tX * c.angB[0] + tY * c.angB[1] + tZ * c.angB[2] + B.inv_mass;
    } else entryB = 0;
    float emi = entryA + entryB;
    float updateRate = 1 / s.dt;
    c.softness = s.softness * emi * updateRate;
    c.v2i = -1 / (c.softness + emi);
    if (ct.depth >= 0) {
        c.bias = minf(maxf(0, ct.depth - s.allowed_penetration) * s.penetration_recovery_stiffness * updateRate,
                      s.max_penetration_recovery_speed);
        if (p.bounce > 0) {
            float bounceVelocity = -pen_relvel(c, A, B);
            if (bounceVelocity > 0) {
                float lowThreshold = s.bounce_velocity_threshold * 0.3f;
                float velocityFraction = clampf(
                    (bounceVelocity - lowThreshold) / (s.bounce_velocity_threshold - lowThreshold + 1e-7f), 0, 1);
                float bouncinessVelocity = velocityFraction * bounceVelocity * p.bounce;
                c.bias = maxf(bouncinessVelocity, c.bias);
            }
        }
    } else {
        c.bias = ct.depth * updateRate;
    }
}

// ContactPenetrationConstraint.ExclusiveUpdate (ContactPenetrationConstraint.cs:240-271)
void pen_warm(const PenC& c, Body& A, Body& B) {
    float lin[3] = {c.acc * c.lin[0], c.acc * c.lin[1], c.acc * c.lin[2]};
    float ang[3];
    if (c.aDyn) {
        ang[0] = c.acc * c.angA[0]; ang[1] = c.acc * c.angA[1]; ang[2] = c.acc * c.angA[2];
        apply_linear(A, lin);
        apply_angular(A, ang);
    }
    if (c.bDyn) {
        lin[0] = -lin[0]; lin[1] = -lin[1]; lin[2] = -lin[2];
        ang[0] = c.acc * c.angB[0]; ang[1] = c.acc * c.angB[1]; ang[2] = c.acc * c.angB[2];
        apply_linear(B, lin);
        apply_angular(B, ang);
    }
}

// ContactPenetrationConstraint.SolveIteration (ContactPenetrationConstraint.cs:278-321)
float pen_solve(PenC& c, Body& A, Body& B) {
    float lambda = (pen_relvel(c, A, B) - c.bias + c.softness * c.acc) * c.v2i;
    float prev = c.acc;
    c.acc = maxf(0, c.acc + lambda);
    lambda = c.acc - prev;
    float lin[3] = {lambda * c.lin[0], lambda * c.lin[1], lambda * c.lin[2]};
    float ang[3];
    if (c.aDyn) {
        ang[0] = lambda * c.angA[0]; ang[1] = lambda * c.angA[1]; ang[2] = lambda * c.angA[2];
        apply_linear(A, lin);
        apply_angular(A, ang);
    }
    if (c.bDyn) {
        lin[0] = -lin[0]; lin[1] = -lin[1]; lin[2] = -lin[2];
        ang[0] = lambda * c.angB[0]; ang[1] = lambda * c.angB[1]; ang[2] = lambda * c.angB[2];
        apply_linear(B, lin);
        apply_angular(B, ang);
    }
    return std::fabs(lambda);
}

// SlidingFrictionTwoAxis.Update (SlidingFrictionTwoAxis.cs:242-440)
void sliding_update(refcpu_world* w, Pair& p, Body& A, Body& B) {
    Sliding& sl = p.sl;
    sl.aDyn = A.dynamic;
    sl.bDyn = B.dynamic;
    sl.count = p.npen;
    const V3& c0 = p.slot[p.pen[0].slot].pos;
    switch (sl.count) {
    case 1: sl.center = c0; break;
    case 2:
        sl.center = add(c0, p.slot[p.pen[1].slot].pos);
        sl.center.x *= .5f; sl.center.y *= .5f; sl.center.z *= .5f;
        break;
    case 3:
        sl.center = add(c0, p.slot[p.pen[1].slot].pos);
        sl.center = add(p.slot[p.pen[2].slot].pos, sl.center);
        sl.center.x *= .333333333f; sl.center.y *= .333333333f; sl.center.z *= .333333333f;
        break;
    case 4:
        sl.center = add(c0, p.slot[p.pen[1].slot].pos);
        sl.center = add(p.slot[p.pen[2].slot].pos, sl.center);
        sl.center = add(p.slot[p.pen[3].slot].pos, sl.center);
        sl.center.x *= .25f; sl.center.y *= .25f; sl.center.z *= .25f;
        break;
    default: sl.center = V3{-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f}; break;
    }
    sl.ra = sub(sl.center, A.pos);
    V3 velA = add(cross(A.w, sl.ra), A.v);
    sl.rb = sub(sl.center, B.pos);
    V3 velB = add(cross(B.w, sl.rb), B.v);
    sl.relvel = sub(velA, velB);
    V3 normal = p.slot[p.pen[0].slot].normal;
    float nvs = normal.x * sl.relvel.x + normal.y * sl.relvel.y + normal.z * sl.relvel.z;
    sl.relvel.x -= nvs * normal.x;
    sl.relvel.y -= nvs * normal.y;
    sl.relvel.z -= nvs * normal.z;
    float length = length_sq(sl.relvel);
    const bepu_settings& s = w->s;
    if (length > 1e-7f) {
        length = std::sqrt(length);
        float inv = 1 / length;
        sl.lin[0] = sl.relvel.x * inv;
        sl.lin[1] = sl.relvel.y * inv;
        sl.lin[2] = sl.relvel.z * inv;
        sl.friction = length > s.static_friction_velocity_threshold ? p.kinetic : p.static_f;
    } else {
        sl.friction = p.static_f;
        if (!(sl.lin[0] != 0 || sl.lin[1] != 0 || sl.lin[2] != 0)) {
            V3 axis1 = cross(normal, V3{1, 0, 0});
            length = length_sq(axis1);
            if (length > 1e-7f) {
                length = std::sqrt(length);
                float inv = 1 / length;
                sl.lin[0] = axis1.x * inv;
                sl.lin[1] = axis1.y * inv;
                sl.lin[2] = axis1.z * inv;
            } else {
                axis1 = normalize(cross(normal, V3{0, 1, 0}));
                sl.lin[0] = axis1.x;
                sl.lin[1] = axis1.y;
                sl.lin[2] = axis1.z;
            }
        }
    }
    float* L = sl.lin;
    L[3] = (L[1] * normal.z) - (L[2] * normal.y);
    L[4] = (L[2] * normal.x) - (L[0] * normal.z);
    L[5] = (L[0] * normal.y) - (L[1] * normal.x);
    const V3& ra = sl.ra;
    const V3& rb = sl.rb;
    float* aA = sl.angA;
    float* aB = sl.angB;
    aA[0] = (ra.y * L[2]) - (ra.z * L[1]);
    aA[1] = (ra.z * L[0]) - (ra.x * L[2]);
    aA[2] = (ra.x * L[1]) - (ra.y * L[0]);
    aA[3] = (ra.y * L[5]) - (ra.z * L[4]);
    aA[4] = (ra.z * L[3]) - (ra.x * L[5]);
    aA[5] = (ra.x * L[4]) - (ra.y * L[3]);
    aB[0] = (L[1] * rb.z) - (L[2] * rb.y);
    aB[1] = (L[2] * rb.x) - (L[0] * rb.z);
    aB[2] = (L[0] * rb.y) - (L[1] * rb.x);
    aB[3] = (L[4] * rb.z) - (L[5] * rb.y);
    aB[4] = (L[5] * rb.x) - (L[3] * rb.z);
    aB[5] = (L[3] * rb.y) - (L[4] * rb.x);
    // entry = J * I^-1 * J^T (+ m^-1 on the diagonal): Matrix2x3.Multiply, Transpose,
    // Matrix2x2.Multiply (Matrix2x3.cs:94-111,225-235; Matrix2x2.cs:245-252)
    auto entry = [](const float* J, const M3& I, float invMass, float e[4]) {
        float t[6];
        t[0] = J[0] * I.m11 + J[1] * I.m21 + J[2] * I.m31;
        t[1] = J[0] * I.m12 + J[1] * I.m22 + J[2] * I.m32;
        t[2] = J[0] * I.m13 + J[1] * I.m23 + J[2] * I.m33;
        t[3] = J[3] * I.m11 + J[4] * I.m21 + J[5] * I.m31;
        t[4] = J[3] * I.m12 + J[4] * I.m22 + J[5] * I.m32;
        t[5] = J[3] * I.m13 + J[4] * I.m23 + J[5] * I.m33;
        // transpose of J: M11=J11 M12=J21 M21=J12 M22=J22 M31=J13 M32=J23
        e[0] = t[0] * J[0] + t[1] * J[1] + t[2] * J[2];
        e[1] = t[0] * J[3] + t[1] * J[4] + t[2] * J[5];
        e[2] = t[3] * J[0] + t[4] * J[1] + t[5] * J[2];
        e[3] = t[3] * J[3] + t[4] * J[4] + t[5] * J[5];
        e[0] += invMass;
        e[3] += invMass;
    };
    float eA[4] = {0, 0, 0, 0}, eB[4] = {0, 0, 0, 0};
    if (sl.aDyn) entry(aA, A.inertia_inv, A.inv_mass, eA);
    if (sl.bDyn) entry(aB, B.inertia_inv, B.inv_mass, eB);
    float m11 = -eA[0] - eB[0];
    float m12 = -eA[1] - eB[1];
    float m21 = -eA[2] - eB[2];
    float m22 = -eA[3] - eB[3];
    // Matrix2x2.Invert (Matrix2x2.cs:160-174)
    float di = 1 / (m11 * m22 - m12 * m21);
    sl.v2i[0] = m22 * di;
    sl.v2i[1] = -m12 * di;
    sl.v2i[2] = -m21 * di;
    sl.v2i[3] = m11 * di;
}

inline void sliding_apply(const Sliding& sl, float lx, float ly, Body& A, Body& B) {
    const float* L = sl.lin;
    float lin[3] = {lx * L[0] + ly * L[3], lx * L[1] + ly * L[4], lx * L[2] + ly * L[5]};
    float ang[3];
    if (sl.aDyn) {
        ang[0] = lx * sl.angA[0] + ly * sl.angA[3];
        ang[1] = lx * sl.angA[1] + ly * sl.angA[4];
        ang[2] = lx * sl.angA[2] + ly * sl.angA[5];
        apply_linear(A, lin);
        apply_angular(A, ang);
    }
    if (sl.bDyn) {
        lin[0] = -lin[0]; lin[1] = -lin[1]; lin[2] = -lin[2];
        ang[0] = lx * sl.angB[0] + ly * sl.angB[3];
        ang[1] = lx * sl.angB[1] + ly * sl.angB[4];
        ang[2] = lx * sl.angB[2] + ly * sl.angB[5];
        apply_linear(B, lin);
        apply_angular(B, ang);
    }
}

// SlidingFrictionTwoAxis.SolveIteration (SlidingFrictionTwoAxis.cs:155-234)
float sliding_solve(Pair& p, Body& A, Body& B) {
    Sliding& sl = p.sl;
    float dvx = A.v.x + (A.w.y * sl.ra.z) - (A.w.z * sl.ra.y);
    float dvy = A.v.y + (A.w.z * sl.ra.x) - (A.w.x * sl.ra.z);
    float dvz = A.v.z + (A.w.x * sl.ra.y) - (A.w.y * sl.ra.x);
    dvx += -B.v.x - (B.w.y * sl.rb.z) + (B.w.z * sl.rb.y);
    dvy += -B.v.y - (B.w.z * sl.rb.x) + (B.w.x * sl.rb.z);
    dvz += -B.v.z - (B.w.x * sl.rb.y) + (B.w.y * sl.rb.x);
    const float* L = sl.lin;
    float lx = dvx * L[0] + dvy * L[1] + dvz * L[2];
    float ly = dvx * L[3] + dvy * L[4] + dvz * L[5];
    float x = lx;
    lx = x * sl.v2i[0] + ly * sl.v2i[2];
    ly = x * sl.v2i[1] + ly * sl.v2i[3];
    float px = sl.acc[0], py = sl.acc[1];
    sl.acc[0] += lx;
    sl.acc[1] += ly;
    float length = sl.acc[0] * sl.acc[0] + sl.acc[1] * sl.acc[1];  // Vector2.LengthSquared
    float maxF = 0;
    for (int i = 0; i < sl.count; i++) maxF += p.pen[i].acc;
    maxF *= sl.friction;
    if (length > maxF * maxF) {
        length = maxF / std::sqrt(length);
        sl.acc[0] *= length;
        sl.acc[1] *= length;
    }
    lx = sl.acc[0] - px;
    ly = sl.acc[1] - py;
    sliding_apply(sl, lx, ly, A, B);
    return std::fabs(lx) + std::fabs(ly);
}

// TwistFrictionConstraint.RelativeVelocity (TwistFrictionConstraint.cs:46-56)
inline float twist_relvel(const Twist& t, const Body& A, const Body& B) {
    float lambda = A.w.x * t.ang[0] + A.w.y * t.ang[1] + A.w.z * t.ang[2];
    lambda -= B.w.x * t.ang[0] + B.w.y * t.ang[1] + B.w.z * t.ang[2];
    return lambda;
}

// TwistFrictionConstraint.Update (TwistFrictionConstraint.cs:114-175)
void twist_update(refcpu_world* w, Pair& p, Body& A, Body& B) {
    Twist& t = p.tw;
    t.aDyn = A.dynamic;
    t.bDyn = B.dynamic;
    V3 normal = p.slot[p.pen[0].slot].normal;
    t.ang[0] = normal.x; t.ang[1] = normal.y; t.ang[2] = normal.z;
    float entryA, entryB, tX, tY, tZ;
    if (t.aDyn) {
        const M3& I = A.inertia_inv;
        tX = t.ang[0] * I.m11 + t.ang[1] * I.m21 + t.ang[2] * I.m31;
        tY = t.ang[0] * I.m12 + t.ang[1] * I.m22 + t.ang[2] * I.m32;
        tZ = t.ang[0] * I.m13 + t.ang[1] * I.m23 + t.ang[2] * I.m33;
        entryA = tX * t.ang[0] + tY * t.ang[1] + tZ * t.ang[2] + A.inv_mass;
    } else entryA = 0;
    if (t.bDyn) {
        const M3& I = B.inertia_inv;
        tX = t.ang[0] * I.m11 + t.ang[1] * I.m21 + t.ang[2] * I.m31;
        tY = t.ang[0] * I.m12 + t.ang[1] * I.m22 + t.ang[2] * I.m32;
        tZ = t.ang[0] * I.m13 + t.ang[1] * I.m23 + t.ang[2] * I.m33;
        entryB = tX * t.ang[0] + tY * t.ang[1] + tZ * t.ang[2] + B.inv_mass;
    } else entryB = 0;
    t.v2i = -1 / (entryA + entryB);
    float rav = twist_relvel(t, A, B);
    const V3& rs = p.sl.relvel;
    const bepu_settings& s = w->s;
    t.friction = std::fabs(rav) > s.static_friction_velocity_threshold ||
                         std::fabs(rs.x) + std::fabs(rs.y) + std::fabs(rs.z) > s.static_friction_velocity_threshold
                     ? p.kinetic
                     : p.static_f;
    t.friction *= s.twist_friction_factor;
    t.count = p.npen;
    for (int i = 0; i < t.count; i++) t.lever[i] = length(sub(p.slot[p.pen[i].slot].pos, p.sl.center));
}

inline void twist_apply(const Twist& t, float lambda, Body& A, Body& B) {
    float ang[3] = {lambda * t.ang[0], lambda * t.ang[1], lambda * t.ang[2]};
    if (t.aDyn) apply_angular(A, ang);
    if (t.bDyn) {
        ang[0] = -ang[0]; ang[1] = -ang[1]; ang[2] = -ang[2];
        apply_angular(B, ang);
    }
}

// TwistFrictionConstraint.SolveIteration (TwistFrictionConstraint.cs:65-107)
float twist_solve(Pair& p, Body& A, Body& B) {
    Twist& t = p.tw;
    float lambda = twist_relvel(t, A, B);
    lambda *= t.v2i;
    float prev = t.acc;
    float maxF = 0;
    for (int i = 0; i < t.count; i++) maxF += t.lever[i] * p.pen[i].acc;
    maxF *= t.friction;
    t.acc = clampf(t.acc + lambda, -maxF, maxF);
    lambda = t.acc - prev;
    twist_apply(t, lambda, A, B);
    return std::fabs(lambda);
}

// ContactFrictionConstraint.RelativeVelocity (ContactFrictionConstraint.cs:114-126)
inline float fric_relvel(const FricC& f, const Body& A, const Body& B) {
    float velocity = 0;
    velocity += A.v.x * f.lin[0] + A.v.y * f.lin[1] + A.v.z * f.lin[2] + A.w.x * f.angA[0] + A.w.y * f.angA[1] +
                A.w.z * f.angA[2];
    velocity += -B.v.x * f.lin[0] - B.v.y * f.lin[1] - B.v.z * f.lin[2] + B.w.x * f.angB[0] + B.w.y * f.angB[1] +
                B.w.z * f.angB[2];
    return velocity;
}

// ContactFrictionConstraint.Update (ContactFrictionConstraint.cs:180-287)
void fric_update(refcpu_world* w, Pair& p, FricC& f, const PenC& pc, Body& A, Body& B) {
    f.aDyn = A.dynamic;
    f.bDyn = B.dynamic;
    V3 ra = pc.ra, rb = pc.rb;
    V3 velA = add(cross(A.w, ra), A.v);
    V3 velB = add(cross(B.w, rb), B.v);
    V3 rv = sub(velA, velB);
    V3 normal = p.slot[pc.slot].normal;
    float nvs = normal.x * rv.x + normal.y * rv.y + normal.z * rv.z;
    rv.x -= nvs * normal.x;
    rv.y -= nvs * normal.y;
    rv.z -= nvs * normal.z;
    float length = length_sq(rv);
    if (length > 1e-7f) {
        length = std::sqrt(length);
        f.lin[0] = rv.x / length;
        f.lin[1] = rv.y / length;
        f.lin[2] = rv.z / length;
        f.friction = length > w->s.static_friction_velocity_threshold ? p.kinetic : p.static_f;
    } else {
        if (f.lin[0] != 0 || f.lin[1] != 0 || f.lin[2] != 0) {
            f.friction = p.static_f;
        } else {
            f.ais = false;
            return;
        }
    }
    f.angA[0] = (ra.y * f.lin[2]) - (ra.z * f.lin[1]);
    f.angA[1] = (ra.z * f.lin[0]) - (ra.x * f.lin[2]);
    f.angA[2] = (ra.x * f.lin[1]) - (ra.y * f.lin[0]);
    f.angB[0] = (f.lin[1] * rb.z) - (f.lin[2] * rb.y);
    f.angB[1] = (f.lin[2] * rb.x) - (f.lin[0] * rb.z);
    f.angB[2] = (f.lin[0] * rb.y) - (f.lin[1] * rb.x);
    float entryA, entryB, tX, tY, tZ;
    if (f.aDyn) {
        const M3& I = A.inertia_inv;
        tX = f.angA[0] * I.m11 + f.angA[1] * I.m21 + f.angA[2] * I.m31;
        tY = f.angA[0] * I.m12 + f.angA[1] * I.m22 + f.angA[2] * I.m32;
        tZ = f.angA[0] * I.m13 + f.angA[1] * I.m23 + f.angA[2] * I.m33;
        entryA = tX * f.angA[0] + tY * f.angA[1] + tZ * f.angA[2] + A.inv_mass;
    } else entryA = 0;
    if (f.bDyn) {
        const M3& I = B.inertia_inv;
        tX = f.angB[0] * I.m11 + f.angB[1] * I.m21 + f.angB[2] * I.m31;
        tY = f.angB[0] * I.m12 + f.angB[1] * I.m22 + f.angB[2] * I.m32;
        tZ = f.angB[0] * I.m13 + f.angB[1] * I.m23 + f.angB[2] * I.m33;
        entryB = tX * f.angB[0] + tY * f.angB[1] + tZ * f.angB[2] + B.inv_mass;
    } else entryB = 0;
    f.v2i = -1 / (entryA + entryB);
}

inline void fric_apply(const FricC& f, float lambda, Body& A, Body& B) {
    float lin[3] = {lambda * f.lin[0], lambda * f.lin[1], lambda * f.lin[2]};
    float ang[3];
    if (f.aDyn) {
        ang[0] = lambda * f.angA[0]; ang[1] = lambda * f.angA[1]; ang[2] = lambda * f.angA[2];
        apply_linear(A, lin);
        apply_angular(A, ang);
    }
    if (f.bDyn) {
        lin[0] = -lin[0]; lin[1] = -lin[1]; lin[2] = -lin[2];
        ang[0] = lambda * f.angB[0]; ang[1] = lambda * f.angB[1]; ang[2] = lambda * f.angB[2];
        apply_linear(B, lin);
        apply_angular(B, ang);
    }
}

// ContactFrictionConstraint.SolveIteration (ContactFrictionConstraint.cs:131-174)
float fric_solve(FricC& f, const PenC& pc, Body& A, Body& B) {
    float lambda = fric_relvel(f, A, B) * f.v2i;
    float prev = f.acc;
    float maxForce = f.friction * pc.acc;
    f.acc = clampf(f.acc + lambda, -maxForce, maxForce);
    lambda = f.acc - prev;
    fric_apply(f, lambda, A, B);
    return std::fabs(lambda);
}

// Solver.UnsafePrestep -> UpdateSolverActivity -> group Update
// (Solver.cs:265-275; ContactManifoldConstraint.cs:128-145; SolverGroup.cs:70-76;
//  ConvexContactManifoldConstraint.cs:171-177; NonConvexContactManifoldConstraint.cs:180-186)
void prestep(refcpu_world* w, Pair& p) {
    Body& A = w->bodies[p.a];
    Body& B = w->bodies[p.b];
    p.group_ais = A.dynamic || B.dynamic;  // dynamic bodies are always active here
    if (!p.group_ais) return;
    p.group_cur = 0;
    p.group_zero = 0;
    for (int i = 0; i < p.npen; ++i) {
        PenC& c = p.pen[i];
        c.ais = true;
        c.cur = 0;
        c.zero = 0;
        pen_update(w, p, c, A, B);
    }
    if (convex_kind(p.kind)) {
        p.sl.ais = p.sl.active;
        p.sl.cur = p.sl.zero = 0;
        if (p.sl.ais) sliding_update(w, p, A, B);
        p.tw.ais = p.tw.active;
        p.tw.cur = p.tw.zero = 0;
        if (p.tw.ais) twist_update(w, p, A, B);
    } else {
        for (int i = 0; i < p.nfric; ++i) {
            FricC& f = p.fric[i];
            f.ais = true;
            f.cur = 0;
            f.zero = 0;
            fric_update(w, p, f, p.pen[i], A, B);
        }
    }
}

// Solver.UnsafeExclusiveUpdate -> group ExclusiveUpdate (Solver.cs:277-283)
void warm_start(refcpu_world* w, Pair& p) {
    if (!p.group_ais) return;
    Body& A = w->bodies[p.a];
    Body& B = w->bodies[p.b];
    for (int i = 0; i < p.npen; ++i)
        if (p.pen[i].ais) pen_warm(p.pen[i], A, B);
    if (convex_kind(p.kind)) {
        if (p.sl.ais) sliding_apply(p.sl, p.sl.acc[0], p.sl.acc[1], A, B);
        if (p.tw.ais) twist_apply(p.tw, p.tw.acc, A, B);
    } else {
        for (int i = 0; i < p.nfric; ++i) {
            const FricC& f = p.fric[i];
            if (f.ais) fric_apply(f, f.acc, A, B);
        }
    }
}

// SolverGroup.SolveUpdateable (SolverGroup.cs:117-150)
template <typename F>
inline void solve_sub(refcpu_world* w, bool& ais, int& cur, int& zero, int& active, F solve) {
    if (!ais) return;
    cur++;
    if (cur <= w->s.iteration_limit) {
        if (solve() < w->s.min_impulse) {
            zero++;
            if (zero > w->s.min_iteration_count) ais = false;
            else active++;
        } else {
            zero = 0;
            active++;
        }
    } else {
        ais = false;
    }
}

// Solver.UnsafeSolveIteration -> ContactManifoldConstraint.SolveIteration
// (Solver.cs:285-314; ConvexContactManifoldConstraint.cs:199-212; NonConvex...:201-212)
void solve_iteration(refcpu_world* w, Pair& p) {
    if (!p.group_ais) return;
    p.group_cur++;
    if (p.group_cur > w->s.iteration_limit) {
        p.group_ais = false;
        return;
    }
    Body& A = w->bodies[p.a];
    Body& B = w->bodies[p.b];
    int active = 0;
    for (int i = 0; i < p.npen; ++i) {
        PenC& c = p.pen[i];
        solve_sub(w, c.ais, c.cur, c.zero, active, [&] { return pen_solve(c, A, B); });
    }
    if (convex_kind(p.kind)) {
        solve_sub(w, p.sl.ais, p.sl.cur, p.sl.zero, active, [&] { return sliding_solve(p, A, B); });
        solve_sub(w, p.tw.ais, p.tw.cur, p.tw.zero, active, [&] { return twist_solve(p, A, B); });
    } else {
        for (int i = 0; i < p.nfric; ++i) {
            FricC& f = p.fric[i];
            solve_sub(w, f.ais, f.cur, f.zero, active, [&] { return fric_solve(f, p.pen[i], A, B); });
        }
    }
    p.group_ais = active > 0;
    // The group returns minimumImpulse + 1, so its own zero-impulse counter resets.
    p.group_zero = 0;
}

// ---------------------------------------------------------------------------
// BallSocketJoint (Constraints/TwoEntity/Joints/BallSocketJoint.cs)

// SolverUpdateable.UpdateSolverActivity (SolverUpdateable.cs:202-218): a
// joint is solved while the island of one of its dynamic connections is
// active (a kinematic connection has no island).
bool joint_active(const refcpu_world* w, const JointC& j) {
    if (j.dead) return false;
    const Body& A = w->bodies[j.a];
    const Body& B = w->bodies[j.b];
    return (A.dynamic && !A.asleep) || (B.dynamic && !B.asleep);
}

// Matrix3x3.CreateCrossProduct (BEPUutilities/Matrix3x3.cs:378-389)
inline M3 m3_cross(const V3& v) { return M3{0, -v.z, v.y, v.z, 0, -v.x, -v.y, v.x, 0}; }
// Matrix3x3.Subtract (Matrix3x3.cs:973-997)
inline M3 m3_sub(const M3& a, const M3& b) {
    return M3{a.m11 - b.m11, a.m12 - b.m12, a.m13 - b.m13, a.m21 - b.m21, a.m22 - b.m22,
              a.m23 - b.m23, a.m31 - b.m31, a.m32 - b.m32, a.m33 - b.m33};
}
// Matrix3x3.CreateScale (Matrix3x3.cs:437-440)
inline M3 m3_diag(float s) { return M3{s, 0, 0, 0, s, 0, 0, 0, s}; }

// Solver.UnsafePrestep -> BallSocketJoint.Update (Solver.cs:265-275,
// BallSocketJoint.cs:229-300): world offsets, the softened effective mass
// of the point constraint, and the bias velocity toward closing the gap.
void joint_prestep(refcpu_world* w, JointC& j) {
    j.ais = joint_active(w, j);
    if (!j.ais) return;
    j.cur = 0;
    j.zero = 0;
    const Body& A = w->bodies[j.a];
    const Body& B = w->bodies[j.b];
    const float dt = w->s.dt;
    j.offA = m3_transform(j.localA, A.R);
    j.offB = m3_transform(j.localB, B.R);
    // SpringSettings.ComputeErrorReductionAndSoftness (SpringSettings.cs:109-124), update rate 1/dt
    const float rate = 1 / dt;
    const float multiplier = 1 / (dt * j.stiffness + j.damping);
    const float erp = j.stiffness * multiplier;
    j.softness = rate * multiplier;
    const M3 rA = m3_cross(j.offA), rB = m3_cross(j.offB);
    M3 k;
    if (A.dynamic && B.dynamic) {
        const M3 lin = m3_diag(A.inv_mass + B.inv_mass);
        M3 angA = m3_mul(rA, A.inertia_inv), angB = m3_mul(rB, B.inertia_inv);
        angA = m3_mul(angA, rA);
        angB = m3_mul(angB, rB);
        k = m3_sub(lin, angA);
        k = m3_sub(k, angB);
    } else if (A.dynamic) {
        const M3 lin = m3_diag(A.inv_mass);
        M3 angA = m3_mul(rA, A.inertia_inv);
        angA = m3_mul(angA, rA);
        k = m3_sub(lin, angA);
    } else {
        const M3 lin = m3_diag(B.inv_mass);
        M3 angB = m3_mul(rB, B.inertia_inv);
        angB = m3_mul(angB, rB);
        k = m3_sub(lin, angB);
    }
    k.m11 += j.softness;
    k.m22 += j.softness;
    k.m33 += j.softness;
    j.mass = m3_invert(k);
    V3 err = add(B.pos, j.offB);
    err = sub(err, A.pos);
    err = sub(err, j.offA);
    j.bias = mul(err, -erp);
    const float len = length_sq(j.bias);
    if (len > j.max_corr_sq) {
        const float m = j.max_corr / std::sqrt(len);
        j.bias.x *= m;
        j.bias.y *= m;
        j.bias.z *= m;
    }
}

// The joint's impulse on its connections (BallSocketJoint.cs:308-328, 364-383).
inline void joint_apply(const JointC& j, const V3& impulse, Body& A, Body& B) {
    if (A.dynamic) {
        const V3 l = neg(impulse);
        const float lv[3] = {l.x, l.y, l.z};
        apply_linear(A, lv);
        const V3 t = cross(j.offA, l);
        const float tv[3] = {t.x, t.y, t.z};
        apply_angular(A, tv);
    }
    if (B.dynamic) {
        const float lv[3] = {impulse.x, impulse.y, impulse.z};
        apply_linear(B, lv);
        const V3 t = cross(j.offB, impulse);
        const float tv[3] = {t.x, t.y, t.z};
        apply_angular(B, tv);
    }
}

// Solver.UnsafeExclusiveUpdate -> BallSocketJoint.ExclusiveUpdate (warm start)
void joint_warm(refcpu_world* w, JointC& j) {
    if (!j.ais) return;
    joint_apply(j, j.acc, w->bodies[j.a], w->bodies[j.b]);
}

// Solver.UnsafeSolveIteration (Solver.cs:285-314) -> BallSocketJoint.SolveIteration
// (BallSocketJoint.cs:335-388)
void joint_solve(refcpu_world* w, JointC& j) {
    if (!j.ais) return;
    j.cur++;
    if (j.cur > w->s.iteration_limit) {
        j.ais = false;
        return;
    }
    Body& A = w->bodies[j.a];
    Body& B = w->bodies[j.b];
    const V3 aVel = add(A.v, cross(A.w, j.offA));
    const V3 bVel = add(B.v, cross(B.w, j.offB));
    V3 lambda;
    lambda.x = aVel.x - bVel.x + j.bias.x - j.softness * j.acc.x;
    lambda.y = aVel.y - bVel.y + j.bias.y - j.softness * j.acc.y;
    lambda.z = aVel.z - bVel.z + j.bias.z - j.softness * j.acc.z;
    lambda = m3_transform(lambda, j.mass);
    j.acc = add(j.acc, lambda);
    joint_apply(j, lambda, A, B);
    const float impulse = std::fabs(lambda.x) + std::fabs(lambda.y) + std::fabs(lambda.z);
    if (impulse < w->s.min_impulse) {
        j.zero++;
        if (j.zero > w->s.min_iteration_count) j.ais = false;
    } else {
        j.zero = 0;
    }
}

void con_prestep(refcpu_world* w, const CRef& c) {
    if (c.p) prestep(w, *c.p);
    else joint_prestep(w, *c.j);
}
void con_warm(refcpu_world* w, const CRef& c) {
    if (c.p) warm_start(w, *c.p);
    else joint_warm(w, *c.j);
}
void con_solve(refcpu_world* w, const CRef& c) {
    if (c.p) solve_iteration(w, *c.p);
    else joint_solve(w, *c.j);
}

// ---------------------------------------------------------------------------
// The reference's single-threaded solve order (Solver.UpdateSingleThreaded,
// Solver.cs:237-263; BEPUutilities/PermutationMapper.cs:27-127): prestep and
// exclusive update (warm start) in list order, then iterationLimit x n visits
// in the order (i * prime + offset) mod n, where each step advances the
// permutation index k, prime = primes[k mod |primes|] and offset = prime * k.
// A prime above n makes i -> i * prime mod n a bijection, so every block of n
// visits is one fixed permutation of the list.  Two things cannot follow the
// reference exactly: its list order (the order constraints were added, which
// follows its broad phase's overlap order) is taken here as pair-key order,
// and its prime table is replaced by primes of the same range generated here
// (the smallest prime at or above 350,000,000 + 4,999,999 k, k = 0..127).

int64_t perm_prime(int64_t k) {
    static std::vector<int64_t> table;
    static std::mutex m;
    std::lock_guard<std::mutex> g(m);
    if (table.empty()) {
        for (int64_t j = 0; j < 128; ++j) {
            int64_t c = 350000000 + 4999999 * j;
            for (;; ++c) {
                bool prime = c > 1 && (c % 2 != 0 || c == 2);
                for (int64_t d = 3; prime && d * d <= c; d += 2)
                    if (c % d == 0) prime = false;
                if (prime) break;
            }
            table.push_back(c);
        }
    }
    return table[(size_t)(k % (int64_t)table.size())];
}

// The visit order of one step: entry i of the iterations is list[order(i)].
inline int64_t permuted_index(int64_t i, int64_t n, int64_t k) {
    const int64_t prime = perm_prime(k);
    return (i * prime + prime * k) % n;
}

void permuted_solve(refcpu_world* w, const std::vector<CRef>& cons, bool descending) {
    std::vector<CRef> list(cons);
    std::sort(list.begin(), list.end(), [&](const CRef& a, const CRef& b) {
        const uint64_t ka = a.key(), kb = b.key();
        return descending ? ka > kb : ka < kb;
    });
    for (const CRef& c : list) con_warm(w, c);
    ++w->perm_index;
    const int64_t n = (int64_t)list.size();
    const int64_t total = (int64_t)w->s.iteration_limit * n;
    for (int64_t i = 0; i < total; ++i) con_solve(w, list[(size_t)permuted_index(i, n, w->perm_index)]);
}

// ---------------------------------------------------------------------------
// Activity after the narrow phase

// A pair that gained contacts adds a constraint to the solver, and adding its
// island connection merges the islands and activates the result
// (StandardPairHandler.cs:158-168 -> Solver.Add -> DeactivationManager.Add /
// Merge, DeactivationManager.cs:323-388); a new constraint with an active
// kinematic body activates the island too (a merge with no island).  So an
// island of this step's constraints (frozen ones included) is active when a
// member is awake or a constraint holds an active kinematic body.  Frozen
// pairs whose island became active are stale and removed (RemoveStaleOverlaps,
// NarrowPhase.cs:300-333: not updated, one entry active); the next step
// creates them again.  The islands the step ends with, the labels the next
// deactivation stage uses, are the components of the constraints left.
// A joint between two dynamic bodies joins their islands for as long as it
// lives (its SimulationIslandConnection).
inline bool joint_edge(const refcpu_world* w, const JointC& j) {
    return !j.dead && w->bodies[j.a].dynamic && w->bodies[j.b].dynamic;
}

void activity_pass(refcpu_world* w, PairTable& t) {
    const uint32_t n = (uint32_t)w->bodies.size();
    UnionFind uf(n);
    for (const Pair& p : t.vals)
        if (p.ccount > 0 && w->bodies[p.a].dynamic && w->bodies[p.b].dynamic) uf.unite(p.a, p.b);
    for (const JointC& j : w->joints)
        if (joint_edge(w, j)) uf.unite(j.a, j.b);
    std::vector<uint8_t> act(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const Body& b = w->bodies[i];
        if (b.dynamic && !b.removed && !b.asleep) act[uf.find(i)] = 1;
    }
    auto kinematic_wake = [&](uint32_t a, uint32_t b) {
        const Body& A = w->bodies[a];
        const Body& B = w->bodies[b];
        if (kinematic_active(A) && B.dynamic) act[uf.find(b)] = 1;
        if (kinematic_active(B) && A.dynamic) act[uf.find(a)] = 1;
    };
    for (const Pair& p : t.vals)
        if (p.ccount > 0) kinematic_wake(p.a, p.b);
    for (const JointC& j : w->joints)
        if (!j.dead) kinematic_wake(j.a, j.b);
    for (uint32_t i = 0; i < n; ++i) {
        Body& b = w->bodies[i];
        if (b.dynamic && !b.removed && act[uf.find(i)]) b.asleep = false;
    }
    size_t m = 0;
    for (size_t k = 0; k < t.keys.size(); ++k) {
        const Pair& p = t.vals[k];
        if (p.frozen && (body_active(w->bodies[p.a]) || body_active(w->bodies[p.b]))) continue;
        if (m != k) {
            t.keys[m] = t.keys[k];
            t.vals[m] = t.vals[k];
        }
        ++m;
    }
    t.keys.resize(m);
    t.vals.resize(m);
    UnionFind live(n);
    for (const Pair& p : t.vals)
        if (p.ccount > 0 && w->bodies[p.a].dynamic && w->bodies[p.b].dynamic) live.unite(p.a, p.b);
    for (const JointC& j : w->joints)
        if (joint_edge(w, j)) live.unite(j.a, j.b);
    for (uint32_t i = 0; i < n; ++i) {
        Body& b = w->bodies[i];
        if (b.dynamic && !b.removed) b.label = live.find(i);
    }
}

// A constraint removed this step queues a split attempt, flushed at the next
// deactivation update (DeactivationManager.Remove / FlushSplits,
// DeactivationManager.cs:232-283, 401-417): when its two dynamic bodies are
// no longer connected, both are activated (TryToSplit, DeactivationManager.cs:
// 432-555), which restarts their time below the limit.  Every attempt is
// flushed at once (the reference flushes max(3, 1%) of them per step).
void split_activation(refcpu_world* w, const PairTable& old, const PairTable& now) {
    for (const Pair& p : old.vals) {
        if (p.ccount <= 0) continue;
        Body& A = w->bodies[p.a];
        Body& B = w->bodies[p.b];
        if (!A.dynamic || !B.dynamic || A.removed || B.removed) continue;
        const Pair* q = now.find(pair_key(p.a, p.b));
        if (q && q->ccount > 0) continue;
        if (label_of(w, p.a) != label_of(w, p.b)) {
            activate(A);
            activate(B);
        }
    }
}

// ---------------------------------------------------------------------------
// One time step (Space.DoTimeStep, Space.cs:496-521)

void step_once(refcpu_world* w) {
    // The body stages are per-body independent, the pair updates per-pair
    // independent, and the solver runs colour by colour (no two constraints of
    // a colour share a dynamic body), so the multithreaded step computes the
    // same bits as the single-threaded one.
    deactivation_stage(w);
    par_for(w, w->bodies.size(), 4096, [&](size_t b0, size_t b1) {
        for (size_t i = b0; i < b1; ++i) {
            force_update(w, w->bodies[i]);
            bounding_box(w, w->bodies[i]);
        }
    });
    if (w->broadphase_mode == 1) broadphase_bruteforce(w, &w->overlaps);
    else if (w->broadphase_mode == 2) broadphase_sweep(w, &w->overlaps);
    else broadphase_grid(w, &w->overlaps);

    // NarrowPhase.UpdateSingleThreaded: update (or create) one pair per overlap,
    // then drop pairs whose overlap vanished (NarrowPhase.cs:254-340).  The
    // broad phase reports no overlap of two inactive entries (BroadPhase.cs:
    // 97-102): such a pair is kept as it was, not updated, and none is created.
    std::vector<Pair> updated(w->overlaps.size());
    std::vector<uint8_t> keep(w->overlaps.size(), 1);
    par_for(w, w->overlaps.size(), 256, [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; ++k) {
            uint64_t key = w->overlaps[k];
            const uint32_t ka = (uint32_t)(key >> 32), kb = (uint32_t)(key & 0xffffffffu);
            const Pair* old = w->pairs.find(key);
            Pair& p = updated[k];
            if (!body_active(w->bodies[ka]) && !body_active(w->bodies[kb])) {
                if (old) {
                    p = *old;
                    p.frozen = true;
                } else {
                    keep[k] = 0;
                }
                continue;
            }
            if (old) p = *old;
            else p = make_pair_state(w, ka, kb);
            p.frozen = false;
            if (p.kind == BEPU_PAIR_BOX_BOX) update_box_box(w, p);
            else if (p.kind == BEPU_PAIR_CONVEX) update_general_convex(w, p);
            else if (p.kind == BEPU_PAIR_MESH_CONVEX || p.kind == BEPU_PAIR_MESH_SPHERE) update_mesh(w, p);
            else update_single_contact(w, p);
        }
    });
    PairTable next;
    next.keys.reserve(w->overlaps.size());
    next.vals.reserve(w->overlaps.size());
    for (size_t k = 0; k < w->overlaps.size(); ++k)
        if (keep[k]) {
            next.keys.push_back(w->overlaps[k]);
            next.vals.push_back(std::move(updated[k]));
        }
    activity_pass(w, next);
    split_activation(w, w->pairs, next);
    if (!w->mesh_testers.empty()) {
        // The tester tables of mesh pairs that ended are free again.
        std::vector<char> used(w->mesh_testers.size(), 0);
        for (const Pair& p : next.vals)
            if (p.mesh_slot >= 0) used[p.mesh_slot] = 1;
        w->mesh_free.clear();
        for (int i = (int)used.size() - 1; i >= 0; --i)
            if (!used[i]) w->mesh_free.push_back(i);
    }
    w->pairs = std::move(next);
    w->overlaps = w->pairs.keys;

    // Solver.UpdateSolverActivity (SolverUpdateable.cs:202-218): the
    // constraints of inactive islands are not solved.
    std::vector<CRef> cons;
    float maxpen = 0;
    for (Pair& p : w->pairs.vals) {
        if (p.ccount > 0 && !(w->bodies[p.a].dynamic && w->bodies[p.a].asleep) &&
            !(w->bodies[p.b].dynamic && w->bodies[p.b].asleep))
            cons.push_back(CRef{&p, nullptr});
        for (int i = 0; i < p.ccount; ++i) maxpen = std::max(maxpen, p.slot[p.clist[i]].depth);
    }
    for (JointC& j : w->joints)
        if (joint_active(w, j)) cons.push_back(CRef{nullptr, &j});
    w->max_pen = maxpen;
    colour_constraints(w, cons);
    w->order.clear();
    for (const CRef& c : cons)
        if (c.p) w->order.push_back(c.key());

    // Solver.UpdateSingleThreaded (Solver.cs:239-263) in colour order.
    par_for(w, cons.size(), 512, [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; ++c) con_prestep(w, cons[c]);
    });
    std::vector<size_t> cstart(1, 0);
    for (size_t c = 1; c < cons.size(); ++c)
        if (cons[c].color() != cons[c - 1].color()) cstart.push_back(c);
    cstart.push_back(cons.size());
    auto by_colour = [&](auto&& fn) {
        for (size_t g = 0; g + 1 < cstart.size(); ++g) {
            size_t base = cstart[g];
            par_for(w, cstart[g + 1] - base, 512, [&](size_t c0, size_t c1) {
                for (size_t c = c0; c < c1; ++c) fn(cons[base + c]);
            });
        }
    };
    if (w->solve_order != 0) {
        permuted_solve(w, cons, w->solve_order == 2);
    } else {
        by_colour([&](const CRef& c) { con_warm(w, c); });
        for (int it = 0; it < w->s.iteration_limit; ++it) by_colour([&](const CRef& c) { con_solve(w, c); });
    }

    // ContinuousPositionUpdater.UpdateSingleThreaded (ContinuousPositionUpdater.cs:134-154)
    par_for(w, w->bodies.size(), 4096, [&](size_t b0, size_t b1) {
        for (size_t i = b0; i < b1; ++i) integrate(w, w->bodies[i]);
    });
    if (w->ccd) continuous_update(w);
    w->steps++;
}

}  // namespace

// ---------------------------------------------------------------------------
// C API used by the tests

extern "C" {

void refcpu_settings_default(bepu_settings* s) {
    s->gravity[0] = 0; s->gravity[1] = -9.81f; s->gravity[2] = 0;
    s->dt = 1.0f / 60.0f;
    s->iteration_limit = 10;
    s->min_iteration_count = 1;
    s->min_impulse = .001f;
    s->allowed_penetration = .01f;
    s->default_margin = .04f;
    s->max_contact_distance = .1f;
    s->contact_invalidation_length_sq = .01f;
    s->contact_min_separation_sq = .0009f;
    s->bounce_velocity_threshold = 1;
    s->max_penetration_recovery_speed = 2;
    s->penetration_recovery_stiffness = .2f;
    s->static_friction_velocity_threshold = .2f;
    s->twist_friction_factor = 1;
    s->softness = .001f;
    s->use_stabilization = 1;
    s->velocity_lower_limit = .26f;
    s->low_velocity_time_minimum = 1;
    s->linear_damping = .03f;
    s->angular_damping = .15f;
    s->max_deactivation_attempts = 100;
    s->core_shape_scaling = .8f;
}

refcpu_world* refcpu_create(const bepu_settings* s) {
    refcpu_world* w = new refcpu_world();
    if (s) w->s = *s;
    else refcpu_settings_default(&w->s);
    w->s.core_shape_scaling = clampf(w->s.core_shape_scaling, 0, .99f);  // MotionSettings.cs:22-34
    return w;
}
void refcpu_destroy(refcpu_world* w) { delete w; }
// sizeof of the header structs as this library was compiled, so that a
// binding can check its layout against them: 0 settings, 1 motion state,
// 2 material, 3 contact, 4 manifold, 5 impulses, 6 step stats.
uint32_t refcpu_sizeof(int which) {
    switch (which) {
    case 0: return sizeof(bepu_settings);
    case 1: return sizeof(bepu_motion_state);
    case 2: return sizeof(bepu_material);
    case 3: return sizeof(bepu_contact);
    case 4: return sizeof(bepu_manifold);
    case 5: return sizeof(bepu_impulses);
    case 6: return sizeof(bepu_step_stats);
    }
    return 0;
}
// Threads used by refcpu_step (1: single-threaded).  Results do not depend on it.
void refcpu_set_threads(refcpu_world* w, int n) {
    w->threads = n < 1 ? 1 : n;
    if (w->threads > 1) w->pool.reset(new Pool(w->threads));
    else w->pool.reset();
}
void refcpu_set_broadphase(refcpu_world* w, int mode) { w->broadphase_mode = mode; }

static int finite3(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return 0;
    return 1;
}

int refcpu_add_box(refcpu_world* w, const bepu_motion_state* st, float width, float height, float length,
                   float mass, const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!st || !(width > 0) || !(height > 0) || !(length > 0) || std::isnan(mass) || std::isinf(mass))
        return BEPU_E_INVALID_ARG;
    if (!finite3(st->position, 3) || !finite3(st->orientation, 4)) return BEPU_E_INVALID_ARG;
    Body b{};
    std::memset(&b, 0, sizeof(b));
    b.shape = BEPU_SHAPE_BOX;
    b.margin = w->s.default_margin;
    // BoxShape(width, height, length) (BoxShape.cs:56-62) and ComputeDescription (:118-138)
    b.half = V3{width * 0.5f, height * 0.5f, length * 0.5f};
    float w2 = width * width, h2 = height * height, l2 = length * length;
    const float inv12 = 1 / 12.0f;
    M3 vd{0, 0, 0, 0, 0, 0, 0, 0, 0};
    vd.m11 = (h2 + l2) * inv12;
    vd.m22 = (w2 + l2) * inv12;
    vd.m33 = (w2 + h2) * inv12;
    init_body(w, &b, st, mass, m, flags, vd);
    b.min_radius = 0.5f * std::min(width, std::min(height, length));  // BoxShape.cs:134
    if (id) *id = (uint32_t)w->bodies.size();
    w->bodies.push_back(b);
    return BEPU_OK;
}

int refcpu_add_sphere(refcpu_world* w, const bepu_motion_state* st, float radius, float mass,
                      const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!st || !(radius > 0) || std::isnan(mass) || std::isinf(mass)) return BEPU_E_INVALID_ARG;
    if (!finite3(st->position, 3) || !finite3(st->orientation, 4)) return BEPU_E_INVALID_ARG;
    Body b{};
    std::memset(&b, 0, sizeof(b));
    b.shape = BEPU_SHAPE_SPHERE;
    b.radius = radius;
    b.margin = radius;
    // SphereShape.ComputeDescription (SphereShape.cs:52-67)
    float diag = ((2.0f / 5.0f) * radius * radius);
    M3 vd{diag, 0, 0, 0, diag, 0, 0, 0, diag};
    init_body(w, &b, st, mass, m, flags, vd);
    b.min_radius = radius;  // SphereShape.cs:62
    if (id) *id = (uint32_t)w->bodies.size();
    w->bodies.push_back(b);
    return BEPU_OK;
}

int refcpu_add_capsule(refcpu_world* w, const bepu_motion_state* st, float length, float radius, float mass,
                       const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!st || !(length >= 0) || !(radius > 0) || std::isnan(mass) || std::isinf(mass)) return BEPU_E_INVALID_ARG;
    if (!finite3(st->position, 3) || !finite3(st->orientation, 4)) return BEPU_E_INVALID_ARG;
    Body b{};
    std::memset(&b, 0, sizeof(b));
    b.shape = BEPU_SHAPE_CAPSULE;
    // CapsuleShape(length, radius) and ComputeDescription (CapsuleShape.cs:40-44, 76-92)
    b.half_length = length * 0.5f;
    b.radius = radius;
    b.margin = radius;
    float effectiveLength = length + radius / 2;
    float diag = (.0833333333f * effectiveLength * effectiveLength + .25f * radius * radius);
    M3 vd{diag, 0, 0, 0, .5f * radius * radius, 0, 0, 0, diag};
    init_body(w, &b, st, mass, m, flags, vd);
    b.min_radius = radius;  // CapsuleShape.cs:89
    if (id) *id = (uint32_t)w->bodies.size();
    w->bodies.push_back(b);
    return BEPU_OK;
}

// ConvexHullShape(localSurfaceVertices, description) (ConvexHullShape.cs:114-129): the
// hull is given as its recentred surface vertices, its volume distribution
// (row-major 3x3) and its minimum radius without the margin, as
// ConvexHullShape.ComputeDescription produces them.
int refcpu_add_convex_hull(refcpu_world* w, const bepu_motion_state* st, const float* verts, uint32_t nverts,
                           const float* volume_distribution, float minimum_radius, float mass,
                           const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!st || !verts || nverts == 0 || !volume_distribution || std::isnan(mass) || std::isinf(mass))
        return BEPU_E_INVALID_ARG;
    if (!finite3(st->position, 3) || !finite3(st->orientation, 4) || !finite3(verts, 3 * nverts))
        return BEPU_E_INVALID_ARG;
    Body b{};
    std::memset(&b, 0, sizeof(b));
    b.shape = BEPU_SHAPE_CONVEX_HULL;
    b.margin = w->s.default_margin;
    b.hull_off = (int)w->hull_verts.size();
    b.hull_n = (int)nverts;
    for (uint32_t i = 0; i < nverts; ++i) w->hull_verts.push_back(V3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]});
    const float* v = volume_distribution;
    M3 vd{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
    init_body(w, &b, st, mass, m, flags, vd);
    b.min_radius = minimum_radius + b.margin;  // ConvexHullShape.cs:170
    if (id) *id = (uint32_t)w->bodies.size();
    w->bodies.push_back(b);
    return BEPU_OK;
}

// new StaticMeshShape(vertices, indices, worldTransform) (StaticMeshShape.cs:36-47,
// TransformableMeshData.cs:40-47): `transform` is the AffineTransform's linear
// part (Matrix3x3 M11..M33, row-major) and translation, or null for the
// identity; its MeshBoundingBoxTree is built here (TriangleMesh.cs:46-50).
int refcpu_create_mesh_shape(refcpu_world* w, const float* verts, uint32_t nverts, const int32_t* idx, uint32_t nidx,
                             const float* transform, uint32_t* id) {
    if (!verts || !idx || nverts == 0 || nidx == 0 || nidx % 3 != 0) return BEPU_E_INVALID_ARG;
    if (!finite3(verts, 3 * nverts) || (transform && !finite3(transform, 12))) return BEPU_E_INVALID_ARG;
    for (uint32_t i = 0; i < nidx; ++i)
        if (idx[i] < 0 || (uint32_t)idx[i] >= nverts) return BEPU_E_INVALID_ARG;
    MeshShape m;
    for (uint32_t i = 0; i < nverts; ++i) m.verts.push_back(V3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]});
    m.idx.assign(idx, idx + nidx);
    if (transform) {
        const float* t = transform;
        m.xf.m = M3{t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7], t[8]};
        m.xf.t = V3{t[9], t[10], t[11]};
    } else {
        m.xf.m = M3{1, 0, 0, 0, 1, 0, 0, 0, 1};
        m.xf.t = V3{0, 0, 0};
    }
    mesh_build_tree(&m);
    if (id) *id = (uint32_t)w->mesh_shapes.size();
    w->mesh_shapes.push_back(std::move(m));
    return BEPU_OK;
}

// new StaticMesh(...) added to the space (StaticMesh.cs): a static collidable
// (kinematic collision group, never active) whose box is the tree's root box.
int refcpu_add_static_mesh(refcpu_world* w, uint32_t shape, int sidedness, int improve_boundary,
                           const bepu_material* m, uint32_t* id) {
    if (shape >= w->mesh_shapes.size() || sidedness < 0 || sidedness > 2) return BEPU_E_INVALID_ARG;
    Body b{};
    std::memset(&b, 0, sizeof(b));
    b.shape = BEPU_SHAPE_STATIC_MESH;
    bepu_motion_state st{};
    st.orientation[3] = 1;
    init_body(w, &b, &st, 0, m, 0, M3{0, 0, 0, 0, 0, 0, 0, 0, 0});
    b.time_below = 1;  // never active (StaticCollidable.IsActive, StaticCollidable.cs:72-75)
    b.mesh = (int)shape;
    b.sidedness = sidedness;
    b.improve_boundary = improve_boundary != 0;
    const MeshShape& ms = w->mesh_shapes[shape];
    b.bmin = ms.nodes[0].bmin;
    b.bmax = ms.nodes[0].bmax;
    if (id) *id = (uint32_t)w->bodies.size();
    w->bodies.push_back(b);
    return BEPU_OK;
}

// new TerrainShape(heights, triangleOrganization) (TerrainShape.cs:60-72):
// heights[column * length + row] for `width` columns and `length` rows.
int refcpu_create_terrain_shape(refcpu_world* w, const float* heights, uint32_t width, uint32_t length,
                                int organization, uint32_t* id) {
    if (!heights || width < 2 || length < 2 || organization < 0 || organization > 1) return BEPU_E_INVALID_ARG;
    if (!finite3(heights, (int)(width * length))) return BEPU_E_INVALID_ARG;
    TerrainShapeRef t;
    t.width = (int)width;
    t.length = (int)length;
    t.organization = organization;
    t.heights.assign(heights, heights + (size_t)width * length);
    if (id) *id = (uint32_t)w->terrain_shapes.size();
    w->terrain_shapes.push_back(std::move(t));
    return BEPU_OK;
}

// new Terrain(shape, worldTransform) added to the space (Terrain.cs): its
// sidedness follows the transform's winding; its box is the shape's under
// the transform (Terrain.UpdateBoundingBox, thickness 0).
int refcpu_add_terrain(refcpu_world* w, uint32_t shape, const float* transform, int improve_boundary,
                       const bepu_material* m, uint32_t* id) {
    if (shape >= w->terrain_shapes.size() || !transform || !finite3(transform, 12)) return BEPU_E_INVALID_ARG;
    TerrainInst ti;
    ti.shape = w->terrain_shapes[shape];
    ti.shape.xf.m = M3{transform[0], transform[1], transform[2], transform[3], transform[4], transform[5],
                       transform[6], transform[7], transform[8]};
    ti.shape.xf.t = V3{transform[9], transform[10], transform[11]};
    ti.sidedness = terrain_sidedness(ti.shape.xf);
    Body b{};
    std::memset(&b, 0, sizeof(b));
    b.shape = BEPU_SHAPE_TERRAIN;
    bepu_motion_state st{};
    st.orientation[3] = 1;
    init_body(w, &b, &st, 0, m, 0, M3{0, 0, 0, 0, 0, 0, 0, 0, 0});
    b.time_below = 1;  // never active
    b.terrain = (int)w->terrains.size();
    b.sidedness = ti.sidedness;
    b.improve_boundary = improve_boundary != 0;
    terrain_bounding_box(ti.shape, &b.bmin, &b.bmax);
    // The thickness offset (Terrain.UpdateBoundingBox): Normalize(Down) * 0.
    V3 off = mul(normalize(V3{-ti.shape.xf.m.m21, -ti.shape.xf.m.m22, -ti.shape.xf.m.m23}), 0.0f);
    if (off.x < 0) b.bmin.x += off.x; else b.bmax.x += off.x;
    if (off.y < 0) b.bmin.y += off.y; else b.bmax.y += off.y;
    if (off.z < 0) b.bmin.z += off.z; else b.bmax.z += off.z;
    w->terrains.push_back(ti);
    if (id) *id = (uint32_t)w->bodies.size();
    w->bodies.push_back(b);
    return BEPU_OK;
}

// The largest per-pair list sizes the mesh manifold updates met (tests).
int refcpu_mesh_peaks(refcpu_world* w, int32_t* out6) {
    for (int k = 0; k < 6; ++k) out6[k] = w->mesh_peak[k];
    return BEPU_OK;
}

// TerrainShape.GetOverlaps of a terrain body over a box in its local space (tests).
int refcpu_terrain_overlaps(refcpu_world* w, uint32_t body, const float* lmin, const float* lmax, int32_t* out,
                            uint32_t cap, uint32_t* n) {
    if (body >= w->bodies.size() || w->bodies[body].terrain < 0) return BEPU_E_INVALID_ARG;
    std::vector<int> tris;
    terrain_overlaps(w->terrains[w->bodies[body].terrain].shape, V3{lmin[0], lmin[1], lmin[2]},
                     V3{lmax[0], lmax[1], lmax[2]}, &tris);
    for (uint32_t k = 0; k < tris.size() && k < cap; ++k) out[k] = tris[k];
    if (n) *n = (uint32_t)tris.size();
    return BEPU_OK;
}

// One triangle of a terrain body (tests): world != 0 gives TerrainShape.GetTriangle
// in world space; world == 0 the GetLocalIndices vertices in the shape's local
// space with their sequential indices.
int refcpu_terrain_triangle(refcpu_world* w, uint32_t body, int tri, int world, float* out9, int32_t* idx3) {
    if (body >= w->bodies.size() || w->bodies[body].terrain < 0) return BEPU_E_INVALID_ARG;
    const TerrainShapeRef& t = w->terrains[w->bodies[body].terrain].shape;
    V3 v[3];
    if (world) {
        terrain_world_triangle(t, tri, v);
    } else {
        int col[3], row[3];
        terrain_local_indices(t, tri, col, row);
        for (int k = 0; k < 3; ++k) {
            v[k] = V3{(float)col[k], t.h(col[k], row[k]), (float)row[k]};
            if (idx3) idx3[k] = row[k] * t.width + col[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        out9[3 * k] = v[k].x;
        out9[3 * k + 1] = v[k].y;
        out9[3 * k + 2] = v[k].z;
    }
    return BEPU_OK;
}

// The box a terrain pair's manifold queries (tests): the convex body's shape
// in the terrain's local space, grown by its velocity over one step.
int refcpu_terrain_query_box(refcpu_world* w, uint32_t convex, uint32_t body, float* out6) {
    if (body >= w->bodies.size() || w->bodies[body].terrain < 0 || convex >= w->bodies.size() ||
        is_static(w->bodies[convex]))
        return BEPU_E_INVALID_ARG;
    V3 mn, mx;
    terrain_query_box(w, w->bodies[convex], w->terrains[w->bodies[body].terrain].shape.xf, &mn, &mx);
    const float o[6] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z};
    std::memcpy(out6, o, sizeof(o));
    return BEPU_OK;
}

// The flattened tree of a mesh shape (tests): per node bmin[3], bmax[3], a, b.
int refcpu_get_mesh_tree(refcpu_world* w, uint32_t shape, float* boxes, int32_t* links, uint32_t cap, uint32_t* n) {
    if (shape >= w->mesh_shapes.size()) return BEPU_E_INVALID_ARG;
    const MeshShape& m = w->mesh_shapes[shape];
    *n = (uint32_t)m.nodes.size();
    for (uint32_t k = 0; k < m.nodes.size() && k < cap; ++k) {
        const MeshNode& nd = m.nodes[k];
        float* o = boxes + 6 * k;
        o[0] = nd.bmin.x; o[1] = nd.bmin.y; o[2] = nd.bmin.z;
        o[3] = nd.bmax.x; o[4] = nd.bmax.y; o[5] = nd.bmax.z;
        links[2 * k] = nd.a;
        links[2 * k + 1] = nd.b;
    }
    return BEPU_OK;
}

// MeshBoundingBoxTree.GetOverlaps for a box (tests): triangles in traversal order.
int refcpu_mesh_overlaps(refcpu_world* w, uint32_t shape, const float* bmin, const float* bmax, int32_t* out,
                         uint32_t cap, uint32_t* n) {
    if (shape >= w->mesh_shapes.size()) return BEPU_E_INVALID_ARG;
    std::vector<int> t;
    mesh_overlaps(w->mesh_shapes[shape], V3{bmin[0], bmin[1], bmin[2]}, V3{bmax[0], bmax[1], bmax[2]}, &t);
    *n = (uint32_t)t.size();
    for (uint32_t k = 0; k < t.size() && k < cap; ++k) out[k] = t[k];
    return BEPU_OK;
}

int refcpu_step(refcpu_world* w, int n) {
    for (int i = 0; i < n; ++i) step_once(w);
    return BEPU_OK;
}

// Runs only the body stages and the broadphase of one step (no narrowphase,
// solver or integration), for AABB/pair-set checks on a frozen scene.
int refcpu_broadphase_only(refcpu_world* w) {
    for (Body& b : w->bodies) candidacy(w, b);
    for (Body& b : w->bodies) force_update(w, b);
    for (Body& b : w->bodies) bounding_box(w, b);
    if (w->broadphase_mode == 1) broadphase_bruteforce(w, &w->overlaps);
    else if (w->broadphase_mode == 2) broadphase_sweep(w, &w->overlaps);
    else broadphase_grid(w, &w->overlaps);
    return BEPU_OK;
}

int refcpu_get_motion_states(refcpu_world* w, bepu_motion_state* out, uint32_t cap, uint32_t* n) {
    uint32_t cnt = (uint32_t)w->bodies.size();
    if (n) *n = cnt;
    for (uint32_t i = 0; i < cnt && i < cap; ++i) {
        const Body& b = w->bodies[i];
        bepu_motion_state& s = out[i];
        s.position[0] = b.pos.x; s.position[1] = b.pos.y; s.position[2] = b.pos.z;
        s.orientation[0] = b.q.x; s.orientation[1] = b.q.y; s.orientation[2] = b.q.z; s.orientation[3] = b.q.w;
        s.linear_velocity[0] = b.v.x; s.linear_velocity[1] = b.v.y; s.linear_velocity[2] = b.v.z;
        s.angular_velocity[0] = b.w.x; s.angular_velocity[1] = b.w.y; s.angular_velocity[2] = b.w.z;
    }
    return cnt <= cap ? BEPU_OK : BEPU_E_CAPACITY;
}

int refcpu_get_aabbs(refcpu_world* w, float* out, uint32_t cap, uint32_t* n) {
    uint32_t cnt = (uint32_t)w->bodies.size();
    if (n) *n = cnt;
    for (uint32_t i = 0; i < cnt && i < cap; ++i) {
        const Body& b = w->bodies[i];
        float* o = out + 6 * i;
        o[0] = b.bmin.x; o[1] = b.bmin.y; o[2] = b.bmin.z;
        o[3] = b.bmax.x; o[4] = b.bmax.y; o[5] = b.bmax.z;
    }
    return cnt <= cap ? BEPU_OK : BEPU_E_CAPACITY;
}

int refcpu_get_pairs(refcpu_world* w, uint64_t* keys, uint32_t cap, uint32_t* n) {
    uint32_t cnt = (uint32_t)w->overlaps.size();
    if (n) *n = cnt;
    for (uint32_t i = 0; i < cnt && i < cap; ++i) keys[i] = w->overlaps[i];
    return cnt <= cap ? BEPU_OK : BEPU_E_CAPACITY;
}

int refcpu_get_manifolds(refcpu_world* w, bepu_manifold* out, uint32_t cap, uint32_t* n) {
    uint32_t cnt = 0;
    for (size_t k = 0; k < w->pairs.keys.size(); ++k) {
        const Pair& p = w->pairs.vals[k];
        if (p.ccount == 0) continue;
        if (cnt < cap) {
            bepu_manifold& m = out[cnt];
            std::memset(&m, 0, sizeof(m));
            m.key = w->pairs.keys[k];
            m.body_a = p.a;
            m.body_b = p.b;
            m.count = p.ccount;
            m.kind = p.kind;
            for (int i = 0; i < p.ccount; ++i) {
                const ContactObj& c = p.slot[p.clist[i]];
                bepu_contact& o = m.contacts[i];
                o.position[0] = c.pos.x; o.position[1] = c.pos.y; o.position[2] = c.pos.z;
                o.normal[0] = c.normal.x; o.normal[1] = c.normal.y; o.normal[2] = c.normal.z;
                o.penetration_depth = c.depth;
                o.id = c.id;
            }
        }
        cnt++;
    }
    if (n) *n = cnt;
    return cnt <= cap ? BEPU_OK : BEPU_E_CAPACITY;
}

// Accumulated impulses of every constraint (a sleeping island's keep the
// values they had when it went to sleep), by colour of the last solve, then key.
int refcpu_get_impulses(refcpu_world* w, bepu_impulses* out, uint32_t cap, uint32_t* n) {
    std::vector<const Pair*> live;
    for (const Pair& p : w->pairs.vals)
        if (p.ccount > 0) live.push_back(&p);
    std::sort(live.begin(), live.end(), [](const Pair* a, const Pair* b) {
        return a->color != b->color ? a->color < b->color : pair_key(a->a, a->b) < pair_key(b->a, b->b);
    });
    uint32_t cnt = (uint32_t)live.size();
    if (n) *n = cnt;
    for (uint32_t i = 0; i < cnt && i < cap; ++i) {
        const Pair& p = *live[i];
        bepu_impulses& o = out[i];
        std::memset(&o, 0, sizeof(o));
        o.key = pair_key(p.a, p.b);
        o.color = p.color;
        o.count = p.npen;
        for (int k = 0; k < p.npen; ++k) o.normal[k] = p.pen[k].acc;
        for (int k = 0; k < p.nfric; ++k) o.friction[k] = p.fric[k].acc;
        o.sliding[0] = p.sl.acc[0];
        o.sliding[1] = p.sl.acc[1];
        o.twist = p.tw.acc;
    }
    return cnt <= cap ? BEPU_OK : BEPU_E_CAPACITY;
}

int refcpu_get_stats(refcpu_world* w, bepu_step_stats* s) {
    std::memset(s, 0, sizeof(*s));
    s->bodies = (uint32_t)w->bodies.size();
    s->pairs = (uint32_t)w->overlaps.size();
    uint32_t m = 0, c = 0;
    for (const Pair& p : w->pairs.vals)
        if (p.ccount > 0) { m++; c += p.ccount; }
    s->manifolds = m;
    s->contacts = c;
    s->colors = (uint32_t)w->ncolors;
    s->steps = w->steps;
    s->max_penetration = w->max_pen;
    return BEPU_OK;
}

// Space.Remove(entity) (Space.cs:364-482).  An unknown or already removed id
// is the reference's ArgumentException ("not a member of this space").
// The removed member's island loses it, and the dynamic bodies it had
// constraints with are activated (Space.Remove -> DeactivationManager.Remove
// -> RemoveSimulationIslandFromMember, DeactivationManager.cs:565-656).
int refcpu_remove(refcpu_world* w, uint32_t id) {
    if (id >= w->bodies.size() || w->bodies[id].removed) return BEPU_E_NOT_FOUND;
    w->bodies[id].removed = true;
    for (const Pair& p : w->pairs.vals) {
        if (p.ccount <= 0 || (p.a != id && p.b != id)) continue;
        Body& other = w->bodies[p.a == id ? p.b : p.a];
        if (other.dynamic) activate(other);
    }
    // A joint of the removed body stops acting; its other connection wakes.
    for (JointC& j : w->joints) {
        if (j.dead || (j.a != id && j.b != id)) continue;
        j.dead = true;
        Body& other = w->bodies[j.a == id ? j.b : j.a];
        if (other.dynamic && !other.removed) activate(other);
    }
    return BEPU_OK;
}

// new BallSocketJoint(connectionA, connectionB, anchor) (BallSocketJoint.cs:57-63)
// added to the space: the offsets from each body's position to the anchor,
// stored in each body's local frame (the OffsetA / OffsetB setters,
// BallSocketJoint.cs:108-130); SpringSettings defaults (SpringSettings.cs:
// 34-35), no corrective velocity limit (Joint.cs:14-15).
int refcpu_add_ball_socket(refcpu_world* w, uint32_t a, uint32_t b, const float* anchor, uint32_t* id) {
    if (!anchor || a == b) return BEPU_E_INVALID_ARG;
    if (a >= w->bodies.size() || b >= w->bodies.size() || w->bodies[a].removed || w->bodies[b].removed)
        return BEPU_E_NOT_FOUND;
    const Body& A = w->bodies[a];
    const Body& B = w->bodies[b];
    if (!A.dynamic && !B.dynamic) return BEPU_E_INVALID_ARG;  // "Cannot constrain two kinematic bodies."
    if (is_static(A) || is_static(B)) return BEPU_E_INVALID_ARG;  // joints connect entities
    JointC j;
    std::memset(&j, 0, sizeof(j));
    j.a = a;
    j.b = b;
    j.local = (uint32_t)w->joints.size();
    const V3 p{anchor[0], anchor[1], anchor[2]};
    j.localA = m3_transform_transpose(sub(p, A.pos), A.R);
    j.localB = m3_transform_transpose(sub(p, B.pos), B.R);
    j.stiffness = 600000;
    j.damping = 90000;
    j.max_corr = FLT_MAX;
    j.max_corr_sq = FLT_MAX;
    j.color = -1;
    if (id) *id = (uint32_t)w->joints.size();
    w->joints.push_back(j);
    return BEPU_OK;
}

// SpringSettings.Stiffness / Damping (clamped at 0, SpringSettings.cs:61-80)
// and Joint.MaxCorrectiveVelocity (Joint.cs:24-38).
int refcpu_set_joint_settings(refcpu_world* w, uint32_t id, float stiffness, float damping, float max_corrective) {
    if (id >= w->joints.size()) return BEPU_E_NOT_FOUND;
    stiffness = std::max(0.0f, stiffness);
    damping = std::max(0.0f, damping);
    // "Constraints cannot have both 0 stiffness and 0 damping." (SpringSettings.cs:119)
    if (stiffness == 0 && damping == 0) return BEPU_E_INVALID_ARG;
    JointC& j = w->joints[id];
    j.stiffness = stiffness;
    j.damping = damping;
    j.max_corr = std::max(0.0f, max_corrective);
    j.max_corr_sq = j.max_corr >= FLT_MAX ? FLT_MAX : j.max_corr * j.max_corr;
    return BEPU_OK;
}

// BallSocketJoint.TotalImpulse of every joint (3 floats each).
int refcpu_get_joint_impulses(refcpu_world* w, float* out, uint32_t cap, uint32_t* n) {
    if (n) *n = (uint32_t)w->joints.size();
    if (!out) return BEPU_OK;
    if (cap < w->joints.size()) return BEPU_E_CAPACITY;
    for (size_t k = 0; k < w->joints.size(); ++k) {
        out[3 * k] = w->joints[k].acc.x;
        out[3 * k + 1] = w->joints[k].acc.y;
        out[3 * k + 2] = w->joints[k].acc.z;
    }
    return BEPU_OK;
}

// Solve order of the following steps: 0 colour by colour, 1 the reference's
// permuted single-threaded order, 2 the same over the list in descending key
// order (test instrument: how far the colour order moves impulses and stacks
// from the reference's kind of order, beside how far the reference's own
// order moves with the order constraints were added).
int refcpu_set_solve_order(refcpu_world* w, int mode) {
    if (mode < 0 || mode > 2) return BEPU_E_INVALID_ARG;
    w->solve_order = mode;
    return BEPU_OK;
}

// The permuted visit order of step k over n constraints: out[i] for i < count.
int refcpu_permuted_order(int64_t n, int64_t k, int64_t count, int64_t* out) {
    if (n <= 0 || k < 0 || count < 0) return BEPU_E_INVALID_ARG;
    for (int64_t i = 0; i < count; ++i) out[i] = permuted_index(i, n, k);
    return BEPU_OK;
}

// Entity.ActivityInformation.Activate() for the listed bodies.
int refcpu_activate(refcpu_world* w, const uint32_t* ids, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= w->bodies.size() || w->bodies[ids[k]].removed) return BEPU_E_NOT_FOUND;
        if (is_static(w->bodies[ids[k]])) return BEPU_E_INVALID_ARG;  // a static collidable has no activity
    }
    for (uint32_t k = 0; k < n; ++k) activate(w->bodies[ids[k]]);
    return BEPU_OK;
}

// Entity.ActivityInformation.IsActive of every body (0 for a removed one).
int refcpu_get_activity(refcpu_world* w, uint8_t* out, uint32_t cap, uint32_t* n) {
    uint32_t cnt = (uint32_t)w->bodies.size();
    if (n) *n = cnt;
    for (uint32_t i = 0; i < cnt && i < cap; ++i) out[i] = body_active(w->bodies[i]) ? 1 : 0;
    return cnt <= cap ? BEPU_OK : BEPU_E_CAPACITY;
}

// Entity state setters, for each listed body the parts `fields` selects:
// Position (EntityBase.cs:51-62), Orientation (normalised; orientation matrix
// and world inertia refreshed, EntityBase.cs:68-84), LinearVelocity
// (:181-195), AngularVelocity (:125-140); all four: MotionState (:218-235).
int refcpu_set_motion_states(refcpu_world* w, const uint32_t* ids, const bepu_motion_state* st, uint32_t n,
                             uint32_t fields) {
    if (fields == 0 || (fields & ~BEPU_STATE_ALL)) return BEPU_E_INVALID_ARG;
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= w->bodies.size() || w->bodies[ids[k]].removed) return BEPU_E_NOT_FOUND;
        if (is_static(w->bodies[ids[k]])) return BEPU_E_INVALID_ARG;  // a static collidable is not an entity
        const bepu_motion_state& m = st[k];
        if (((fields & BEPU_STATE_POSITION) && !finite3(m.position, 3)) ||
            ((fields & BEPU_STATE_ORIENTATION) && !finite3(m.orientation, 4)) ||
            ((fields & BEPU_STATE_LINEAR_VELOCITY) && !finite3(m.linear_velocity, 3)) ||
            ((fields & BEPU_STATE_ANGULAR_VELOCITY) && !finite3(m.angular_velocity, 3)))
            return BEPU_E_INVALID_ARG;
    }
    for (uint32_t k = 0; k < n; ++k) {
        Body& b = w->bodies[ids[k]];
        const bepu_motion_state& m = st[k];
        if (fields & BEPU_STATE_POSITION) b.pos = V3{m.position[0], m.position[1], m.position[2]};
        if (fields & BEPU_STATE_ORIENTATION) {
            b.q = qnormalize(Q4{m.orientation[0], m.orientation[1], m.orientation[2], m.orientation[3]});
            b.R = m3_from_quat(b.q);
            b.inertia_inv = m3_mul(m3_mul_transposed(b.R, b.local_inertia_inv), b.R);
        }
        if (fields & BEPU_STATE_LINEAR_VELOCITY)
            b.v = V3{m.linear_velocity[0], m.linear_velocity[1], m.linear_velocity[2]};
        if (fields & BEPU_STATE_ANGULAR_VELOCITY)
            b.w = V3{m.angular_velocity[0], m.angular_velocity[1], m.angular_velocity[2]};
        activate(b);  // every one of these setters activates the entity
    }
    return BEPU_OK;
}

// Solver.IterationLimit, ForceUpdater.Gravity, TimeStepSettings.TimeStepDuration
// and the other settings changed between steps (Constraints/Solver.cs:19,
// OtherSpaceStages/ForceUpdater.cs:19-29, TimeStepSettings.cs:19).  Per-entity
// values taken at creation (damping) keep their values, as the reference's do.
int refcpu_set_settings(refcpu_world* w, const bepu_settings* s) {
    if (!s || s->min_iteration_count > 6 || s->iteration_limit < 0 || !(s->dt > 0) || std::isnan(s->core_shape_scaling))
        return BEPU_E_INVALID_ARG;
    w->s = *s;
    w->s.core_shape_scaling = clampf(w->s.core_shape_scaling, 0, .99f);  // MotionSettings.cs:22-34
    return BEPU_OK;
}

// Entity.PositionUpdateMode setter (EntityBase.cs:1050-1068) for each listed body.
int refcpu_set_position_update_mode(refcpu_world* w, const uint32_t* ids, uint32_t n, int mode) {
    if ((n && !ids) || mode < 0 || mode > 2) return BEPU_E_INVALID_ARG;
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= w->bodies.size() || w->bodies[ids[k]].removed) return BEPU_E_NOT_FOUND;
        if (is_static(w->bodies[ids[k]])) return BEPU_E_INVALID_ARG;
    }
    for (uint32_t k = 0; k < n; ++k) w->bodies[ids[k]].pmode = mode;
    if (mode && n) w->ccd = true;
    return BEPU_OK;
}

// ContactReducer.ReduceContacts on a hand-built manifold: four contacts
// (positions xyz and depths, list order) with a shared normal and one
// candidate.  Returns the list index removed (-1: none); *add_candidate
// tells whether the candidate joins.
int refcpu_reduce_contacts(const float* pos12, const float* depth4, const float* normal3, const float* cand_pos3,
                           float cand_depth, int* add_candidate) {
    Pair p;
    std::memset(&p, 0, sizeof(p));
    for (int i = 0; i < 4; ++i) {
        p.slot[i].pos = V3{pos12[3 * i], pos12[3 * i + 1], pos12[3 * i + 2]};
        p.slot[i].normal = V3{normal3[0], normal3[1], normal3[2]};
        p.slot[i].depth = depth4[i];
        p.used[i] = true;
        p.clist[i] = i;
    }
    p.ccount = 4;
    ContactData cand;
    cand.pos = V3{cand_pos3[0], cand_pos3[1], cand_pos3[2]};
    cand.normal = V3{normal3[0], normal3[1], normal3[2]};
    cand.depth = cand_depth;
    cand.id = 0;
    bool add = false;
    int r = reduce_contacts(p, cand, &add);
    *add_candidate = add ? 1 : 0;
    return r;
}

// Stand-alone collision tests for closed-form checks.
int refcpu_box_box(const float* halfA, const float* posA, const float* qA, const float* halfB, const float* posB,
                   const float* qB, float* axis_out, float* distance_out, float* contacts_out /* 8 x (x,y,z,depth) */,
                   int* ids_out) {
    float d;
    V3 ax;
    BoxContactList m;
    bool hit = box_box_colliding(V3{halfA[0], halfA[1], halfA[2]}, V3{halfB[0], halfB[1], halfB[2]},
                                 V3{posA[0], posA[1], posA[2]}, Q4{qA[0], qA[1], qA[2], qA[3]},
                                 V3{posB[0], posB[1], posB[2]}, Q4{qB[0], qB[1], qB[2], qB[3]}, &d, &ax, &m);
    axis_out[0] = ax.x; axis_out[1] = ax.y; axis_out[2] = ax.z;
    *distance_out = d;
    if (!hit) return -1;
    for (int i = 0; i < m.n; ++i) {
        contacts_out[4 * i + 0] = m.e[i].pos.x;
        contacts_out[4 * i + 1] = m.e[i].pos.y;
        contacts_out[4 * i + 2] = m.e[i].pos.z;
        contacts_out[4 * i + 3] = m.e[i].depth;
        ids_out[i] = m.e[i].id;
    }
    return m.n;
}

int refcpu_sphere_sphere(float ra, float rb, const float* pa, const float* pb, float* out7) {
    ContactData c{};
    if (!sphere_sphere_colliding(ra, rb, V3{pa[0], pa[1], pa[2]}, V3{pb[0], pb[1], pb[2]}, .1f, &c)) return 0;
    out7[0] = c.pos.x; out7[1] = c.pos.y; out7[2] = c.pos.z;
    out7[3] = c.normal.x; out7[4] = c.normal.y; out7[5] = c.normal.z; out7[6] = c.depth;
    return 1;
}

int refcpu_box_sphere(const float* half, const float* bp, const float* bq, float r, const float* sp, float* out7) {
    ContactData c{};
    if (!box_sphere_colliding(V3{half[0], half[1], half[2]}, V3{bp[0], bp[1], bp[2]}, Q4{bq[0], bq[1], bq[2], bq[3]},
                              r, V3{sp[0], sp[1], sp[2]}, .1f, &c))
        return 0;
    out7[0] = c.pos.x; out7[1] = c.pos.y; out7[2] = c.pos.z;
    out7[3] = c.normal.x; out7[4] = c.normal.y; out7[5] = c.normal.z; out7[6] = c.depth;
    return 1;
}

uint64_t refcpu_mix64(uint64_t k) { return mix64(k); }

}  // extern "C"
