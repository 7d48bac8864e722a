// world.h -- device-resident layout of one simulation world (or a batch of
// worlds sharing one set of arrays) and the kernel launch interface.
//
// Layout in HBM (all arrays structure-of-arrays, float4-packed so that every
// per-element access is a coalesced 16-byte load):
//   bodies       pos4(x,y,z,-) quat4 linv4(v,-) angv4(w,-) inv_inertia rows r0..r2
//                (r0.w = inverse mass), local inverse inertia l0..l2, shape4
//                (half extents / radius), aabb_min4/aabb_max4, stab4
//                (previous |v|^2+|w|^2, time below limit, slowing flag,
//                stabilization allowed), damp2, material4, flags (dynamic).
//   broadphase   morton keys + sorted leaf ids, LBVH nodes (Karras), pair list.
//   pairs        PairState records double-buffered across steps and an
//                open-addressing key -> slot table per buffer.
//   solver       constraint list, colours, per-colour ranges and the rows of
//                every contact constraint in solve order (SolverRows).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bepu_gpu.h"
#include "dmath.h"

namespace bepu {

// Persistent per-pair state: manifold contacts plus everything the
// reference keeps alive between steps (accumulated impulses for warm start,
// friction directions, the sphere manifolds' previouslyColliding flag).
struct PairState {
    uint64_t key;
    uint32_t a, b;        // A/B as the pair handler orders them
    int32_t kind;         // BEPU_PAIR_*
    int32_t ccount;       // contacts in the manifold's contact list
    int32_t npen;         // penetration constraints (list order)
    uint32_t flags;       // bit0 prev_colliding, bit1 sliding/twist active, bits 4..7 slot used
    float kinetic, static_f, bounce;
    int8_t clist[4];      // contact list -> slot (FastRemoveAt order)
    int8_t pslot[4];      // penetration constraint list -> slot (ordered RemoveAt)
    float4 cpos[4];       // slot: position xyz, penetration depth
    float4 cnrm[4];       // slot: normal xyz, id (int bits)
    float penacc[4];      // per penetration constraint (list order)
    float fracc[4];       // non-convex ContactFrictionConstraint accumulated impulse
    float4 frlin[4];      // its remembered direction (xyz)
    float slacc[2];       // SlidingFrictionTwoAxis accumulated impulse
    float twacc;          // TwistFrictionConstraint accumulated impulse
    int32_t color;        // colour of the constraint in the last solve (-1 before)
    float sllin[6];       // SlidingFrictionTwoAxis linearA (remembered across steps)
    int32_t tester;       // general convex pairs: the pair tester's state (ConvexState::sep.w) after the last test
    float toi;            // CollidablePairHandler.timeOfImpact: 1 for a new pair, recomputed in the steps
                          // it has a continuous body and an active one, kept otherwise
};

// PF_FROZEN: not updated this step (both bodies inactive when the broad phase
// ran; the state is last step's); PF_DEAD: a frozen pair whose island woke in
// this step, stale and removed (the next step creates it again); PF_VOID: an
// overlap of two inactive bodies with no pair to keep (no state).  Dead and
// void slots are not pairs: no constraint, no readback, no history.
enum { PF_PREV_COLLIDING = 1u, PF_FRICTION_ACTIVE = 2u, PF_FROZEN = 256u, PF_DEAD = 512u, PF_VOID = 1024u };

// The fields of a pair slot that the island, deactivation and statistics
// passes read, beside the 336-byte state (one 16-byte load per pair instead
// of two cache lines of the state).  Written wherever the state is stored;
// `flags` carries the PF_FROZEN / PF_DEAD / PF_VOID bits.
struct PairHot {
    uint32_t a, b;
    int32_t ccount;
    uint32_t flags;
};
__host__ __device__ __forceinline__ PairHot hot_of(const PairState& s) { return PairHot{s.a, s.b, s.ccount, s.flags}; }

// State of a general convex pair (a pair with a hull or a capsule) that the
// PairState does not hold: the GeneralConvexPairTester's state machine
// (sep.w: state as int bits), its boolean-GJK warm-start axis (sep.xyz) and
// deep-contact search direction (dir.xyz), and the manifold's supplement
// data per contact slot (local offsets on A and B; offa.w: base depth).
struct ConvexState {
    float4 sep;
    float4 dir;
    float4 offa[4];
    float4 offb[4];
};
// BF_GENERAL: a shape whose pairs take the general convex tester (hull or
// capsule); BF_HULL: a convex hull; BF_SPHERE: a sphere.
// BF_ALWAYS_ACTIVE: SimulationIslandMember.IsAlwaysActive; BF_CANDIDATE: its
// IsDeactivationCandidate.  Activity: BF_ASLEEP, a dynamic body whose island
// sleeps; BF_KIN_ACTIVE, a kinematic body active this step (its flag, the
// time below the limit in stab.y, is <= 0); BF_NODEACT, the island may not
// deactivate at its next check (an active kinematic body touched it);
// BF_WAKE, activated since the last step (its island is active at the next).
// BF_PASSIVE / BF_CONTINUOUS: PositionUpdateMode Passive / Continuous (none:
// Discrete).  BF_MESH: a static mesh or a terrain (a static collidable:
// never dynamic, never active, its box fixed).
enum { BF_DYNAMIC = 1u, BF_PRESENT = 2u, BF_GENERAL = 4u, BF_HULL = 8u, BF_SPHERE = 16u, BF_ALWAYS_ACTIVE = 32u,
       BF_CANDIDATE = 64u, BF_ASLEEP = 128u, BF_KIN_ACTIVE = 256u, BF_NODEACT = 512u, BF_WAKE = 1024u,
       BF_PASSIVE = 2048u, BF_CONTINUOUS = 4096u, BF_MESH = 8192u };
constexpr uint32_t BF_CCD_MODE = BF_PASSIVE | BF_CONTINUOUS;

// One body as built on the host (Entity construction), copied to the device
// in one block and scattered into the structure-of-arrays by k_scatter_bodies.
struct BodyRecord {
    float4 pos, quat, linv, angv, ii0, ii1, ii2, li0, li1, li2, shape, stab, mat;
    float2 damp;
    uint32_t flags;   // BF_*
    uint32_t wid;     // sub-world
    uint32_t wbase;   // its first body id
    uint32_t born;    // index of the first step the body takes part in
    uint32_t target;  // body id
    float minrad;     // the shape's MinimumRadius
};
static_assert(sizeof(BodyRecord) == 240, "BodyRecord layout");

// A BallSocketJoint (Constraints/TwoEntity/Joints/BallSocketJoint.cs): its
// connections, local anchors, spring settings and accumulated impulse.  In
// the solver's constraint ids, those at and above pair_cap name joints (id -
// pair_cap); those below are pair slots.
struct JointState {
    uint32_t a, b;      // ConnectionA / ConnectionB
    uint32_t local;     // ordinal among its sub-world's joints: colouring key kJointKey | local
    uint32_t flags;     // JF_DEAD: a connection was removed, the joint no longer acts
    float4 la;          // localAnchorA; w: SpringSettings.Stiffness
    float4 lb;          // localAnchorB; w: SpringSettings.Damping
    float4 acc;         // accumulatedImpulse; w: MaxCorrectiveVelocity
    float max_corr_sq;  // its square (float max when unlimited)
    int32_t color;      // colour in the last solve (-1 before)
    uint32_t pad[2];
};
static_assert(sizeof(JointState) == 80, "JointState layout");
enum { JF_DEAD = 1u };
// The joints' colouring keys lie above every world-local pair key.
constexpr uint64_t kJointKey = 1ull << 63;

struct RowConst;
struct RowVar;
struct MeshTester;

struct DeviceWorld {
    bepu_settings s;
    uint32_t nbodies;
    uint32_t body_cap;
    uint32_t pair_cap;
    uint32_t hash_cap;  // power of two
    // bodies
    float4 *pos, *quat, *linv, *angv;
    float4 *ii0, *ii1, *ii2;   // world inverse inertia rows; ii0.w = inverse mass
    float4 *li0, *li1, *li2;   // local inverse inertia rows
    float4* shape;             // w = shape kind; box: half extents; sphere: (radius, 0, 0);
                               // capsule: (radius, half length, 0); hull: (margin, first vertex, vertex count as uint bits)
    float4* hull_verts;        // vertices of every hull shape, back to back (xyz)
    uint32_t hull_nverts;      // vertices in hull_verts
    uint32_t hull_max_verts;   // vertices of the largest hull shape
    uint32_t small_island_max; // islands of at most this many constraints are packed per wavefront
    uint32_t pair_lists;       // 1: k_pair_prep lists this step's pairs by tester; 0: every pair is a
                               // closed-form pair of one kind and k_narrowphase takes them in order
    float4 *bmin, *bmax;
    float4* stab;              // prev velocity, time below, slowing, allow stabilization
    float2* damp;              // linear, angular damping
    float4* mat;               // static, kinetic, bounce
    uint32_t* bflags;          // BF_* bits (0: absent -- a removed world's slot)
    uint32_t* wid;             // sub-world of every body (bodies of different sub-worlds never interact)
    uint32_t* wbase;           // first body id of the body's sub-world
    uint32_t* wlevels;         // per sub-world: grid levels holding at least one of its bodies
    uint32_t* wlev_count;      // per sub-world and grid level (16 each): bodies at that level
    uint32_t* wlev_list;       // the first kLevelList of them (a sparse coarse level is tested directly)
    uint32_t* wfault;          // per sub-world: 1 once a body's state became non-finite (the world is evicted)
    uint32_t nworlds;
    uint32_t* born;            // index of the first step a body takes part in
    // continuous collision detection (run only when `ccd`: a body was ever
    // made passive or continuous)
    uint32_t ccd;
    float* minrad;             // per body: the shape's MinimumRadius
    uint32_t* toi;             // per body: smallest time of impact of its pairs this step (float bits)
    float4* qold;              // per passive / continuous body: its orientation before this step's update
    // static meshes (present once a mesh shape was created): per mesh shape
    // 4 float4 in mesh_info (the rows of its world transform's linear part
    // with its vertex, index and node bases in .w, its translation with its
    // triangle count in .w); local vertices (xyz), vertex indices; tree nodes
    // as 2 float4 each, (box min, child a or -1 for a leaf) and (box max,
    // child b or the leaf's triangle), children relative to the shape's node
    // base.  Mesh pairs keep their tester tables per pair slot (kMeshTesters
    // entries, double-buffered like the pair states; the count is in
    // PairState::tester) and are listed for k_mesh as (slot, last slot).
    uint32_t meshes;
    float4* mesh_info;
    float4* mesh_verts;
    int32_t* mesh_idx;
    float4* mesh_nodes;
    MeshTester* mtest_cur;
    MeshTester* mtest_old;
    uint2* mesh_work;
    // terrains (their pairs share the mesh pairs' tables and list): per
    // terrain shape 6 float4 in terrain_info (the rows of its transform's
    // linear part with its height base, width and length in .w, its
    // translation with its QuadTriangleOrganization in .w, its world box min
    // and max); heights column by column (heights[column * length + row]).
    float4* terrain_info;
    float* terrain_heights;
    // joints
    JointState* joints;
    uint32_t njoints;
    uint32_t joint_cap;
    // broadphase
    uint32_t* morton;          // Morton codes (tree broadphase)
    uint32_t* morton_sorted;
    uint32_t* leaf_ids;        // body id per sorted leaf
    uint32_t* leaf_ids_tmp;
    float4 *node_min, *node_max;  // implicit tree, levels stored back to back (see tree_layout)
    uint32_t tree_levels;
    uint32_t tree_off[32];
    uint32_t tree_cnt[32];
    uint32_t* bounds;          // [0..5] ordered-int scene bounds of AABB centres, [6] grid base cell edge (float bits)
    uint32_t* bounds_part;     // per-block partial bounds (6 per block of 256 bodies)
    uint32_t bp_kind;          // BEPU_BROADPHASE_*
    // hash grid
    uint64_t* cell_key;        // per body: level and cell
    uint32_t* gnext;           // per body: next body of its cell's list
    uint64_t* gt_key;          // cell table: packed cell key -> list head
    uint32_t* gt_head;
    uint32_t gt_cap;           // power of two
    // pairs
    uint64_t* pair_keys;       // this step's overlap list
    uint32_t* counters;        // see CTR_*
    PairState* pairs_cur;      // this step's pair states (index = overlap index)
    PairState* pairs_old;      // last step's
    PairHot* hot_cur;          // this step's PairHot per slot
    PairHot* hot_old;          // last step's
    ConvexState* conv_cur;     // general convex pairs: this step's extra state (index = overlap index)
    ConvexState* conv_old;     // last step's
    uint2* conv_work;          // this step's general convex pairs: (overlap index, last step's index or ~0),
                               // in three regions of pair_cap entries by the pair tester's state; a
                               // region holds pairs of two hulls from its front, the others from its back
    uint32_t* np_work;         // this step's closed-form pairs (overlap index): box-box from the front,
                               // the others (a sphere with a box or a sphere) from the back
    uint64_t* hash_keys_cur;
    uint32_t* hash_vals_cur;
    uint64_t* hash_keys_old;
    uint32_t* hash_vals_old;
    // islands (connected components of the constraint graph over dynamic bodies)
    uint32_t* cons;            // constraint -> pair slot (unordered)
    uint32_t* uf_parent;       // union-find forest over bodies
    uint32_t* isl_root;        // island (smallest dynamic body id) of every dynamic body
    uint32_t* cons_root;       // island of every constraint
    uint64_t* isl_count;       // per root: constraints << 32 | dynamic bodies (zero between steps)
    uint64_t* isl_start;       // exclusive scan of isl_count (n + 1)
    uint32_t* isl_cons;        // island-contiguous constraint list (pair slots)
    uint32_t* isl_bodies;      // island-contiguous body list
    uint32_t* body_local;      // index of a body inside its island's body list
    uint32_t* isl_medium;      // roots of islands solved by one workgroup each
    uint32_t* cls_list;        // small island roots, class k at [k * cls_cap, k * cls_cap + count)
    uint32_t cls_cap;          // per-class capacity (>= number of islands)
    // global path (large islands)
    uint32_t* gcons;           // pair slots of constraints in large islands
    int* color;                // per gcons entry
    uint64_t* prio;
    uint32_t* inc_count;       // body incidence counts (body_cap + 1)
    uint32_t* inc_offset;
    uint32_t* inc_list;        // 2 * pair_cap
    uint32_t* color_count;     // kMaxColors
    uint32_t* color_start;     // kMaxColors + 1
    uint32_t* color_cursor;
    uint32_t* scan_part;       // per-workgroup partial sums of the global path's scan (4096)
    uint32_t* ordered;         // solve order -> gcons index
    RowConst* grow_c;          // rows in solve order
    RowVar* grow_v;
    uint32_t* grow_pair;
    // deactivation (sleeping); the stages run only when `sleeping` (a dynamic
    // body that is not always active was added)
    uint32_t sleeping;
    uint32_t* comp_root;       // per dynamic body: its island at the end of the last step (label = a body id)
    uint32_t* lab_act;         // per label: an island member was activated (zeroed each step)
    uint64_t* lab_count;       // per label: candidates << 32 | members
    uint32_t* lab_min;         // per label: smallest present member (the island's turn position)
    uint32_t* lab_flags;       // per label: LF_* bits
    uint32_t* act2;            // per root of this step's union-find: the island is active
    uint32_t* uf2;             // union-find of the constraints left after stale pairs are removed
    uint32_t* deact_m;         // per body: its island's members if it is the island's smallest member (n + 1)
    uint32_t* deact_e;         // exclusive scan of deact_m (n + 1)
    uint32_t* wrange;          // per sub-world: first body, body count, generation (host-written)
    uint32_t* wcursor;         // per sub-world: last island looked at by the rotation
    uint32_t* wcurgen;         // per sub-world: generation the cursor belongs to
    uint32_t* wkey;            // per sub-world: furthest island looked at this step (turn position + 1)
    uint32_t* wsleepy;         // per sub-world: holds a dynamic body that is not always active (this step)
};

// Key -> pair slot table (open addressing, linear probing; empty key ~0).
__device__ __forceinline__ uint32_t hash_slot(uint64_t key, uint32_t cap) { return (uint32_t)mix64(key) & (cap - 1); }

__device__ __forceinline__ int hash_find(const uint64_t* keys, const uint32_t* vals, uint32_t cap, uint64_t key) {
    uint32_t h = hash_slot(key, cap);
    for (uint32_t probe = 0; probe < cap; ++probe) {
        uint64_t k = keys[h];
        if (k == key) return (int)vals[h];
        if (k == ~0ull) return -1;
        h = (h + 1) & (cap - 1);
    }
    return -1;
}

__device__ __forceinline__ bool hash_insert(uint64_t* keys, uint32_t* vals, uint32_t cap, uint64_t key, uint32_t val) {
    uint32_t h = hash_slot(key, cap);
    for (uint32_t probe = 0; probe < cap; ++probe) {
        unsigned long long prev = atomicCAS((unsigned long long*)&keys[h], ~0ull, (unsigned long long)key);
        if (prev == ~0ull || prev == key) {
            vals[h] = val;
            return true;
        }
        h = (h + 1) & (cap - 1);
    }
    return false;
}

// Label flags of the deactivation stage.
enum { LF_ACTIVE = 1u, LF_NODEACT = 2u, LF_CHECKED = 4u, LF_SLEEP = 8u };
constexpr uint32_t kNoCursor = 0xffffffffu;

// Whether a body is inactive for the broad and narrow phases (IsActive false):
// a sleeping dynamic body or a kinematic body whose flag is above 0.
__host__ __device__ __forceinline__ bool body_inactive(uint32_t f) {
    return (f & BF_DYNAMIC) ? (f & BF_ASLEEP) != 0 : (f & BF_KIN_ACTIVE) == 0;
}

// A kinematic member's flag after its candidacy update
// (SimulationIslandMember.cs:96-139): inactive (> 0) stays; at zero velocity
// -1 -> 0 -> 1; moving (or always active) -1.
__host__ __device__ __forceinline__ float kinematic_flag_next(float flag, float velocity, bool always_active) {
    if (flag > 0.0f) return flag;
    if (velocity == 0.0f && !always_active) return flag == 0.0f ? 1.0f : 0.0f;
    return -1.0f;
}

constexpr int kMaxColors = 128;
constexpr uint32_t kLevelList = 4;
constexpr uint32_t kMaxTreeLevels = 32;

// counters[] indices
enum {
    CTR_PAIRS = 0,
    CTR_CONS = 1,         // CONS..MAXPEN: last step's statistics, filled on request
    CTR_COLORS = 2,
    CTR_CONTACTS = 3,
    CTR_MAXPEN = 4,       // float bits (non-negative), max via atomicMax on uint
    CTR_OVERFLOW = 5,     // pair/hash capacity overflow flag
    CTR_ROUND = 6,        // colouring rounds
    CTR_CONVEX = 7,       // general convex pairs of two hulls queued for k_convex this step: Separated,
    CTR_CONVEX_SHALLOW = 12,  // Shallow and
    CTR_CONVEX_DEEP = 14,     // Deep tester state (the other general convex pairs: CTR_CONVEX_OTHER0 + state)
    CTR_UNCOLORED0 = 8,   // 8..10: uncoloured counts of rotating colouring rounds
    CTR_GCONS = 11,       // constraints on the global path
    CTR_ISL_MEDIUM = 13,
    CTR_GBAR = 15,        // arrivals at the global solver's grid barriers
    CTR_CLS0 = 16,        // 16..79: small islands per packing class
    CTR_SIDE0 = 80,       // 80..143: bodies per half-octave bin of the largest AABB side
    CTR_GBAR_ABORT = 144, // set when a grid barrier wait ran out of time
    CTR_CONVEX_OTHER0 = 145,  // 145..147: general convex pairs with at most one hull, per tester state
    CTR_NP_BOX = 148,     // closed-form pairs listed for k_narrowphase: box-box
    CTR_NP_SINGLE = 149,  // and the others
    CTR_STICKY = 150,     // STICKY_* bits of every step since the last bepu_clear_error (never reset by a step)
    CTR_PAIRS_PREV = 151, // the last step's pair count (copied by the step reset)
    CTR_DEAD = 152,       // frozen pairs removed as stale this step
    CTR_STEPS = 153,      // steps started (the step reset adds one; never cleared)
    CTR_MESH = 154,       // mesh pairs listed for k_mesh this step
    CTR_MESH_TESTERS = 155,  // statistics: live triangle testers of the last step's mesh pairs
    CTR_COUNT = 156
};

// Sticky error bits: the start of each step folds the previous step's
// overflow flag and grid barrier abort into CTR_STICKY before clearing them.
// STICKY_NONFINITE is not fatal: it tells that some sub-world was evicted
// (its flag in wfault) while the others go on.
enum { STICKY_OVERFLOW = 1u, STICKY_GBAR_ABORT = 2u, STICKY_NONFINITE = 4u };

// Every counter sits on its own 256-byte line: device atomics on one line
// serialise (about 10 ns each on MI355X), atomics on different lines do not.
constexpr uint32_t kCtrStride = 64;
__host__ __device__ constexpr uint32_t ctr(uint32_t i) { return i * kCtrStride; }

// Index of the step being run (the world's step count when it started).  It
// is kept in device memory, not passed by the host, so that the launches of
// one step replay unchanged for any later step.
__device__ __forceinline__ uint32_t step_index(const DeviceWorld& w) { return w.counters[ctr(CTR_STEPS)] - 1u; }

// Pair key with world-local body ids: the colouring priority of a constraint
// depends only on its own world, so a world solves identically alone or
// batched with others.
__device__ __forceinline__ uint64_t local_pair_key(const DeviceWorld& w, uint64_t key, uint32_t a) {
    uint64_t base = w.wbase[a];
    return key - ((base << 32) | base);
}

// A joint in the solver: not dead, and the island of a dynamic connection
// active (SolverUpdateable.UpdateSolverActivity, SolverUpdateable.cs:202-218).
__device__ __forceinline__ bool solver_joint(const DeviceWorld& w, const JointState& j) {
    if (j.flags & JF_DEAD) return false;
    const uint32_t fa = w.bflags[j.a], fb = w.bflags[j.b];
    return ((fa & BF_DYNAMIC) && !(fa & BF_ASLEEP)) || ((fb & BF_DYNAMIC) && !(fb & BF_ASLEEP));
}
// A joint joins the islands of its connections when both are dynamic.
__device__ __forceinline__ bool joint_edge(const DeviceWorld& w, const JointState& j) {
    return !(j.flags & JF_DEAD) && (w.bflags[j.a] & BF_DYNAMIC) && (w.bflags[j.b] & BF_DYNAMIC);
}
// The two bodies of solver constraint `id` (a pair slot or pair_cap + joint).
__device__ __forceinline__ void cons_ends(const DeviceWorld& w, uint32_t id, uint32_t& a, uint32_t& b) {
    if (id >= w.pair_cap) {
        const JointState& j = w.joints[id - w.pair_cap];
        a = j.a;
        b = j.b;
    } else {
        const PairHot h = w.hot_cur[id];
        a = h.a;
        b = h.b;
    }
}
// Its colouring priority: mix64 of its world-local key.
__device__ __forceinline__ uint64_t cons_priority(const DeviceWorld& w, uint32_t id) {
    if (id >= w.pair_cap) return mix64(kJointKey | w.joints[id - w.pair_cap].local);
    const PairState& s = w.pairs_cur[id];
    return mix64(local_pair_key(w, s.key, s.a));
}

// Host-side launchers (one per stage), defined in the kernel sources.
void launch_body_pre(const DeviceWorld& w, hipStream_t st);
void launch_broadphase(const DeviceWorld& w, void* sort_temp, size_t sort_temp_bytes, hipStream_t st);
void launch_narrowphase(const DeviceWorld& w, uint32_t block, hipStream_t st);
// General convex pairs queued by the narrowphase (GJK / MPR).
void launch_pair_prep(const DeviceWorld& w, hipStream_t st);
// `lds_mode` 1: each pair's hull vertices are copied to LDS before its test
// (when no hull has more than kConvexStageVerts vertices); 0: read from HBM;
// 2: each step picks one of the two on the device from its pair counts.
void launch_convex(const DeviceWorld& w, uint32_t lds_mode, hipStream_t st);
// Mesh pairs queued by k_pair_prep (StaticMeshContactManifold.Update).
void launch_mesh(const DeviceWorld& w, hipStream_t st);
constexpr uint32_t kConvexStageVerts = 16;
// Last step's constraint, contact, colour and penetration statistics into the counters.
void launch_stats(const DeviceWorld& w, uint32_t npairs, hipStream_t st);
void launch_islands(const DeviceWorld& w, void* scan_temp, size_t scan_temp_bytes, hipStream_t st);
void launch_island_solve(const DeviceWorld& w, hipStream_t st);
void launch_global_solver(const DeviceWorld& w, hipStream_t st);
size_t island_scan_temp_bytes(uint32_t n);
void launch_integrate(const DeviceWorld& w, hipStream_t st);
// Continuous position update (after launch_integrate, when w.ccd): times of
// impact of the pairs with a continuous body, then the linear motion of the
// passive and continuous bodies clamped to their smallest time of impact.
void launch_ccd(const DeviceWorld& w, hipStream_t st);
// Joints: their island edges in this step's union-find (after the narrow
// phase, before the activity pass and the islands); set-up of joint `id`
// from the bodies' current poses; its spring settings; the joints of removed
// bodies end (their present dynamic connections are activated).
void launch_joint_edges(const DeviceWorld& w, hipStream_t st);
void launch_joint_init(const DeviceWorld& w, uint32_t id, uint32_t a, uint32_t b, uint32_t local, float4 anchor,
                       hipStream_t st);
void launch_joint_settings(const DeviceWorld& w, uint32_t id, float stiffness, float damping, float max_corr,
                           float max_corr_sq, hipStream_t st);
void launch_joint_sweep(const DeviceWorld& w, hipStream_t st);
size_t broadphase_sort_temp_bytes(uint32_t n);
// Level offsets/counts of the implicit tree over n leaves (level 0 = leaves).
void tree_layout(uint32_t n, uint32_t* levels, uint32_t* off, uint32_t* cnt);
void launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t st);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
void launch_step_reset(const DeviceWorld& w, hipStream_t st);
// Body records -> structure-of-arrays; motion states of listed bodies -> 4 float4 each.
void launch_scatter_bodies(const DeviceWorld& w, const BodyRecord* recs, uint32_t n, hipStream_t st);
void launch_gather_states(const DeviceWorld& w, const uint32_t* ids, uint32_t n, float4* out, hipStream_t st);
// Motion states written between steps (Entity.MotionState setter): one record per listed body.
struct StateRecord {
    float4 pos, quat, linv, angv;
    uint32_t target;
    uint32_t fields;  // BEPU_STATE_* parts to write, or kStateUpdateMode
    uint32_t pad[2];  // kStateUpdateMode: pad[0] = the BF_CCD_MODE bits to set
};
// A StateRecord that sets the body's position update mode (and activates nothing).
constexpr uint32_t kStateUpdateMode = 1u << 16;
static_assert(sizeof(StateRecord) == 80, "StateRecord layout");
void launch_set_states(const DeviceWorld& w, const StateRecord* recs, uint32_t n, hipStream_t st);
// Marks bodies absent (Space.Remove): no AABB, no pairs, not dynamic from the next step on.
void launch_remove_bodies(const DeviceWorld& w, const uint32_t* ids, uint32_t n, hipStream_t st);
// Deactivation: the stage at the start of a step (candidacy, island turn),
// the activity pass after the narrow phase (merges wake islands, stale frozen
// pairs, split activation), and the activation of the bodies a removed body
// had constraints with.
void launch_deactivation(const DeviceWorld& w, void* scan_temp, size_t scan_temp_bytes, hipStream_t st);
void launch_activity(const DeviceWorld& w, hipStream_t st);
void launch_removal_wake(const DeviceWorld& w, uint32_t id, hipStream_t st);
size_t deactivation_scan_temp_bytes(uint32_t n);

}  // namespace bepu
