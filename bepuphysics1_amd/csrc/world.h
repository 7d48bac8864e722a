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
    float pad0;
    float sllin[6];       // SlidingFrictionTwoAxis linearA (remembered across steps)
    float pad1[2];
};

enum { PF_PREV_COLLIDING = 1u, PF_FRICTION_ACTIVE = 2u };

// Solver rows of all constraints in solve order; index = position in order.
struct SolverRows {
    int2* bodies;       // A, B
    int* meta;          // kind | npen << 4 | aDyn << 8 | bDyn << 9
    int* pair;          // pair slot
    uint32_t* act;      // bit k (k < 10): sub-constraint k active in solver; bit 31: group active
    uint32_t* zero;     // 3-bit zero-impulse iteration counter per sub-constraint
    float4* pen;        // [k][3][cap]: (lin.xyz, softness) (angA.xyz, bias) (angB.xyz, v2i)
    float* penacc;      // [k][cap]
    float4* sl;         // [9][cap]: lin0(xyz,friction) lin1 angA0 angA1 angB0 angB1 v2i(4) ra rb
    float2* slacc;
    float4* tw;         // [3][cap]: (ang.xyz, v2i) (lever0..3) (friction, count, -, -)
    float* twacc;
    float4* fr;         // [k][3][cap]: (lin.xyz, friction) (angA.xyz, v2i) (angB.xyz, -)
    float* fracc;       // [k][cap]
};

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
    float4* shape;             // half extents (box) or (radius,0,0) ; w = shape kind
    float4 *bmin, *bmax;
    float4* stab;              // prev velocity, time below, slowing, allow stabilization
    float2* damp;              // linear, angular damping
    float4* mat;               // static, kinetic, bounce
    uint32_t* bflags;          // bit0 dynamic
    // broadphase
    uint32_t* morton;
    uint32_t* morton_sorted;
    uint32_t* leaf_ids;        // body id per sorted leaf
    uint32_t* leaf_ids_tmp;
    int2* node_child;          // internal node children (>=0 internal, <0 leaf ~idx)
    int* node_parent;          // parent for internal + leaf nodes (leaves offset by n-1)
    float4 *node_min, *node_max;
    uint32_t* node_visit;
    uint32_t* bounds;          // ordered-int encoded scene bounds (6)
    // pairs
    uint64_t* pair_keys;       // this step's overlap list
    uint32_t* counters;        // see CTR_*
    PairState* pairs_cur;      // this step's pair states (index = overlap index)
    PairState* pairs_old;      // last step's
    uint64_t* hash_keys_cur;
    uint32_t* hash_vals_cur;
    uint64_t* hash_keys_old;
    uint32_t* hash_vals_old;
    // solver
    uint32_t* cons;            // constraint -> pair slot (unordered)
    int* color;                // per constraint
    uint64_t* prio;            // per constraint
    uint32_t* inc_count;       // body incidence counts (body_cap + 1)
    uint32_t* inc_offset;
    uint32_t* inc_list;        // 2 * pair_cap
    uint32_t* color_count;     // kMaxColors
    uint32_t* color_start;     // kMaxColors + 1
    uint32_t* color_cursor;
    uint32_t* ordered;         // solve order -> constraint index
    SolverRows rows;
};

constexpr int kMaxColors = 128;

// counters[] indices
enum {
    CTR_PAIRS = 0,
    CTR_CONS = 1,
    CTR_COLORS = 2,
    CTR_CONTACTS = 3,
    CTR_MAXPEN = 4,       // float bits (non-negative), max via atomicMax on uint
    CTR_OVERFLOW = 5,     // pair/hash capacity overflow flag
    CTR_ROUND = 6,        // colouring rounds
    CTR_UNCOLORED0 = 8,   // 8..10: uncoloured counts of rotating colouring rounds
    CTR_COUNT = 16
};

// Host-side launchers (one per stage), defined in the kernel sources.
void launch_body_pre(const DeviceWorld& w, hipStream_t st);
void launch_broadphase(const DeviceWorld& w, void* sort_temp, size_t sort_temp_bytes, hipStream_t st);
void launch_narrowphase(const DeviceWorld& w, hipStream_t st);
void launch_solver_full(const DeviceWorld& w, void* scan_temp, size_t scan_temp_bytes, hipStream_t st);
size_t solver_scan_temp_bytes(uint32_t n);
void launch_integrate(const DeviceWorld& w, hipStream_t st);
size_t broadphase_sort_temp_bytes(uint32_t n);
void launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t st);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);

}  // namespace bepu
