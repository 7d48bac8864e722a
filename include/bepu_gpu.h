/*
 * bepu_gpu.h -- C ABI of the MI355X (gfx950) rigid-body step library.
 *
 * The library replaces the contiguous block of BEPUphysics v1 stages that
 * Space.DoTimeStep runs between the deactivation candidacy update and the
 * position update (BEPUphysics/Space.cs:496-521):
 *
 *   ForceUpdater -> BoundingBoxUpdater -> BroadPhase -> NarrowPhase
 *   -> Solver -> PositionUpdater
 *
 * for convex entities (Box, Sphere, ConvexHull, Capsule) with the default
 * collision rules and material blending.  All state of a world (bodies,
 * pair cache, manifolds, accumulated impulses, friction memory) lives on the
 * device between calls; host buffers are only touched by the add/get calls.
 *
 * Conventions
 *  - Status codes replace the reference's exceptions (see bepu_status).
 *  - Buffers passed in are caller-owned host memory.
 *  - A world handle is not thread safe; like Space.Update it is driven by a
 *    single caller (BEPUphysics/Space.cs:527-531).
 *  - Body ids are dense uint32 in the order of the add calls.
 *  - Pair keys are (min(idA,idB) << 32) | max(idA,idB).
 *  - Quaternions are stored x, y, z, w (BEPUutilities/Quaternion.cs).
 */
#ifndef BEPU_GPU_H
#define BEPU_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEPU_ABI_VERSION 1

/* Maps 1:1 onto the exceptions the reference throws on the same misuse:
 * ArgumentException on double add / foreign remove (Space.cs:243-244,366-367),
 * InvalidOperationException for a bad mass (Entities/EntityBase.cs:879-880). */
typedef enum bepu_status {
    BEPU_OK = 0,
    BEPU_E_INVALID_ARG = 1,
    BEPU_E_ALREADY_ADDED = 2,
    BEPU_E_NOT_FOUND = 3,
    BEPU_E_OUT_OF_MEMORY = 4,
    BEPU_E_DEVICE = 5,
    BEPU_E_NONFINITE = 6,
    BEPU_E_CAPACITY = 7,
    BEPU_E_UNSUPPORTED = 8
} bepu_status;

/* One POD block per world (or per batch of worlds).  Every field defaults to
 * the reference's static default; bepu_settings_default fills them in.
 * Sources: Settings/CollisionDetectionSettings.cs:11-87,
 * Settings/CollisionResponseSettings.cs:12-43, Constraints/SolverSettings.cs:62,70,
 * Constraints/Solver.cs:15, DeactivationManagement/DeactivationManager.cs:18-20,60,
 * Entities/EntityBase.cs:1164-1165, Materials/MaterialManager.cs:29-39,
 * BEPUutilities/Toolbox.cs:20-25. */
typedef struct bepu_settings {
    float gravity[3];                  /* ForceUpdater.Gravity; demos use (0,-9.81,0) */
    float dt;                          /* TimeStepSettings.TimeStepDuration (1/60) */
    int32_t iteration_limit;           /* Solver.IterationLimit (10) */
    int32_t min_iteration_count;       /* SolverSettings.DefaultMinimumIterationCount (1) */
    float min_impulse;                 /* SolverSettings.DefaultMinimumImpulse (0.001) */
    float allowed_penetration;         /* CollisionDetectionSettings.AllowedPenetration (0.01) */
    float default_margin;              /* CollisionDetectionSettings.DefaultMargin (0.04) */
    float max_contact_distance;        /* CollisionDetectionSettings.maximumContactDistance (0.1) */
    float contact_invalidation_length_sq;   /* (0.01) */
    float contact_min_separation_sq;        /* (0.0009) */
    float bounce_velocity_threshold;   /* CollisionResponseSettings.BouncinessVelocityThreshold (1) */
    float max_penetration_recovery_speed;   /* (2) */
    float penetration_recovery_stiffness;   /* (0.2) */
    float static_friction_velocity_threshold; /* (0.2) */
    float twist_friction_factor;       /* (1) */
    float softness;                    /* (0.001) */
    int32_t use_stabilization;         /* DeactivationManager.UseStabilization (1) */
    float velocity_lower_limit;        /* DeactivationManager.VelocityLowerLimit (0.26) */
    float low_velocity_time_minimum;   /* DeactivationManager.LowVelocityTimeMinimum (1) */
    float linear_damping;              /* Entity default LinearDamping (0.03) */
    float angular_damping;             /* Entity default AngularDamping (0.15) */
    int32_t max_deactivation_attempts; /* DeactivationManager.MaximumDeactivationAttemptsPerFrame (100):
                                          island members looked at per step and sub-world when
                                          deciding which islands go to sleep */
    float core_shape_scaling;          /* MotionSettings.CoreShapeScaling (0.8; clamped to [0, 0.99]
                                          as its setter does, Settings/MotionSettings.cs:22-34):
                                          the continuous collision test sweeps spheres of this
                                          fraction of each shape's minimum radius */
} bepu_settings;

/* BEPUphysics/EntityStateManagement/MotionState.cs:10-55 */
typedef struct bepu_motion_state {
    float position[3];
    float orientation[4];  /* x, y, z, w */
    float linear_velocity[3];
    float angular_velocity[3];
} bepu_motion_state;

/* BEPUphysics/Materials/Material.cs:10,29,49; pair values are the product
 * of the two bodies' values (MaterialManager.DefaultMaterialBlender). */
typedef struct bepu_material {
    float static_friction;   /* default 1.0 */
    float kinetic_friction;  /* default 0.8 */
    float bounciness;        /* default 0.0 */
} bepu_material;

/* Per-body flags. */
#define BEPU_FLAG_ALWAYS_ACTIVE 1u      /* SimulationIslandMember.IsAlwaysActive */
#define BEPU_FLAG_NO_STABILIZATION 2u   /* SimulationIslandMember.AllowStabilization = false */
#define BEPU_FLAG_CCD_PASSIVE 4u        /* Entity.PositionUpdateMode = Passive (EntityBase.cs:1046-1068) */
#define BEPU_FLAG_CCD_CONTINUOUS 8u     /* Entity.PositionUpdateMode = Continuous */

/* Shape kinds. */
#define BEPU_SHAPE_BOX 0
#define BEPU_SHAPE_SPHERE 1
#define BEPU_SHAPE_CONVEX_HULL 2
#define BEPU_SHAPE_CAPSULE 3
/* A static triangle mesh (BroadPhaseEntries/StaticMesh.cs): a static
 * collidable, not an entity, that never moves and is never active. */
#define BEPU_SHAPE_STATIC_MESH 4
/* A terrain (BroadPhaseEntries/Terrain.cs): a static heightfield collidable. */
#define BEPU_SHAPE_TERRAIN 5

/* One contact of a manifold, in the manifold's contact-list order. */
typedef struct bepu_contact {
    float position[3];
    float normal[3];
    float penetration_depth;
    int32_t id;
} bepu_contact;

/* One pair's manifold (only pairs with at least one contact are reported). */
typedef struct bepu_manifold {
    uint64_t key;
    uint32_t body_a, body_b;       /* A/B as the pair handler orders them */
    int32_t count;
    int32_t kind;                  /* BEPU_PAIR_* */
    bepu_contact contacts[4];
} bepu_manifold;

#define BEPU_PAIR_BOX_BOX 0
#define BEPU_PAIR_BOX_SPHERE 1
#define BEPU_PAIR_SPHERE_SPHERE 2
#define BEPU_PAIR_CONVEX 3
/* A convex (box, capsule, hull) with a static mesh or a terrain
 * (StaticMeshConvexPairHandler, TerrainConvexPairHandler) and a sphere with
 * either (StaticMeshSpherePairHandler, TerrainSpherePairHandler): body_a is
 * the convex, body_b the mesh or terrain, the manifold holds up to four
 * contacts from its triangles with normals from the convex towards it. */
#define BEPU_PAIR_MESH_CONVEX 4
#define BEPU_PAIR_MESH_SPHERE 5

/* TriangleSidedness (BEPUutilities/TriangleSidedness.cs) of a static mesh's triangles. */
#define BEPU_SIDED_DOUBLE 0
#define BEPU_SIDED_CLOCKWISE 1
#define BEPU_SIDED_COUNTERCLOCKWISE 2

/* Accumulated impulses of one contact constraint after a step, listed in
 * solve order (colour, then key).  `normal[i]` is the i-th penetration
 * constraint (ContactPenetrationConstraint.accumulatedImpulse);
 * convex pairs fill `sliding` (SlidingFrictionTwoAxis) and `twist`
 * (TwistFrictionConstraint); non-convex pairs (sphere-sphere, box-sphere, mesh)
 * fill `friction[i]` (ContactFrictionConstraint). */
typedef struct bepu_impulses {
    uint64_t key;
    int32_t color;
    int32_t count;
    float normal[4];
    float friction[4];
    float sliding[2];
    float twist;
    float pad;
} bepu_impulses;

typedef struct bepu_step_stats {
    uint32_t bodies;
    uint32_t pairs;           /* broadphase overlaps */
    uint32_t manifolds;       /* pairs with >= 1 contact (solver constraints) */
    uint32_t contacts;
    uint32_t colors;
    uint32_t steps;           /* steps run since creation */
    float max_penetration;    /* deepest contact of the last step */
    uint32_t global_manifolds; /* constraints (manifolds and joints) in islands too large for one workgroup */
    uint32_t convex_pairs[3]; /* general convex pairs (any hull or capsule) tested in the last step,
                                 by pair tester state: separated, shallow, deep */
    uint32_t mesh_pairs;      /* static mesh and terrain pairs updated in the last step */
    uint32_t mesh_testers;    /* triangle pair testers those pairs kept (a sphere pair keeps none) */
} bepu_step_stats;

typedef struct bepu_world_s* bepu_world;

/* Fills every field with the reference defaults listed above
 * (BEPUphysicsDemos/ConfigurationHelper.cs:39-68 documents them). */
void bepu_settings_default(bepu_settings* out);

/* Space() constructor analogue (Space.cs:199-233).  `capacity_bodies` sizes
 * the device arrays; adding more bodies than that returns BEPU_E_CAPACITY. */
bepu_status bepu_world_create(const bepu_settings* settings, uint32_t capacity_bodies,
                              bepu_world* out);
/* The same with the capacity of broadphase pairs (manifold slots) given:
 * 0 takes the default max(1024, 4 * capacity_bodies).  Dense stacks and
 * close-packed piles hold up to about 11 pairs per body; a step that finds
 * more pairs than the capacity drops the excess and every later readback
 * returns BEPU_E_CAPACITY (see bepu_clear_error).  No reference counterpart
 * (the reference's pair lists grow on the managed heap). */
bepu_status bepu_world_create_ex(const bepu_settings* settings, uint32_t capacity_bodies,
                                 uint32_t capacity_pairs, bepu_world* out);
bepu_status bepu_world_destroy(bepu_world world);

/* Space.Add(new Box(motionState, width, height, length, mass)) -- Space.cs:241-358,
 * Entities/Prefabs/Box.cs:60-64.  mass <= 0 makes a kinematic body (BecomeKinematic),
 * a NaN/inf mass returns BEPU_E_INVALID_ARG (EntityBase.cs:879-880). */
bepu_status bepu_add_box(bepu_world world, const bepu_motion_state* state, float width,
                         float height, float length, float mass,
                         const bepu_material* material, uint32_t flags, uint32_t* id_out);
/* Space.Add(new Sphere(motionState, radius, mass)) -- Entities/Prefabs/Sphere.cs:70-74. */
bepu_status bepu_add_sphere(bepu_world world, const bepu_motion_state* state, float radius,
                            float mass, const bepu_material* material, uint32_t flags,
                            uint32_t* id_out);
/* Space.Add(new Capsule(motionState, length, radius, mass)) -- Entities/Prefabs/Capsule.cs:149-153,
 * CapsuleShape.cs:40-92: a segment of `length` along the local y axis, swept by `radius`
 * (the collision margin). */
bepu_status bepu_add_capsule(bepu_world world, const bepu_motion_state* state, float length,
                             float radius, float mass, const bepu_material* material,
                             uint32_t flags, uint32_t* id_out);

/* Convex hulls.  bepu_compute_hull is ConvexHullShape.ComputeDescription
 * (ConvexHullShape.cs:154-175): the hull of `points` (xyz triples;
 * ConvexHullHelper.GetConvexHull), its volume distribution and volume
 * (InertiaHelper.ComputeShapeDistribution) and its unique surface vertices
 * recentred on the volume centroid `center_out`.  Pure host computation; any
 * output pointer may be null; at most `capacity` vertices are written and
 * `nvertices_out` receives their number.  A point set without volume returns
 * BEPU_E_INVALID_ARG ("Point set is degenerate", ConvexHullHelper.cs:302-367). */
bepu_status bepu_compute_hull(const float* points, uint32_t npoints, float* vertices_out,
                              uint32_t capacity, uint32_t* nvertices_out,
                              float* volume_distribution_out /* 9, row-major */,
                              float* center_out /* 3 */, float* volume_out);
/* InertiaHelper.ComputeMinimumRadius (CollisionShapes/ConvexShapes/InertiaHelper.cs:529-570)
 * of the hull of `points` about its volume centroid: the distance from the centroid to the
 * nearest hull face plane.  A hull shape's MinimumRadius is this plus its collision margin
 * (ConvexHullShape.cs:170); continuous collision detection sweeps a fraction of it.
 * BEPU_E_INVALID_ARG for a point set without volume. */
bepu_status bepu_compute_hull_minimum_radius(const float* points, uint32_t npoints, float* radius_out);
/* new ConvexHullShape(points, out center) registered with the world: many bodies
 * may share one hull shape, as entities share a ConvexHullShape instance. */
bepu_status bepu_create_hull_shape(bepu_world world, const float* points, uint32_t npoints,
                                   uint32_t* shape_id_out, float* center_out /* 3, may be null */);
/* Space.Add(new ConvexHull(motionState, vertices, mass)) -- Entities/Prefabs/ConvexHull.cs:83-87:
 * the body's position is the motion state's (the shape's local origin is its centroid),
 * collision margin CollisionDetectionSettings.DefaultMargin. */
bepu_status bepu_add_convex_hull(bepu_world world, const bepu_motion_state* state,
                                 uint32_t hull_shape_id, float mass, const bepu_material* material,
                                 uint32_t flags, uint32_t* id_out);

/* Static meshes.  bepu_create_mesh_shape is new StaticMeshShape(vertices,
 * indices, worldTransform) (CollisionShapes/StaticMeshShape.cs:36-47,
 * DataStructures/TransformableMeshData.cs): `vertices` xyz triples, `indices`
 * three per triangle, `transform` the AffineTransform as 12 floats (its
 * Matrix3x3 M11..M33 row by row, then the translation; a vertex v maps to
 * v * M + t) or null for the identity.  The shape's MeshBoundingBoxTree
 * (DataStructures/MeshBoundingBoxTree.cs:60-69) is built here; a tree deeper
 * than 64 levels returns BEPU_E_CAPACITY.  Many static meshes may share one
 * shape.  bepu_add_static_mesh is Space.Add(new StaticMesh(...))
 * (BroadPhaseEntries/StaticMesh.cs): a static collidable with the given
 * TriangleSidedness (BEPU_SIDED_*; the reference's default DoubleSided),
 * ImproveBoundaryBehavior (default 1) and material.  It is not an entity:
 * state setters, activation, position update modes and joints refuse it
 * (BEPU_E_INVALID_ARG); bepu_remove removes it.  A mesh pair keeps at most 32
 * triangle testers and meets at most 32 contact candidates per step; a step
 * that needs more sets the sticky capacity error. */
bepu_status bepu_create_mesh_shape(bepu_world world, const float* vertices, uint32_t nvertices,
                                   const int32_t* indices, uint32_t nindices, const float* transform,
                                   uint32_t* shape_id_out);
bepu_status bepu_add_static_mesh(bepu_world world, uint32_t mesh_shape_id, int32_t sidedness,
                                 int32_t improve_boundary, const bepu_material* material, uint32_t* id_out);

/* Whole-world snapshots (no reference counterpart: the reference exposes
 * per-entity MotionState only, EntityStateManagement/MotionState.cs:10-55).
 * A snapshot holds every body's state, the pair cache with its manifolds,
 * accumulated impulses and pair tester state, the triangle tester tables,
 * islands, sleeping labels and turns, joints, counters, the settings and the
 * step count; stepping on from a restored snapshot gives the same bits as
 * stepping on from the moment it was taken.  bepu_snapshot_size gives the
 * bytes; bepu_save_snapshot writes them to caller memory (bodies staged since
 * the last step are uploaded first); bepu_restore_snapshot loads one into
 * the world it came from, or any world built the same way, provided the
 * structure (bodies and removals, sub-worlds, joints, shapes, capacities) is
 * unchanged: otherwise BEPU_E_INVALID_ARG and nothing is written. */
bepu_status bepu_snapshot_size(bepu_world world, uint64_t* bytes_out);
bepu_status bepu_save_snapshot(bepu_world world, void* buffer, uint64_t capacity, uint64_t* bytes_out);
bepu_status bepu_restore_snapshot(bepu_world world, const void* buffer, uint64_t bytes);

/* Terrains.  bepu_create_terrain_shape is new TerrainShape(heights,
 * organization) (CollisionShapes/TerrainShape.cs:60-72) together with the
 * Terrain's WorldTransform (BroadPhaseEntries/Terrain.cs:41-75): `heights`
 * column by column (heights[column * length + row], the reference's
 * heights[column, row]; `width` columns along local x, `length` rows along
 * local z, both at least 2), `organization` the QuadTriangleOrganization
 * (0 BottomLeftUpperRight, 1 BottomRightUpperLeft), `transform` the
 * AffineTransform as 12 floats (Matrix3x3 M11..M33 row by row, then the
 * translation; the sample (c, heights[c, r], r) maps to its v * M + t).  The
 * terrain's sidedness follows the transform's winding and its box is fixed
 * here (TerrainShape.GetBoundingBox).  bepu_add_terrain is Space.Add(new
 * Terrain(shape, transform)) (Terrain.cs:172-184) with ImproveBoundaryBehavior
 * (the reference's default 1) and material; its Thickness is 0 (the
 * reference's default), the thickness raycast is not offered.  A terrain is a
 * static collidable like a static mesh: it is not an entity, setters,
 * activation, update modes and joints refuse it, bepu_remove removes it, and
 * its pairs share the mesh pairs' limits. */
bepu_status bepu_create_terrain_shape(bepu_world world, const float* heights, uint32_t width, uint32_t length,
                                      int32_t organization, const float* transform /* 12 */,
                                      uint32_t* shape_id_out);
bepu_status bepu_add_terrain(bepu_world world, uint32_t terrain_shape_id, int32_t improve_boundary,
                             const bepu_material* material, uint32_t* id_out);

/* Batched worlds.  A world handle may hold many independent sub-worlds, each
 * the analogue of one Space: bodies of different sub-worlds never form pairs,
 * and every sub-world's result is bit-identical to stepping it alone (its
 * arithmetic, including the constraint colouring, depends on world-local ids
 * only).  bepu_begin_world starts a new sub-world whose bodies are those added
 * after the call (a sub-world without bodies is reused); sub-world 0 exists
 * from creation.  One bepu_step steps every sub-world.  No reference
 * counterpart: the reference steps one Space per Update call. */
bepu_status bepu_begin_world(bepu_world world, uint32_t* world_id_out);
bepu_status bepu_get_world_range(bepu_world world, uint32_t world_id, uint32_t* first_body_out,
                                 uint32_t* count_out);

/* Continuous batching.  One body of a world inserted in a single call. */
typedef struct bepu_body_desc {
    int32_t shape;            /* BEPU_SHAPE_* */
    float dims[3];            /* box: width, height, length; sphere: radius, 0, 0;
                                 capsule: length, radius, 0; hull: shape id (an integer), 0, 0;
                                 static mesh: mesh shape id, sidedness, improve boundary (0/1);
                                 terrain: terrain shape id, improve boundary (0/1), 0;
                                 a static mesh's or terrain's mass, state and flags are ignored */
    float mass;               /* <= 0: kinematic */
    bepu_motion_state state;
    bepu_material material;
    uint32_t flags;           /* BEPU_FLAG_* */
} bepu_body_desc;

/* Inserts a whole new sub-world at any time between steps: its bodies take
 * a free contiguous id range (one left by a removed world if large enough,
 * else the end of the arrays) and join the next step.  A body's first step
 * never picks up state left by a previous occupant of its id.  Returns the
 * world id and its first body id.  bepu_remove_world retires a sub-world:
 * its bodies leave every stage from the next step on and its id range and
 * world id are reused.  bepu_get_motion_states_range reads one world's
 * bodies.  No reference counterpart (each request would be its own Space). */
bepu_status bepu_insert_world(bepu_world world, const bepu_body_desc* bodies, uint32_t count,
                              uint32_t* world_id_out, uint32_t* first_body_out);
bepu_status bepu_remove_world(bepu_world world, uint32_t world_id);
bepu_status bepu_get_motion_states_range(bepu_world world, uint32_t first_body, uint32_t count,
                                         bepu_motion_state* out);
/* The same for many worlds per call (one device copy each way): worlds are
 * inserted in order until one does not fit; `inserted_out` tells how many
 * went in (BEPU_E_CAPACITY when not all).  Bodies of all worlds are packed
 * back to back in `bodies`, counts[i] per world. */
bepu_status bepu_insert_worlds(bepu_world world, const bepu_body_desc* bodies, const uint32_t* counts,
                               uint32_t nworlds, uint32_t* world_ids_out, uint32_t* first_bodies_out,
                               uint32_t* inserted_out);
bepu_status bepu_remove_worlds(bepu_world world, const uint32_t* world_ids, uint32_t count);
bepu_status bepu_get_motion_states_ranges(bepu_world world, const uint32_t* first_bodies,
                                          const uint32_t* counts, uint32_t nranges, bepu_motion_state* out);

/* The same readback, asynchronous: bepu_request_motion_states_ranges queues
 * it behind the steps already queued and returns a ticket at once (the ids
 * are taken from the arguments before it returns);
 * bepu_collect_motion_states waits for that readback alone and writes its
 * states to `out` in request order (the sum of `counts` entries).  Up to 8
 * readbacks may be outstanding (BEPU_E_CAPACITY beyond); a collected ticket
 * is gone (BEPU_E_NOT_FOUND).  The collect status carries the error flags as
 * they stood when the readback ran.  With inserts that never wait for the
 * device either, a serving loop keeps its next step queued while it reads
 * results (no reference counterpart). */
bepu_status bepu_request_motion_states_ranges(bepu_world world, const uint32_t* first_bodies,
                                              const uint32_t* counts, uint32_t nranges, uint32_t* ticket_out);
bepu_status bepu_collect_motion_states(bepu_world world, uint32_t ticket, bepu_motion_state* out);

/* Waits until every queued step of the world has finished on the device.
 * Returns BEPU_E_CAPACITY when any step since the last bepu_clear_error
 * overflowed a capacity (pairs, pair table, grid table, colours), and
 * BEPU_E_DEVICE when the global solver's grid barrier gave up waiting (its
 * workgroups were not all resident).  The same status comes back from every
 * readback (whose outputs are still filled), and bepu_step refuses to run
 * while it is set: the flags are sticky until bepu_clear_error. */
bepu_status bepu_synchronize(bepu_world world);
/* Acknowledges and clears the sticky error of the world (no reference
 * counterpart: the reference's lists grow instead of overflowing). */
bepu_status bepu_clear_error(bepu_world world);

/* The non-finite guard.  The reference validates positions, orientations and
 * velocities with MathChecker in its debug builds (BEPUutilities/MathChecker.cs:23-50,
 * NotFiniteNumberException).  Here a sub-world whose body state is no longer
 * finite after a step is evicted: from the next step on every body of that
 * sub-world leaves the simulation (as if removed), so the other sub-worlds of
 * a batch go on bit-identical.  bepu_get_faulted_worlds lists the evicted
 * sub-worlds still live (sorted); synchronous readbacks whose bodies include
 * one of them return BEPU_E_NONFINITE (outputs still filled: an evicted body
 * keeps the non-finite state it reached); an asynchronous collect reports the
 * states as they are and no status of its own for this, so that it stays
 * free of device waits.  Stepping continues, and bepu_clear_error keeps the
 * evictions.  Bodies appended to an evicted sub-world are evicted with it; a
 * sub-world id reused by an insertion starts unfaulted. */
bepu_status bepu_get_faulted_worlds(bepu_world world, uint32_t* world_ids, uint32_t capacity,
                                    uint32_t* count_out);

/* Deactivation (sleeping), DeactivationManagement/DeactivationManager.cs:
 * a body that is not BEPU_FLAG_ALWAYS_ACTIVE becomes a deactivation
 * candidate after low_velocity_time_minimum below velocity_lower_limit, and
 * its simulation island (the bodies its constraints connect) sleeps once
 * every member is a candidate, checked for up to max_deactivation_attempts
 * island members per step and sub-world, in turn.  A sleeping island keeps
 * its state and its pairs and is not stepped until it is activated: by a new
 * constraint with an awake body, by an active kinematic body touching it, by
 * the removal of a body it touches, by a state setter or by bepu_activate.
 *
 * EntityBase.ActivityInformation.IsActive (SimulationIslandMember.cs:227-250)
 * of every body: 1 active, 0 asleep, an inactive kinematic body or removed. */
bepu_status bepu_get_activity(bepu_world world, uint8_t* active_out, uint32_t capacity, uint32_t* count_out);
/* EntityBase.ActivityInformation.Activate() (SimulationIslandMember.cs:256-277)
 * for the listed bodies: their islands are active at the next step. */
bepu_status bepu_activate(bepu_world world, const uint32_t* ids, uint32_t n);

/* Space.Remove(entity) -- Space.cs:364-482: the body leaves every stage from
 * the next step on, and its pairs and constraints end with it.  An id that is
 * unknown or already removed returns BEPU_E_NOT_FOUND (the reference's
 * ArgumentException, Space.cs:366-367).  Ids are not reused. */
bepu_status bepu_remove(bepu_world world, uint32_t id);
/* Entity state setters for each of `n` listed bodies; `fields` selects which
 * parts of states[i] are written, as the reference's setters do one at a time:
 * Position (Entities/EntityBase.cs:51-62), Orientation (normalised,
 * EntityBase.cs:68-84), LinearVelocity (:181-195), AngularVelocity (:125-140);
 * BEPU_STATE_ALL is the MotionState setter (:218-235).  Takes effect at the
 * next step; pair caches, warm-start impulses and the stabilization state are
 * kept.  BEPU_E_NOT_FOUND for a removed or unknown id, BEPU_E_INVALID_ARG for
 * a non-finite written value or no field (nothing is written then). */
#define BEPU_STATE_POSITION 1u
#define BEPU_STATE_ORIENTATION 2u
#define BEPU_STATE_LINEAR_VELOCITY 4u
#define BEPU_STATE_ANGULAR_VELOCITY 8u
#define BEPU_STATE_ALL 15u
bepu_status bepu_set_motion_states(bepu_world world, const uint32_t* ids, const bepu_motion_state* states,
                                   uint32_t n, uint32_t fields);
/* Continuous collision detection, PositionUpdating/ContinuousPositionUpdater.cs:76-160.
 * Entity.PositionUpdateMode setter (EntityBase.cs:1050-1068) for each listed body:
 *   BEPU_POSITION_UPDATE_DISCRETE    (the default) the position moves by v dt;
 *   BEPU_POSITION_UPDATE_PASSIVE     the position moves by v dt toi, toi the smallest time
 *                                    of impact of the body's pairs (1 when none);
 *   BEPU_POSITION_UPDATE_CONTINUOUS  as passive, and its pairs compute times of impact
 *                                    (ConvexPairHandler.UpdateTimeOfImpact,
 *                                    ConvexPairHandler.cs:37-111): a sphere of
 *                                    core_shape_scaling x the shape's minimum radius is cast
 *                                    along the relative motion against the other shape
 *                                    (GJKToolbox.CCDSphereCast, GJKToolbox.cs:470-495).
 * Orientation always integrates fully.  Takes effect at the next step and does not
 * activate the body.  BEPU_E_NOT_FOUND for a removed or unknown id, BEPU_E_INVALID_ARG
 * for another mode (nothing is written then). */
#define BEPU_POSITION_UPDATE_DISCRETE 0
#define BEPU_POSITION_UPDATE_PASSIVE 1
#define BEPU_POSITION_UPDATE_CONTINUOUS 2
bepu_status bepu_set_position_update_mode(bepu_world world, const uint32_t* ids, uint32_t n, int32_t mode);

/* Two-body joints in the solver (Constraints/TwoEntity/Joints/).  A joint is a
 * constraint of the islands, the colouring and the solve like a contact
 * manifold constraint; joined bodies are one island (they sleep and wake
 * together).
 * new BallSocketJoint(connectionA, connectionB, anchor) added to the space
 * (BallSocketJoint.cs:57-63): keeps the world point `anchor` (xyz) of both
 * bodies together, from their poses after the steps already queued.  Both
 * bodies belong to one sub-world, at least one is dynamic (the reference
 * throws "Cannot constrain two kinematic bodies"; a null connection, the
 * reference's WorldEntity, is given as a kinematic body).  SpringSettings
 * default to stiffness 600000 and damping 90000, no corrective velocity limit.
 * BEPU_E_NOT_FOUND for an unknown or removed body, BEPU_E_INVALID_ARG for two
 * kinematic bodies, one body twice, different sub-worlds or a non-finite
 * anchor, BEPU_E_CAPACITY beyond the world's body capacity in joints.  A
 * joint of a removed body ends (its other connection is activated). */
bepu_status bepu_add_ball_socket(bepu_world world, uint32_t body_a, uint32_t body_b, const float* anchor,
                                 uint32_t* joint_id_out);
/* SpringSettings.Stiffness / Damping (clamped at 0; both 0 is BEPU_E_INVALID_ARG,
 * SpringSettings.cs:61-124) and Joint.MaxCorrectiveVelocity (Joint.cs:24-38;
 * FLT_MAX or more: no limit) of a joint, from the next step on. */
bepu_status bepu_set_joint_settings(bepu_world world, uint32_t joint_id, float stiffness, float damping,
                                    float max_corrective_velocity);
/* BallSocketJoint.TotalImpulse (the accumulated impulse, BallSocketJoint.cs:146-149)
 * of every joint in id order, three floats each; `count_out` gets the number of
 * joints (a NULL `impulses` only counts). */
bepu_status bepu_get_joint_impulses(bepu_world world, float* impulses, uint32_t capacity, uint32_t* count_out);

/* The settings block of a world, readable and writable between steps:
 * Solver.IterationLimit (Constraints/Solver.cs:19), ForceUpdater.Gravity
 * (OtherSpaceStages/ForceUpdater.cs:19-29), TimeStepSettings.TimeStepDuration
 * (TimeStepSettings.cs:19) and the rest.  Every later step uses the new
 * values; what an entity took at creation (its damping, a hull's margin)
 * keeps its value, as the reference's entities do. */
bepu_status bepu_get_settings(bepu_world world, bepu_settings* out);
bepu_status bepu_set_settings(bepu_world world, const bepu_settings* settings);

/* Space.Update() x n_steps (Space.cs:527-531). */
bepu_status bepu_step(bepu_world world, int32_t n_steps);

/* Entity.MotionState for every body, in id order. */
bepu_status bepu_get_motion_states(bepu_world world, bepu_motion_state* out, uint32_t capacity,
                                   uint32_t* count_out);
/* BroadPhase.Overlaps of the last step as a sorted set of pair keys
 * (BroadPhaseSystems/BroadPhase.cs:34-41). */
bepu_status bepu_get_pairs(bepu_world world, uint64_t* keys, uint32_t capacity,
                           uint32_t* count_out);
/* NarrowPhase pair manifolds with >= 1 contact, sorted by key. */
bepu_status bepu_get_manifolds(bepu_world world, bepu_manifold* out, uint32_t capacity,
                               uint32_t* count_out);
/* Solver constraints of the last step in solve order. */
bepu_status bepu_get_impulses(bepu_world world, bepu_impulses* out, uint32_t capacity,
                              uint32_t* count_out);
bepu_status bepu_get_stats(bepu_world world, bepu_step_stats* out);

/* Stage timers.  With timing enabled every step of a bepu_step call records
 * HIP events between its stages on the world's stream (the events cost device
 * time of their own, and the general convex tests then run after the
 * closed-form narrowphase instead of beside it, so that each stage's time is
 * its own); with timing off no events are recorded and
 * bepu_get_stage_times reports the last timed call.  It returns the average device time per
 * stage (milliseconds) over those steps; `names` receives a pointer to a
 * static, comma separated list of stage names matching `ms`.  No reference
 * counterpart: it plays the role of the reference's PROFILE stage timers
 * (ProcessingStage.cs:26-59, Space.cs:486-520). */
bepu_status bepu_set_timing(bepu_world world, int32_t enable);
bepu_status bepu_get_stage_times(bepu_world world, float* ms, uint32_t capacity,
                                 uint32_t* count_out, const char** names);

/* Broadphase algorithm of a world; replaces the Space.BroadPhase setter
 * (Space.cs:110-134), which swaps DynamicHierarchy for SortAndSweep1D or
 * BruteForce.  Both kinds produce the same pair set.  Takes effect at the
 * next step. */
#define BEPU_BROADPHASE_GRID 0  /* hierarchical hash grid (default) */
#define BEPU_BROADPHASE_TREE 1  /* implicit tree over Morton order */
bepu_status bepu_set_broadphase(bepu_world world, int32_t kind);

/* Launch configuration of a world's step and its autotuner.  None of these
 * choices changes any arithmetic: every choice gives bit-identical results,
 * so the tuner times real steps of the world itself.  No reference
 * counterpart (the reference's stages have no launch shapes); it is the
 * autotuner over block size and colouring granularity of the batched form.
 *   narrowphase_block  workgroup size of the closed-form narrowphase: 64, 128 (default) or 256
 *   convex_lds         1: each general convex pair's hull vertices are copied to LDS before
 *                      its test (only when no hull has more than 16 vertices); 0 (default):
 *                      support scans read them from HBM; 2: each step picks one of the two
 *                      on the device (staging while Shallow pairs are the minority)
 *   convex_side_stream 1 (default): the closed-form narrowphase runs on a second stream beside
 *                      the general convex tests; 0: both in turn on the world's stream
 *   small_island_max   islands of at most this many constraints (32 or 64, default 64) are
 *                      packed several to a wavefront; larger ones up to 256 take a workgroup each
 *   step_graph         1: the launches of a step are captured once into a HIP graph (one per
 *                      parity of the step count, captured again after the world or this
 *                      configuration changes) and replayed, which saves the host its launches;
 *                      0 (default): launched one by one.  Steps with stage timing and the tree
 *                      broadphase are always launched one by one; the tuner tries replay only
 *                      on steps shorter than 0.25 ms
 *   ms_per_step        (filled by bepu_autotune) the chosen configuration's time */
typedef struct bepu_tuning {
    uint32_t narrowphase_block;
    uint32_t convex_lds;
    uint32_t convex_side_stream;
    uint32_t small_island_max;
    uint32_t step_graph;
    float ms_per_step;
} bepu_tuning;

bepu_status bepu_get_tuning(bepu_world world, bepu_tuning* out);
/* Takes effect at the next step; BEPU_E_INVALID_ARG for a value outside the sets above. */
bepu_status bepu_set_tuning(bepu_world world, const bepu_tuning* tuning);
/* Steps the world while it searches the configuration, one choice at a time
 * (coordinate descent from the current configuration): every candidate value
 * runs one untimed step and then `trial_steps` timed steps in each of two
 * interleaved rounds, timed by HIP
 * events, and the fastest is kept if it beats the current value by more than
 * 2%.  Runs 2 * (trial_steps + 1) steps of the world
 * per candidate value (24 * (trial_steps + 1) with every choice open; the convex
 * choices are open only in worlds with hulls or capsules), leaves the best
 * configuration set, reports it in `out` and the steps run in `steps_run`
 * (either may be NULL). */
bepu_status bepu_autotune(bepu_world world, uint32_t trial_steps, bepu_tuning* out, uint32_t* steps_run);

/* Readback of body ranges that allocates nothing, for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation).  No reference counterpart.  The same result as
 * bepu_get_motion_states_ranges, staged through a fixed device array of the
 * library (one per GPU, held until the process ends) instead of buffers it
 * allocates: at most BEPU_NOALLOC_MAX_RANGES ranges and
 * BEPU_NOALLOC_MAX_BODIES bodies per call (else BEPU_E_CAPACITY), and no
 * bodies added since the last step (their upload could allocate:
 * BEPU_E_INVALID_ARG).  Every world on one GPU shares that array, which keeps
 * the last call's contents; calls from one process are serialised, and each
 * waits for its copies before it returns. */
#define BEPU_NOALLOC_MAX_RANGES 16384
#define BEPU_NOALLOC_MAX_BODIES 262144
bepu_status bepu_get_motion_states_ranges_noalloc(bepu_world world, const uint32_t* first_bodies,
                                                  const uint32_t* counts, uint32_t nranges,
                                                  bepu_motion_state* out);

/* Library-wide information: 1 if a HIP device is usable. */
int32_t bepu_device_available(void);
const char* bepu_status_string(bepu_status status);

#ifdef __cplusplus
}
#endif

#endif /* BEPU_GPU_H */
