// world.cpp -- C ABI of the gfx950 step library (include/bepu_gpu.h):
// world lifetime, body registration, the per-step launch sequence and
// readback of pairs, manifolds, impulses and stats.
//
// A step is a fixed sequence of kernel launches on the world's stream with no
// host synchronisation inside it: every kernel reads this step's counts from
// device memory, so bepu_step(n) queues n steps back to back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <map>
#include <vector>

#include "../../include/bepu_gpu.h"
#include "bodies_host.h"
#include "hull.h"
#include "mesh.h"
#include "mesh_tree.h"
#include "solver_rows.h"
#include "world_host.h"

using namespace bepu;

namespace {

enum Stage { ST_BODY = 0, ST_BROAD, ST_NARROW, ST_CONVEX, ST_MESH, ST_ISLANDS, ST_ISLAND_SOLVE, ST_GLOBAL_SOLVE, ST_INTEGRATE,
             ST_COUNT };
const char* kStageNames = "body_pre,broadphase,narrowphase,convex,mesh,islands,island_solve,global_solve,integrate";

template <typename T>
bepu_status dalloc(T** p, size_t n) {
    if (n == 0) n = 1;
    if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) return BEPU_E_OUT_OF_MEMORY;
    if (hipMemset(*p, 0, n * sizeof(T)) != hipSuccess) return BEPU_E_DEVICE;
    return BEPU_OK;
}

uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return (uint32_t)p;
}

}  // namespace


static bepu_status check(hipError_t e) { return e == hipSuccess ? BEPU_OK : BEPU_E_DEVICE; }

// The live sub-worlds whose state became non-finite, sorted (reads the
// per-world flags only once a fault was seen: STICKY_NONFINITE).
static bepu_status faulted_worlds(bepu_world w, std::vector<uint32_t>* out) {
    out->clear();
    if (!(w->sticky & STICKY_NONFINITE)) return BEPU_OK;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    std::vector<uint32_t> f(w->world_first.size(), 0);
    if (!f.empty() && hipMemcpy(f.data(), w->d.wfault, f.size() * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return BEPU_E_DEVICE;
    for (uint32_t k = 0; k < f.size(); ++k)
        if (f[k] && w->world_live[k]) out->push_back(k);
    return BEPU_OK;
}

// Whether any of the body ranges holds a body of a faulted world.
static bool ranges_faulted(bepu_world w, const uint32_t* firsts, const uint32_t* counts, uint32_t n) {
    std::vector<uint32_t> bad;
    if (faulted_worlds(w, &bad) != BEPU_OK) return true;
    for (uint32_t k : bad) {
        uint64_t a0 = w->world_first[k], a1 = a0 + w->world_count[k];
        for (uint32_t r = 0; r < n; ++r)
            if ((uint64_t)firsts[r] < a1 && a0 < (uint64_t)firsts[r] + counts[r]) return true;
    }
    return false;
}

// Overflow (pair, hash, colour or list capacity) and grid barrier aborts of
// any step since the last bepu_clear_error, as a status.
bepu_status bepu::sticky_status(uint32_t bits) {
    if (bits & STICKY_GBAR_ABORT) return BEPU_E_DEVICE;
    if (bits & STICKY_OVERFLOW) return BEPU_E_CAPACITY;
    return BEPU_OK;
}

// Waits for the world's stream and folds the device's error flags (the
// sticky word and the last step's own flags) into w->sticky.  Allocates nothing.
bepu_status bepu::poll_errors(bepu_world w) {
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    uint32_t e[3] = {0, 0, 0};
    const uint32_t idx[3] = {CTR_STICKY, CTR_OVERFLOW, CTR_GBAR_ABORT};
    for (int i = 0; i < 3; ++i)
        if (hipMemcpyAsync(&e[i], w->d.counters + ctr(idx[i]), sizeof(uint32_t), hipMemcpyDeviceToHost, w->stream) !=
            hipSuccess)
            return BEPU_E_DEVICE;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    w->sticky |= e[0] | (e[1] ? (uint32_t)STICKY_OVERFLOW : 0u) | (e[2] ? (uint32_t)STICKY_GBAR_ABORT : 0u);
    return sticky_status(w->sticky);
}

extern "C" {

void bepu_settings_default(bepu_settings* s) {
    if (!s) return;
    s->gravity[0] = 0;
    s->gravity[1] = -9.81f;
    s->gravity[2] = 0;
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

int32_t bepu_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

const char* bepu_status_string(bepu_status s) {
    switch (s) {
    case BEPU_OK: return "ok";
    case BEPU_E_INVALID_ARG: return "invalid argument";
    case BEPU_E_ALREADY_ADDED: return "already added";
    case BEPU_E_NOT_FOUND: return "not found";
    case BEPU_E_OUT_OF_MEMORY: return "out of device memory";
    case BEPU_E_DEVICE: return "device error";
    case BEPU_E_NONFINITE: return "non-finite state";
    case BEPU_E_CAPACITY: return "capacity exceeded";
    case BEPU_E_UNSUPPORTED: return "unsupported";
    }
    return "unknown";
}

bepu_status bepu_world_destroy(bepu_world w) {
    if (!w) return BEPU_E_INVALID_ARG;
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    for (void* p : w->allocations) (void)hipFree(p);
    if (w->stage) (void)hipFree(w->stage);
    if (w->id_stage) (void)hipFree(w->id_stage);
    if (w->state_stage) (void)hipFree(w->state_stage);
    if (w->sstage) (void)hipFree(w->sstage);
    if (w->ranges_host) (void)hipHostFree(w->ranges_host);
    if (w->ranges_done) (void)hipEventDestroy(w->ranges_done);
    for (auto& slot : w->insert_ring) {
        if (slot.host) (void)hipHostFree(slot.host);
        if (slot.done) (void)hipEventDestroy(slot.done);
    }
    for (auto& rb : w->rb_ring) {
        if (rb.ids_host) (void)hipHostFree(rb.ids_host);
        if (rb.st_host) (void)hipHostFree(rb.st_host);
        if (rb.err_host) (void)hipHostFree(rb.err_host);
        if (rb.ids_dev) (void)hipFree(rb.ids_dev);
        if (rb.st_dev) (void)hipFree(rb.st_dev);
        if (rb.done) (void)hipEventDestroy(rb.done);
    }
    if (w->d.hull_verts) (void)hipFree(w->d.hull_verts);
    for (void* p : {(void*)w->d.mesh_info, (void*)w->d.mesh_verts, (void*)w->d.mesh_idx, (void*)w->d.mesh_nodes,
                    (void*)w->d.mtest_cur, (void*)w->d.mtest_old, (void*)w->d.mesh_work, (void*)w->d.terrain_info,
                    (void*)w->d.terrain_heights})
        if (p) (void)hipFree(p);
    for (auto& g : w->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto& e : w->ev)
        if (e) (void)hipEventDestroy(e);
    if (w->side) (void)hipStreamSynchronize(w->side);
    if (w->fork) (void)hipEventDestroy(w->fork);
    if (w->join) (void)hipEventDestroy(w->join);
    if (w->side) (void)hipStreamDestroy(w->side);
    if (w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
    return BEPU_OK;
}

static bool settings_valid(const bepu_settings& s) {
    return s.min_iteration_count <= 6 && s.iteration_limit >= 0 && s.dt > 0 && std::isfinite(s.dt) &&
           std::isfinite(s.gravity[0]) && std::isfinite(s.gravity[1]) && std::isfinite(s.gravity[2]) &&
           !std::isnan(s.core_shape_scaling);
}

// MotionSettings.CoreShapeScaling's setter clamps to [0, 0.99] (MotionSettings.cs:22-34).
static void clamp_settings(bepu_settings& s) {
    s.core_shape_scaling = s.core_shape_scaling < 0.0f ? 0.0f : (s.core_shape_scaling > .99f ? .99f : s.core_shape_scaling);
}

bepu_status bepu_world_create(const bepu_settings* settings, uint32_t capacity_bodies, bepu_world* out) {
    return bepu_world_create_ex(settings, capacity_bodies, 0, out);
}

bepu_status bepu_world_create_ex(const bepu_settings* settings, uint32_t capacity_bodies, uint32_t capacity_pairs,
                                 bepu_world* out) {
    if (!out || capacity_bodies == 0 || capacity_bodies > (1u << 28) || capacity_pairs > (1u << 29))
        return BEPU_E_INVALID_ARG;
    *out = nullptr;
    if (!bepu_device_available()) return BEPU_E_DEVICE;
    bepu_world w = new bepu_world_s();
    if (settings) w->d.s = *settings;
    else bepu_settings_default(&w->d.s);
    if (!settings_valid(w->d.s)) {
        delete w;
        return BEPU_E_INVALID_ARG;
    }
    clamp_settings(w->d.s);
    if (hipGetDevice(&w->device) != hipSuccess) {
        delete w;
        return BEPU_E_DEVICE;
    }
    DeviceWorld& d = w->d;
    uint32_t n = capacity_bodies;
    d.body_cap = n;
    d.pair_cap = capacity_pairs ? std::max<uint32_t>(16, capacity_pairs) : std::max<uint32_t>(1024, 4 * n);
    d.small_island_max = 64;
    d.hash_cap = next_pow2(2ull * d.pair_cap);
    bepu_status st = BEPU_OK;
    auto A = [&](auto** p, size_t cnt) {
        if (st != BEPU_OK) return;
        st = dalloc(p, cnt);
        if (st == BEPU_OK) {
            w->allocations.push_back((void*)*p);
            w->state_arrays.push_back({(size_t)((char*)p - (char*)&d), cnt * sizeof(**p)});
        }
    };
    A(&d.pos, n); A(&d.quat, n); A(&d.linv, n); A(&d.angv, n);
    A(&d.ii0, n); A(&d.ii1, n); A(&d.ii2, n); A(&d.li0, n); A(&d.li1, n); A(&d.li2, n);
    A(&d.shape, n); A(&d.bmin, n); A(&d.bmax, n); A(&d.stab, n); A(&d.damp, n); A(&d.mat, n); A(&d.bflags, n);
    // World ids are reused, so they stay below the number of worlds alive at
    // once: at most n worlds with bodies, world 0 and the (possibly empty)
    // world the add calls append to, n + 2 (new_world_id enforces it).
    A(&d.wid, n); A(&d.wbase, n); A(&d.wlevels, (size_t)n + 2); A(&d.born, n);
    A(&d.wlev_count, 16 * ((size_t)n + 2)); A(&d.wlev_list, 16 * kLevelList * ((size_t)n + 2));
    A(&d.wfault, (size_t)n + 2);
    A(&d.morton, n); A(&d.morton_sorted, n); A(&d.leaf_ids, n); A(&d.leaf_ids_tmp, n);
    A(&d.node_min, 2 * (size_t)n + kMaxTreeLevels); A(&d.node_max, 2 * (size_t)n + kMaxTreeLevels);
    A(&d.bounds, 8); A(&d.bounds_part, 6 * ((size_t)n / 256 + 1));
    d.gt_cap = next_pow2(std::max<uint64_t>(1024, 2 * (uint64_t)n));
    A(&d.cell_key, n); A(&d.gnext, n); A(&d.gt_key, d.gt_cap); A(&d.gt_head, d.gt_cap);
    uint32_t P = d.pair_cap;
    A(&d.pair_keys, P); A(&d.counters, (size_t)CTR_COUNT * kCtrStride);
    A(&d.pairs_cur, P); A(&d.pairs_old, P); A(&d.hot_cur, P); A(&d.hot_old, P);
    // General convex pair state: 2 x 160 B + 3 x 8 B per pair slot.
    A(&d.conv_cur, P); A(&d.conv_old, P); A(&d.conv_work, 3 * (size_t)P); A(&d.np_work, P);
    A(&d.hash_keys_cur, d.hash_cap); A(&d.hash_vals_cur, d.hash_cap);
    A(&d.hash_keys_old, d.hash_cap); A(&d.hash_vals_old, d.hash_cap);
    A(&d.cons, P);
    // Solver constraints: the pair slots, then the joints (ids pair_cap + j).
    d.joint_cap = n;
    const size_t C = (size_t)P + d.joint_cap;
    A(&d.joints, d.joint_cap);
    A(&d.uf_parent, n); A(&d.isl_root, n); A(&d.cons_root, C); A(&d.isl_count, n + 1); A(&d.isl_start, n + 1);
    A(&d.isl_cons, C); A(&d.isl_bodies, n); A(&d.body_local, n); A(&d.isl_medium, n);
    d.cls_cap = (uint32_t)std::max<size_t>(1, std::min<size_t>(n, C));
    A(&d.cls_list, 64 * (size_t)d.cls_cap);
    A(&d.gcons, C); A(&d.color, C); A(&d.prio, C);
    A(&d.inc_count, n + 1); A(&d.inc_offset, n + 1); A(&d.inc_list, 2 * C);
    A(&d.color_count, kMaxColors); A(&d.scan_part, 4096); A(&d.color_start, kMaxColors + 1); A(&d.color_cursor, kMaxColors);
    A(&d.ordered, C);
    A(&d.grow_c, C); A(&d.grow_v, C); A(&d.grow_pair, C);
    // deactivation: per body (labels are body ids), per sub-world
    A(&d.comp_root, n); A(&d.lab_act, n); A(&d.lab_count, n); A(&d.lab_min, n); A(&d.lab_flags, n);
    A(&d.act2, n); A(&d.uf2, n); A(&d.deact_m, (size_t)n + 1); A(&d.deact_e, (size_t)n + 1);
    A(&d.wrange, 3 * ((size_t)n + 2)); A(&d.wcursor, (size_t)n + 2); A(&d.wcurgen, (size_t)n + 2);
    A(&d.wkey, (size_t)n + 2); A(&d.wsleepy, (size_t)n + 2);
    // continuous collision detection
    A(&d.minrad, n); A(&d.toi, n); A(&d.qold, n);
    if (st != BEPU_OK) {
        bepu_world_destroy(w);
        return st;
    }
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&w->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&w->join, hipEventDisableTiming) != hipSuccess) {
        bepu_world_destroy(w);
        return BEPU_E_DEVICE;
    }
    w->sort_bytes = broadphase_sort_temp_bytes(n);
    w->scan_bytes = std::max(island_scan_temp_bytes(n + 1), deactivation_scan_temp_bytes(n + 1));
    if (hipMalloc(&w->sort_temp, std::max<size_t>(w->sort_bytes, 16)) != hipSuccess ||
        hipMalloc(&w->scan_temp, std::max<size_t>(w->scan_bytes, 16)) != hipSuccess) {
        bepu_world_destroy(w);
        return BEPU_E_OUT_OF_MEMORY;
    }
    w->allocations.push_back(w->sort_temp);
    w->allocations.push_back(w->scan_temp);
    launch_fill_u64(d.hash_keys_old, ~0ull, d.hash_cap, w->stream);
    launch_fill_u64(d.hash_keys_cur, ~0ull, d.hash_cap, w->stream);
    launch_fill_u32(d.wkey, 0u, (size_t)n + 2, w->stream);
    launch_fill_u32(d.wcurgen, 0u, (size_t)n + 2, w->stream);
    launch_fill_u32(d.wcursor, kNoCursor, (size_t)n + 2, w->stream);
    launch_fill_u32(d.counters, 0u, (size_t)CTR_COUNT * kCtrStride, w->stream);
    launch_fill_u32(d.toi, 0x3f800000u, n, w->stream);  // 1.0f: no time of impact yet
    if (hipHostMalloc((void**)&w->ranges_host, 3 * ((size_t)n + 2) * sizeof(uint32_t), hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreateWithFlags(&w->ranges_done, hipEventDisableTiming) != hipSuccess) {
        bepu_world_destroy(w);
        return BEPU_E_OUT_OF_MEMORY;
    }
    if (hipStreamSynchronize(w->stream) != hipSuccess) {
        bepu_world_destroy(w);
        return BEPU_E_DEVICE;
    }
    *out = w;
    return BEPU_OK;
}

}  // extern "C"

static bool finite_n(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

// Entity construction: Entity.Initialize(collidable, mass) -> BecomeDynamic /
// BecomeKinematic (EntityBase.cs:632-645, 867-905), Orientation setter
// normalisation (EntityBase.cs:68-84).
static bepu_status make_body(bepu_world w, const bepu_motion_state* s, float4 shape, const m3& vd, float mass,
                             const bepu_material* m, uint32_t flags, HostBody& b, float hull_minimum_radius = 0) {
    if (!w || !s) return BEPU_E_INVALID_ARG;
    if (std::isnan(mass) || std::isinf(mass)) return BEPU_E_INVALID_ARG;
    if (!finite_n(s->position, 3) || !finite_n(s->orientation, 4) || !finite_n(s->linear_velocity, 3) ||
        !finite_n(s->angular_velocity, 3))
        return BEPU_E_INVALID_ARG;
    b = HostBody{};
    bool dyn = mass > 0;
    m3 li{0, 0, 0, 0, 0, 0, 0, 0, 0};
    float im = 0;
    if (dyn) {
        li = adaptive_invert(scale3(vd, mass * 2.5f));  // InertiaHelper.InertiaTensorScale = 2.5
        im = 1 / mass;
    }
    q4 q = qnormalize(q4{s->orientation[0], s->orientation[1], s->orientation[2], s->orientation[3]});
    m3 R = mat_from_quat(q);
    m3 I = rotate_inertia(R, li);
    b.pos = make_float4(s->position[0], s->position[1], s->position[2], 0);
    b.quat = make_float4(q.x, q.y, q.z, q.w);
    b.linv = make_float4(s->linear_velocity[0], s->linear_velocity[1], s->linear_velocity[2], 0);
    b.angv = make_float4(s->angular_velocity[0], s->angular_velocity[1], s->angular_velocity[2], 0);
    b.ii0 = make_float4(I.m11, I.m12, I.m13, im);
    b.ii1 = make_float4(I.m21, I.m22, I.m23, 0);
    b.ii2 = make_float4(I.m31, I.m32, I.m33, 0);
    b.li0 = make_float4(li.m11, li.m12, li.m13, 0);
    b.li1 = make_float4(li.m21, li.m22, li.m23, 0);
    b.li2 = make_float4(li.m31, li.m32, li.m33, 0);
    b.shape = shape;
    if (shape.w == (float)BEPU_SHAPE_CAPSULE || shape.w == (float)BEPU_SHAPE_CONVEX_HULL) w->convex_bodies = true;
    if (shape.w == (float)BEPU_SHAPE_BOX) w->box_bodies = true;
    if (shape.w == (float)BEPU_SHAPE_SPHERE) w->sphere_bodies = true;
    // A kinematic body's time below the limit is its activity flag, -1 when
    // added (DeactivationManager.Add activates the member).
    b.stab = make_float4(0, dyn ? 0.0f : -1.0f, 0, (flags & BEPU_FLAG_NO_STABILIZATION) ? 0.0f : 1.0f);
    b.mat = m ? make_float4(m->static_friction, m->kinetic_friction, m->bounciness, 0) : make_float4(1.0f, 0.8f, 0.0f, 0);
    b.damp = make_float2(w->d.s.linear_damping, w->d.s.angular_damping);
    const bool hull = shape.w == (float)BEPU_SHAPE_CONVEX_HULL;
    const bool general = hull || shape.w == (float)BEPU_SHAPE_CAPSULE;
    b.flags = (dyn ? BF_DYNAMIC : 0u) | BF_PRESENT | (general ? BF_GENERAL : 0u) | (hull ? BF_HULL : 0u) |
              (shape.w == (float)BEPU_SHAPE_SPHERE ? BF_SPHERE : 0u) |
              ((flags & BEPU_FLAG_ALWAYS_ACTIVE) ? BF_ALWAYS_ACTIVE : 0u) | (dyn ? 0u : (uint32_t)BF_KIN_ACTIVE);
    // A dynamic body that may sleep turns the deactivation stages on.
    if (dyn && !(flags & BEPU_FLAG_ALWAYS_ACTIVE)) w->d.sleeping = 1;
    // Entity.PositionUpdateMode (EntityBase.cs:1046); Continuous wins when both are given.
    if (flags & BEPU_FLAG_CCD_CONTINUOUS) b.flags |= BF_CONTINUOUS;
    else if (flags & BEPU_FLAG_CCD_PASSIVE) b.flags |= BF_PASSIVE;
    if (b.flags & BF_CCD_MODE) w->d.ccd = 1;
    // ConvexShape.MinimumRadius: BoxShape.cs:134 (half the smallest width:
    // the smallest half extent, exactly), SphereShape.cs:62 and
    // CapsuleShape.cs:89 (the radius), ConvexHullShape.cs:170 (the hull's
    // own plus the margin).
    if (shape.w == (float)BEPU_SHAPE_BOX) b.minrad = std::min(shape.x, std::min(shape.y, shape.z));
    else if (shape.w == (float)BEPU_SHAPE_CONVEX_HULL) b.minrad = hull_minimum_radius + shape.x;
    else b.minrad = shape.x;
    b.born = w->steps;
    return BEPU_OK;
}

static bepu_status stage_body(bepu_world w, HostBody& b, uint32_t* id);

static bepu_status add_body(bepu_world w, const bepu_motion_state* s, float4 shape, const m3& vd, float mass,
                            const bepu_material* m, uint32_t flags, uint32_t* id, float hull_minimum_radius = 0) {
    if (!w) return BEPU_E_INVALID_ARG;
    uint32_t n = w->uploaded + (uint32_t)w->staged.size();
    if (n >= w->d.body_cap) return BEPU_E_CAPACITY;
    HostBody b;
    bepu_status st = make_body(w, s, shape, vd, mass, m, flags, b, hull_minimum_radius);
    if (st != BEPU_OK) return st;
    return stage_body(w, b, id);
}

// Appends a made body to the world the add calls append to.
static bepu_status stage_body(bepu_world w, HostBody& b, uint32_t* id) {
    uint32_t n = w->uploaded + (uint32_t)w->staged.size();
    b.wid = w->cur_world;
    b.wbase = w->world_first[w->cur_world];
    w->staged.push_back(b);
    w->world_count[w->cur_world]++;
    w->ranges_dirty = true;
    if (w->absent.size() < (size_t)n + 1) {
        w->absent.resize((size_t)n + 1, 0);
        w->body_dynamic.resize((size_t)n + 1, 0);
        w->body_mesh.resize((size_t)n + 1, 0);
        w->body_world.resize((size_t)n + 1, 0);
    }
    w->absent[n] = 0;
    w->body_dynamic[n] = (b.flags & BF_DYNAMIC) ? 1 : 0;
    w->body_mesh[n] = (b.flags & BF_MESH) ? 1 : 0;
    w->body_world[n] = b.wid;
    if (id) *id = n;
    return BEPU_OK;
}

// Copies body records to the device in one block and scatters them to their
// target ids, in stream order after the work already queued (which may read
// those ids).  The records pass through a pinned host buffer of the insert
// ring and an asynchronous copy; only growing the device staging waits for
// the queued work.
static bepu_status upload_records(bepu_world w, std::vector<BodyRecord>& recs) {
    uint32_t k = (uint32_t)recs.size();
    if (k == 0) return BEPU_OK;
    const size_t bytes = (size_t)k * sizeof(BodyRecord);
    if (k > w->stage_cap) {
        if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
        uint32_t cap = std::max<uint32_t>(k, 2 * w->stage_cap);
        BodyRecord* p = nullptr;
        if (hipMalloc(&p, (size_t)cap * sizeof(BodyRecord)) != hipSuccess) return BEPU_E_OUT_OF_MEMORY;
        if (w->stage) (void)hipFree(w->stage);
        w->stage = p;
        w->stage_cap = cap;
    }
    auto& slot = w->insert_ring[w->insert_next];
    w->insert_next = (w->insert_next + 1) % bepu_world_s::kInsertRing;
    if (!slot.done && hipEventCreateWithFlags(&slot.done, hipEventDisableTiming) != hipSuccess) return BEPU_E_DEVICE;
    if (slot.pending && hipEventSynchronize(slot.done) != hipSuccess) return BEPU_E_DEVICE;  // its last copy finished
    slot.pending = false;
    if (slot.cap < bytes) {
        size_t cap = std::max(bytes, 2 * slot.cap);
        if (slot.host) (void)hipHostFree(slot.host);
        slot.host = nullptr;
        slot.cap = 0;
        if (hipHostMalloc(&slot.host, cap, hipHostMallocDefault) != hipSuccess) return BEPU_E_OUT_OF_MEMORY;
        slot.cap = cap;
    }
    std::memcpy(slot.host, recs.data(), bytes);
    if (hipMemcpyAsync(w->stage, slot.host, bytes, hipMemcpyHostToDevice, w->stream) != hipSuccess ||
        hipEventRecord(slot.done, w->stream) != hipSuccess)
        return BEPU_E_DEVICE;
    slot.pending = true;
    launch_scatter_bodies(w->d, w->stage, k, w->stream);
    return hipGetLastError() == hipSuccess ? BEPU_OK : BEPU_E_DEVICE;
}

static bepu_status upload_range(bepu_world w, uint32_t o, std::vector<HostBody>& src) {
    for (size_t i = 0; i < src.size(); ++i) src[i].target = o + (uint32_t)i;
    return upload_records(w, src);
}

static bepu_status upload(bepu_world w) {
    if (w->staged.empty()) return BEPU_OK;
    if (upload_range(w, w->uploaded, w->staged) != BEPU_OK) return BEPU_E_DEVICE;
    DeviceWorld& d = w->d;
    w->uploaded += (uint32_t)w->staged.size();
    d.nbodies = w->uploaded;
    d.nworlds = (uint32_t)w->world_first.size();
    tree_layout(d.nbodies, &d.tree_levels, d.tree_off, d.tree_cnt);
    w->staged.clear();
    return BEPU_OK;
}

static uint32_t new_world_id(bepu_world w) {
    uint32_t id;
    if (!w->free_world_ids.empty()) {
        id = w->free_world_ids.back();
        w->free_world_ids.pop_back();
    } else {
        // The per-world device arrays hold body_cap + 2 worlds; the grid key holds 20 bits.
        if (w->world_first.size() >= std::min<uint64_t>(1u << 20, (uint64_t)w->d.body_cap + 2)) return ~0u;
        w->world_first.push_back(0);
        w->world_count.push_back(0);
        w->world_live.push_back(0);
        w->world_gen.push_back(0);
        id = (uint32_t)w->world_first.size() - 1;
    }
    // A new sub-world in this id: its deactivation turn starts afresh.
    w->world_gen[id] = w->next_gen++;
    w->ranges_dirty = true;
    return id;
}

// Copies (first, count, generation) of every sub-world to the device when a
// range changed since the last copy (through a pinned buffer; the previous
// copy from it has finished by then, or is waited for).
static bepu_status push_world_ranges(bepu_world w) {
    if (!w->ranges_dirty) return BEPU_OK;
    if (w->ranges_pending && hipEventSynchronize(w->ranges_done) != hipSuccess) return BEPU_E_DEVICE;
    w->ranges_pending = false;
    const size_t nw = w->world_first.size();
    for (size_t k = 0; k < nw; ++k) {
        w->ranges_host[3 * k] = w->world_first[k];
        w->ranges_host[3 * k + 1] = w->world_live[k] ? w->world_count[k] : 0u;
        w->ranges_host[3 * k + 2] = w->world_gen[k];
    }
    if (hipMemcpyAsync(w->d.wrange, w->ranges_host, 3 * nw * sizeof(uint32_t), hipMemcpyHostToDevice, w->stream) !=
            hipSuccess ||
        hipEventRecord(w->ranges_done, w->stream) != hipSuccess)
        return BEPU_E_DEVICE;
    w->ranges_pending = true;
    w->ranges_dirty = false;
    return BEPU_OK;
}

extern "C" {

bepu_status bepu_add_box(bepu_world w, const bepu_motion_state* s, float width, float height, float length,
                         float mass, const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!(width > 0) || !(height > 0) || !(length > 0)) return BEPU_E_INVALID_ARG;
    // BoxShape(width, height, length): half extents (BoxShape.cs:56-62)
    return add_body(w, s, make_float4(width * 0.5f, height * 0.5f, length * 0.5f, (float)BEPU_SHAPE_BOX),
                    box_volume_distribution(width, height, length), mass, m, flags, id);
}

bepu_status bepu_add_sphere(bepu_world w, const bepu_motion_state* s, float radius, float mass,
                            const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!(radius > 0)) return BEPU_E_INVALID_ARG;
    return add_body(w, s, make_float4(radius, 0, 0, (float)BEPU_SHAPE_SPHERE), sphere_volume_distribution(radius), mass,
                    m, flags, id);
}

bepu_status bepu_add_capsule(bepu_world w, const bepu_motion_state* s, float length, float radius, float mass,
                             const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!(length >= 0) || !std::isfinite(length) || !(radius > 0)) return BEPU_E_INVALID_ARG;
    // CapsuleShape(length, radius): half length, the radius as the collision margin (CapsuleShape.cs:40-44)
    return add_body(w, s, capsule_shape(length, radius), capsule_volume_distribution(length, radius), mass, m, flags,
                    id);
}

bepu_status bepu_compute_hull(const float* points, uint32_t npoints, float* vertices, uint32_t capacity,
                              uint32_t* nvertices, float* volume_distribution, float* center, float* volume) {
    if (!points || npoints == 0) return BEPU_E_INVALID_ARG;
    HullDescription h;
    if (!compute_hull_description(reinterpret_cast<const f3*>(points), npoints, &h)) return BEPU_E_INVALID_ARG;
    uint32_t n = (uint32_t)h.vertices.size();
    if (nvertices) *nvertices = n;
    if (vertices)
        for (uint32_t i = 0; i < n && i < capacity; ++i) {
            vertices[3 * i] = h.vertices[i].x;
            vertices[3 * i + 1] = h.vertices[i].y;
            vertices[3 * i + 2] = h.vertices[i].z;
        }
    if (volume_distribution) {
        const m3& v = h.volume_distribution;
        float t[9] = {v.m11, v.m12, v.m13, v.m21, v.m22, v.m23, v.m31, v.m32, v.m33};
        std::memcpy(volume_distribution, t, sizeof(t));
    }
    if (center) {
        center[0] = h.center.x;
        center[1] = h.center.y;
        center[2] = h.center.z;
    }
    if (volume) *volume = h.volume;
    return BEPU_OK;
}

bepu_status bepu_compute_hull_minimum_radius(const float* points, uint32_t npoints, float* radius) {
    if (!points || npoints == 0 || !radius) return BEPU_E_INVALID_ARG;
    HullDescription h;
    if (!compute_hull_description(reinterpret_cast<const f3*>(points), npoints, &h)) return BEPU_E_INVALID_ARG;
    *radius = h.minimum_radius;
    return BEPU_OK;
}

bepu_status bepu_create_hull_shape(bepu_world w, const float* points, uint32_t npoints, uint32_t* shape_id,
                                   float* center) {
    if (!w || !points || npoints == 0) return BEPU_E_INVALID_ARG;
    HullDescription h;
    if (!compute_hull_description(reinterpret_cast<const f3*>(points), npoints, &h)) return BEPU_E_INVALID_ARG;
    bepu_world_s::HullShape hs;
    hs.first = (uint32_t)w->hull_host.size();
    hs.count = (uint32_t)h.vertices.size();
    hs.volume_distribution = h.volume_distribution;
    hs.minimum_radius = h.minimum_radius;
    for (const f3& v : h.vertices) w->hull_host.push_back(make_float4(v.x, v.y, v.z, 0));
    // The device copy grows by doubling; queued steps may read it, so wait for them.
    uint32_t need = (uint32_t)w->hull_host.size();
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    if (need > w->hull_dev_cap) {
        uint32_t cap = std::max<uint32_t>(need, std::max<uint32_t>(256, 2 * w->hull_dev_cap));
        float4* p = nullptr;
        if (hipMalloc(&p, (size_t)cap * sizeof(float4)) != hipSuccess) {
            w->hull_host.resize(hs.first);
            return BEPU_E_OUT_OF_MEMORY;
        }
        if (w->d.hull_verts) (void)hipFree(w->d.hull_verts);
        w->d.hull_verts = p;
        w->hull_dev_cap = cap;
        if (hipMemcpy(p, w->hull_host.data(), (size_t)need * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess)
            return BEPU_E_DEVICE;
    } else if (hipMemcpy(w->d.hull_verts + hs.first, w->hull_host.data() + hs.first, (size_t)hs.count * sizeof(float4),
                         hipMemcpyHostToDevice) != hipSuccess) {
        return BEPU_E_DEVICE;
    }
    w->hull_shapes.push_back(hs);
    w->d.hull_nverts = need;
    w->d.hull_max_verts = std::max(w->d.hull_max_verts, hs.count);
    if (shape_id) *shape_id = (uint32_t)w->hull_shapes.size() - 1;
    if (center) {
        center[0] = h.center.x;
        center[1] = h.center.y;
        center[2] = h.center.z;
    }
    return BEPU_OK;
}

bepu_status bepu_add_convex_hull(bepu_world w, const bepu_motion_state* s, uint32_t shape_id, float mass,
                                 const bepu_material* m, uint32_t flags, uint32_t* id) {
    if (!w || shape_id >= w->hull_shapes.size()) return BEPU_E_INVALID_ARG;
    const auto& hs = w->hull_shapes[shape_id];
    return add_body(w, s, hull_shape(w->d.s.default_margin, hs.first, hs.count), hs.volume_distribution, mass, m,
                    flags, id, hs.minimum_radius);
}

// Replaces the device copy of a host array with one holding all of `host`.
extern "C++" template <typename T>
static bepu_status mirror(const std::vector<T>& host, T** dev) {
    T* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(host.size(), 1) * sizeof(T)) != hipSuccess) return BEPU_E_OUT_OF_MEMORY;
    if (!host.empty() && hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return BEPU_E_DEVICE;
    }
    if (*dev) (void)hipFree(*dev);
    *dev = p;
    return BEPU_OK;
}

static float bits_float(uint32_t v) {
    float f;
    std::memcpy(&f, &v, sizeof(f));
    return f;
}

// The first mesh or terrain shape brings the mesh pairs' tables (per pair
// slot, double-buffered) and their work list.
static bepu_status mesh_tables(bepu_world w) {
    DeviceWorld& d = w->d;
    if (d.meshes) return BEPU_OK;
    const size_t P = std::max<size_t>(d.pair_cap, 1);
    if (hipMalloc(&d.mtest_cur, P * kMeshTesters * sizeof(MeshTester)) != hipSuccess ||
        hipMalloc(&d.mtest_old, P * kMeshTesters * sizeof(MeshTester)) != hipSuccess ||
        hipMalloc(&d.mesh_work, P * sizeof(uint2)) != hipSuccess)
        return BEPU_E_OUT_OF_MEMORY;
    d.meshes = 1;
    w->state_arrays.push_back({(size_t)((char*)&d.mtest_cur - (char*)&d), P * kMeshTesters * sizeof(MeshTester)});
    w->state_arrays.push_back({(size_t)((char*)&d.mtest_old - (char*)&d), P * kMeshTesters * sizeof(MeshTester)});
    return BEPU_OK;
}

bepu_status bepu_create_mesh_shape(bepu_world w, const float* vertices, uint32_t nvertices, const int32_t* indices,
                                   uint32_t nindices, const float* transform, uint32_t* shape_id) {
    if (!w || !vertices || !indices || nvertices == 0 || nindices == 0 || nindices % 3 != 0 || nvertices > (1u << 28) ||
        nindices > (1u << 28))
        return BEPU_E_INVALID_ARG;
    if (!finite_n(vertices, 3 * (int)nvertices) || (transform && !finite_n(transform, 12))) return BEPU_E_INVALID_ARG;
    for (uint32_t i = 0; i < nindices; ++i)
        if (indices[i] < 0 || (uint32_t)indices[i] >= nvertices) return BEPU_E_INVALID_ARG;
    m3 lin{1, 0, 0, 0, 1, 0, 0, 0, 1};
    f3 tr{0, 0, 0};
    if (transform) {
        lin = m3{transform[0], transform[1], transform[2], transform[3], transform[4], transform[5], transform[6],
                 transform[7], transform[8]};
        tr = f3{transform[9], transform[10], transform[11]};
    }
    std::vector<f3> verts(nvertices);
    for (uint32_t i = 0; i < nvertices; ++i) verts[i] = f3{vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
    std::vector<MeshTreeNode> nodes;
    uint32_t depth = 0;
    build_mesh_tree(verts.data(), indices, nindices, lin, tr, &nodes, &depth);
    // The device walk keeps one pending node per level.
    if (depth > (uint32_t)kMeshStack) return BEPU_E_CAPACITY;
    DeviceWorld& d = w->d;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    bepu_status st = mesh_tables(w);
    if (st != BEPU_OK) return st;
    const uint32_t vbase = (uint32_t)w->mesh_verts_host.size(), ibase = (uint32_t)w->mesh_idx_host.size(),
                   nbase = (uint32_t)(w->mesh_nodes_host.size() / 2);
    for (uint32_t i = 0; i < nvertices; ++i)
        w->mesh_verts_host.push_back(make_float4(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2], 0));
    w->mesh_idx_host.insert(w->mesh_idx_host.end(), indices, indices + nindices);
    for (const MeshTreeNode& n : nodes) {
        w->mesh_nodes_host.push_back(f4_of(n.bmin, bits_float((uint32_t)n.a)));
        w->mesh_nodes_host.push_back(f4_of(n.bmax, bits_float((uint32_t)n.b)));
    }
    w->mesh_info_host.push_back(make_float4(lin.m11, lin.m12, lin.m13, bits_float(vbase)));
    w->mesh_info_host.push_back(make_float4(lin.m21, lin.m22, lin.m23, bits_float(ibase)));
    w->mesh_info_host.push_back(make_float4(lin.m31, lin.m32, lin.m33, bits_float(nbase)));
    w->mesh_info_host.push_back(f4_of(tr, bits_float(nindices / 3)));
    st = mirror(w->mesh_verts_host, &d.mesh_verts);
    if (st == BEPU_OK) st = mirror(w->mesh_idx_host, &d.mesh_idx);
    if (st == BEPU_OK) st = mirror(w->mesh_nodes_host, &d.mesh_nodes);
    if (st == BEPU_OK) st = mirror(w->mesh_info_host, &d.mesh_info);
    if (st != BEPU_OK) return st;
    if (shape_id) *shape_id = w->mesh_shapes;
    w->mesh_shapes++;
    return BEPU_OK;
}

// The body record of a StaticMesh: a static collidable in the kinematic
// collision group (StaticCollidable.cs:21), never active (its flag above 0),
// at the identity (its triangles carry the shape's world transform).
static bepu_status mesh_body(bepu_world w, uint32_t shape_id, int32_t sidedness, int32_t improve,
                             const bepu_material* material, HostBody& rec) {
    if (shape_id >= w->mesh_shapes || sidedness < BEPU_SIDED_DOUBLE || sidedness > BEPU_SIDED_COUNTERCLOCKWISE)
        return BEPU_E_INVALID_ARG;
    bepu_motion_state st{};
    st.orientation[3] = 1;
    const float4 shape =
        make_float4(bits_float(shape_id), (float)sidedness, improve ? 1.0f : 0.0f, (float)BEPU_SHAPE_STATIC_MESH);
    bepu_status s = make_body(w, &st, shape, m3{0, 0, 0, 0, 0, 0, 0, 0, 0}, 0, material, 0, rec);
    if (s != BEPU_OK) return s;
    rec.stab.y = 1.0f;
    rec.flags = (rec.flags & ~(uint32_t)BF_KIN_ACTIVE) | BF_MESH;
    rec.minrad = 0;
    w->mesh_bodies = true;
    return BEPU_OK;
}

bepu_status bepu_add_static_mesh(bepu_world w, uint32_t shape_id, int32_t sidedness, int32_t improve_boundary,
                                 const bepu_material* material, uint32_t* id) {
    if (!w) return BEPU_E_INVALID_ARG;
    uint32_t n = w->uploaded + (uint32_t)w->staged.size();
    if (n >= w->d.body_cap) return BEPU_E_CAPACITY;
    HostBody b;
    bepu_status st = mesh_body(w, shape_id, sidedness, improve_boundary, material, b);
    if (st != BEPU_OK) return st;
    return stage_body(w, b, id);
}

// TerrainShape.GetBoundingBox (TerrainShape.cs:86-140) under the terrain's
// transform, with the reference's else-if (a sample that lowers a running
// minimum is not tried as a maximum), then Terrain.UpdateBoundingBox's
// thickness offset (Terrain.cs:195-212) for a thickness of zero.
static void terrain_box(const float* h, uint32_t width, uint32_t length, const m3& lin, f3 tr, f3* mn, f3* mx) {
    float minX = FLT_MAX, maxX = -FLT_MAX, minY = FLT_MAX, maxY = -FLT_MAX, minZ = FLT_MAX, maxZ = -FLT_MAX;
    f3 minXv{0, 0, 0}, maxXv{0, 0, 0}, minYv{0, 0, 0}, maxYv{0, 0, 0}, minZv{0, 0, 0}, maxZv{0, 0, 0};
    for (uint32_t i = 0; i < width; i++)
        for (uint32_t j = 0; j < length; j++) {
            const f3 v = mat_transform(f3{(float)i, h[(size_t)i * length + j], (float)j}, lin);
            if (v.x < minX) { minX = v.x; minXv = v; } else if (v.x > maxX) { maxX = v.x; maxXv = v; }
            if (v.y < minY) { minY = v.y; minYv = v; } else if (v.y > maxY) { maxY = v.y; maxYv = v; }
            if (v.z < minZ) { minZ = v.z; minZv = v; } else if (v.z > maxZ) { maxZ = v.z; maxZv = v; }
        }
    *mn = f3{minXv.x + tr.x, minYv.y + tr.y, minZv.z + tr.z};
    *mx = f3{maxXv.x + tr.x, maxYv.y + tr.y, maxZv.z + tr.z};
    // Vector3.Normalize(Down) * thickness (Vector3.Normalize takes its
    // reciprocal square root in double).
    const f3 down{-lin.m21, -lin.m22, -lin.m23};
    const float inv = (float)(1.0 / std::sqrt((double)(down.x * down.x + down.y * down.y + down.z * down.z)));
    const f3 off{down.x * inv * 0.0f, down.y * inv * 0.0f, down.z * inv * 0.0f};
    if (off.x < 0) mn->x += off.x; else mx->x += off.x;
    if (off.y < 0) mn->y += off.y; else mx->y += off.y;
    if (off.z < 0) mn->z += off.z; else mx->z += off.z;
}

bepu_status bepu_create_terrain_shape(bepu_world w, const float* heights, uint32_t width, uint32_t length,
                                      int32_t organization, const float* transform, uint32_t* shape_id) {
    if (!w || !heights || !transform || width < 2 || length < 2 || organization < 0 || organization > 1)
        return BEPU_E_INVALID_ARG;
    if ((uint64_t)width * length > (1u << 28) || (uint64_t)width * length * 2 > (uint64_t)INT32_MAX)
        return BEPU_E_INVALID_ARG;
    if (!finite_n(heights, (int)(width * length)) || !finite_n(transform, 12)) return BEPU_E_INVALID_ARG;
    const m3 lin{transform[0], transform[1], transform[2], transform[3], transform[4],
                 transform[5], transform[6], transform[7], transform[8]};
    const f3 tr{transform[9], transform[10], transform[11]};
    DeviceWorld& d = w->d;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    bepu_status st = mesh_tables(w);
    if (st != BEPU_OK) return st;
    f3 mn, mx;
    terrain_box(heights, width, length, lin, tr, &mn, &mx);
    const uint32_t hbase = (uint32_t)w->terrain_heights_host.size();
    w->terrain_heights_host.insert(w->terrain_heights_host.end(), heights, heights + (size_t)width * length);
    w->terrain_info_host.push_back(make_float4(lin.m11, lin.m12, lin.m13, bits_float(hbase)));
    w->terrain_info_host.push_back(make_float4(lin.m21, lin.m22, lin.m23, bits_float(width)));
    w->terrain_info_host.push_back(make_float4(lin.m31, lin.m32, lin.m33, bits_float(length)));
    w->terrain_info_host.push_back(f4_of(tr, bits_float((uint32_t)organization)));
    w->terrain_info_host.push_back(f4_of(mn, 0.0f));
    w->terrain_info_host.push_back(f4_of(mx, 0.0f));
    st = mirror(w->terrain_heights_host, &d.terrain_heights);
    if (st == BEPU_OK) st = mirror(w->terrain_info_host, &d.terrain_info);
    if (st != BEPU_OK) return st;
    if (shape_id) *shape_id = w->terrain_shapes;
    w->terrain_shapes++;
    return BEPU_OK;
}

// The body record of a Terrain: a static collidable like a static mesh, its
// sidedness calibrated from its transform's winding (Terrain.WorldTransform,
// Terrain.cs:41-75: Counterclockwise unless the adjugate transpose turns the
// up row against itself).
static bepu_status terrain_body(bepu_world w, uint32_t shape_id, int32_t improve, const bepu_material* material,
                                HostBody& rec) {
    if (shape_id >= w->terrain_shapes) return BEPU_E_INVALID_ARG;
    const float4* r = w->terrain_info_host.data() + 6 * (size_t)shape_id;
    const m3 m{r[0].x, r[0].y, r[0].z, r[1].x, r[1].y, r[1].z, r[2].x, r[2].y, r[2].z};
    // Matrix3x3.AdjugateTranspose (Matrix3x3.cs:673-705): its up row.
    const f3 up_n{m.m13 * m.m32 - m.m33 * m.m12, m.m11 * m.m33 - m.m13 * m.m31, m.m12 * m.m31 - m.m11 * m.m32};
    const f3 up{m.m21, m.m22, m.m23};
    const float dd = up_n.x * up.x + up_n.y * up.y + up_n.z * up.z;
    const int32_t sided = dd < 0 ? BEPU_SIDED_CLOCKWISE : BEPU_SIDED_COUNTERCLOCKWISE;
    bepu_motion_state st{};
    st.orientation[3] = 1;
    const float4 shape =
        make_float4(bits_float(shape_id), (float)sided, improve ? 1.0f : 0.0f, (float)BEPU_SHAPE_TERRAIN);
    bepu_status s = make_body(w, &st, shape, m3{0, 0, 0, 0, 0, 0, 0, 0, 0}, 0, material, 0, rec);
    if (s != BEPU_OK) return s;
    rec.stab.y = 1.0f;
    rec.flags = (rec.flags & ~(uint32_t)BF_KIN_ACTIVE) | BF_MESH;
    rec.minrad = 0;
    w->mesh_bodies = true;
    return BEPU_OK;
}

bepu_status bepu_add_terrain(bepu_world w, uint32_t shape_id, int32_t improve_boundary, const bepu_material* material,
                             uint32_t* id) {
    if (!w) return BEPU_E_INVALID_ARG;
    uint32_t n = w->uploaded + (uint32_t)w->staged.size();
    if (n >= w->d.body_cap) return BEPU_E_CAPACITY;
    HostBody b;
    bepu_status st = terrain_body(w, shape_id, improve_boundary, material, b);
    if (st != BEPU_OK) return st;
    return stage_body(w, b, id);
}

bepu_status bepu_begin_world(bepu_world w, uint32_t* world_id) {
    if (!w) return BEPU_E_INVALID_ARG;
    if (w->world_count[w->cur_world] > 0) {
        uint32_t id = new_world_id(w);
        if (id == ~0u) return BEPU_E_CAPACITY;
        w->world_first[id] = w->uploaded + (uint32_t)w->staged.size();
        w->world_count[id] = 0;
        w->world_live[id] = 1;
        w->cur_world = id;
        launch_fill_u32(w->d.wfault + id, 0u, 1, w->stream);
    }
    if (world_id) *world_id = w->cur_world;
    return BEPU_OK;
}

bepu_status bepu_get_world_range(bepu_world w, uint32_t world_id, uint32_t* first, uint32_t* count) {
    if (!w || world_id >= w->world_first.size() || !w->world_live[world_id]) return BEPU_E_INVALID_ARG;
    if (first) *first = w->world_first[world_id];
    if (count) *count = w->world_count[world_id];
    return BEPU_OK;
}

static bepu_status body_from_desc(bepu_world w, const bepu_body_desc& b, HostBody& rec) {
    if (b.shape == BEPU_SHAPE_BOX) {
        if (!(b.dims[0] > 0) || !(b.dims[1] > 0) || !(b.dims[2] > 0)) return BEPU_E_INVALID_ARG;
        return make_body(w, &b.state,
                         make_float4(b.dims[0] * 0.5f, b.dims[1] * 0.5f, b.dims[2] * 0.5f, (float)BEPU_SHAPE_BOX),
                         box_volume_distribution(b.dims[0], b.dims[1], b.dims[2]), b.mass, &b.material, b.flags, rec);
    }
    if (b.shape == BEPU_SHAPE_SPHERE) {
        if (!(b.dims[0] > 0)) return BEPU_E_INVALID_ARG;
        return make_body(w, &b.state, make_float4(b.dims[0], 0, 0, (float)BEPU_SHAPE_SPHERE),
                         sphere_volume_distribution(b.dims[0]), b.mass, &b.material, b.flags, rec);
    }
    if (b.shape == BEPU_SHAPE_CAPSULE) {
        if (!(b.dims[0] >= 0) || !std::isfinite(b.dims[0]) || !(b.dims[1] > 0)) return BEPU_E_INVALID_ARG;
        return make_body(w, &b.state, capsule_shape(b.dims[0], b.dims[1]),
                         capsule_volume_distribution(b.dims[0], b.dims[1]), b.mass, &b.material, b.flags, rec);
    }
    if (b.shape == BEPU_SHAPE_CONVEX_HULL) {
        float k = b.dims[0];
        if (!(k >= 0) || k != std::floor(k) || k >= (float)w->hull_shapes.size()) return BEPU_E_INVALID_ARG;
        const auto& hs = w->hull_shapes[(uint32_t)k];
        return make_body(w, &b.state, hull_shape(w->d.s.default_margin, hs.first, hs.count), hs.volume_distribution,
                         b.mass, &b.material, b.flags, rec, hs.minimum_radius);
    }
    if (b.shape == BEPU_SHAPE_STATIC_MESH) {
        float k = b.dims[0], sided = b.dims[1];
        if (!(k >= 0) || k != std::floor(k) || k >= (float)w->mesh_shapes || sided != std::floor(sided))
            return BEPU_E_INVALID_ARG;
        return mesh_body(w, (uint32_t)k, (int32_t)sided, b.dims[2] != 0.0f ? 1 : 0, &b.material, rec);
    }
    if (b.shape == BEPU_SHAPE_TERRAIN) {
        float k = b.dims[0];
        if (!(k >= 0) || k != std::floor(k) || k >= (float)w->terrain_shapes) return BEPU_E_INVALID_ARG;
        return terrain_body(w, (uint32_t)k, b.dims[1] != 0.0f ? 1 : 0, &b.material, rec);
    }
    return BEPU_E_UNSUPPORTED;
}

bepu_status bepu_insert_worlds(bepu_world w, const bepu_body_desc* bodies, const uint32_t* counts, uint32_t nworlds,
                               uint32_t* world_ids, uint32_t* firsts, uint32_t* inserted) {
    if (inserted) *inserted = 0;
    if (!w || (!bodies && nworlds) || (!counts && nworlds)) return BEPU_E_INVALID_ARG;
    if (upload(w) != BEPU_OK) return BEPU_E_DEVICE;
    std::vector<HostBody> recs;
    uint32_t done = 0;
    size_t off = 0;
    bepu_status result = BEPU_OK;
    for (; done < nworlds; ++done) {
        uint32_t n = counts[done];
        if (n == 0) { result = BEPU_E_INVALID_ARG; break; }
        size_t base = recs.size();
        recs.resize(base + n);
        bepu_status st = BEPU_OK;
        for (uint32_t i = 0; i < n && st == BEPU_OK; ++i) st = body_from_desc(w, bodies[off + i], recs[base + i]);
        if (st != BEPU_OK) { recs.resize(base); result = st; break; }
        // First fit among the ranges of removed worlds, else the end of the array.
        uint32_t first = ~0u;
        for (auto it = w->free_ranges.begin(); it != w->free_ranges.end(); ++it)
            if (it->second >= n) {
                first = it->first;
                uint32_t rest = it->second - n;
                w->free_ranges.erase(it);
                if (rest) w->free_ranges[first + n] = rest;
                break;
            }
        bool append = first == ~0u;
        if (append) {
            if ((uint64_t)w->uploaded + n > w->d.body_cap) { recs.resize(base); result = BEPU_E_CAPACITY; break; }
            first = w->uploaded;
        }
        uint32_t id = new_world_id(w);
        if (id == ~0u) {
            if (!append) w->free_ranges[first] = n;
            recs.resize(base);
            result = BEPU_E_CAPACITY;
            break;
        }
        for (uint32_t i = 0; i < n; ++i) {
            recs[base + i].wid = id;
            recs[base + i].wbase = first;
            recs[base + i].target = first + i;
        }
        w->world_first[id] = first;
        w->world_count[id] = n;
        w->world_live[id] = 1;
        launch_fill_u32(w->d.wfault + id, 0u, 1, w->stream);  // a reused id starts unfaulted
        if (w->absent.size() < (size_t)first + n) {
            w->absent.resize((size_t)first + n, 0);
            w->body_dynamic.resize((size_t)first + n, 0);
            w->body_world.resize((size_t)first + n, 0);
        }
        std::fill(w->absent.begin() + first, w->absent.begin() + first + n, 0);
        for (uint32_t i = 0; i < n; ++i) {
            w->body_dynamic[first + i] = (recs[base + i].flags & BF_DYNAMIC) ? 1 : 0;
            w->body_world[first + i] = id;
        }
        if (w->world_joints.size() <= id) w->world_joints.resize((size_t)id + 1, 0);
        w->world_joints[id] = 0;  // a reused id starts its joint ordinals again
        if (append) {
            w->uploaded += n;
            w->cur_world = id;  // later add calls open a new sub-world after this one
        }
        if (world_ids) world_ids[done] = id;
        if (firsts) firsts[done] = first;
        off += n;
    }
    if (!recs.empty()) {
        w->d.nbodies = w->uploaded;
        w->d.nworlds = (uint32_t)w->world_first.size();
        tree_layout(w->d.nbodies, &w->d.tree_levels, w->d.tree_off, w->d.tree_cnt);
        if (upload_records(w, recs) != BEPU_OK) return BEPU_E_DEVICE;
    }
    if (inserted) *inserted = done;
    return done == nworlds ? BEPU_OK : result;
}

bepu_status bepu_insert_world(bepu_world w, const bepu_body_desc* bodies, uint32_t n, uint32_t* world_id,
                              uint32_t* first_body) {
    if (n == 0) return BEPU_E_INVALID_ARG;
    return bepu_insert_worlds(w, bodies, &n, 1, world_id, first_body, nullptr);
}

bepu_status bepu_remove_world(bepu_world w, uint32_t world_id) {
    if (!w || world_id >= w->world_first.size() || !w->world_live[world_id]) return BEPU_E_INVALID_ARG;
    if (upload(w) != BEPU_OK) return BEPU_E_DEVICE;
    uint32_t first = w->world_first[world_id], n = w->world_count[world_id];
    if (n) launch_fill_u32(w->d.bflags + first, 0u, n, w->stream);  // absent: no box, no pairs, not dynamic
    if (n) launch_joint_sweep(w->d, w->stream);  // its joints end
    w->world_live[world_id] = 0;
    w->world_count[world_id] = 0;
    w->ranges_dirty = true;
    if (n) {
        // Merge with free neighbours.
        auto next = w->free_ranges.lower_bound(first);
        if (next != w->free_ranges.end() && next->first == first + n) {
            n += next->second;
            w->free_ranges.erase(next);
        }
        auto prev = w->free_ranges.lower_bound(first);
        if (prev != w->free_ranges.begin()) {
            --prev;
            if (prev->first + prev->second == first) {
                first = prev->first;
                n += prev->second;
                w->free_ranges.erase(prev);
            }
        }
        w->free_ranges[first] = n;
    }
    if (n) std::fill(w->absent.begin() + first, w->absent.begin() + first + n, 1);
    // Free the id first, so that a fresh sub-world for the appends below takes
    // it back: the number of ids never exceeds the worlds alive at once.
    w->free_world_ids.push_back(world_id);
    if (world_id == w->cur_world) {
        // Appends continue in a fresh sub-world.
        uint32_t id = new_world_id(w);
        w->world_first[id] = w->uploaded + (uint32_t)w->staged.size();
        w->world_count[id] = 0;
        w->world_live[id] = 1;
        w->cur_world = id;
    }
    return BEPU_OK;
}

bepu_status bepu_synchronize(bepu_world w) {
    if (!w) return BEPU_E_INVALID_ARG;
    return poll_errors(w);
}

bepu_status bepu_clear_error(bepu_world w) {
    if (!w) return BEPU_E_INVALID_ARG;
    poll_errors(w);
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    // An evicted sub-world stays evicted until it is removed: the bit that
    // says some sub-world may be is kept.
    const uint32_t keep = w->sticky & STICKY_NONFINITE;
    launch_fill_u32(w->d.counters + ctr(CTR_STICKY), keep, 1, w->stream);
    for (uint32_t i : {(uint32_t)CTR_OVERFLOW, (uint32_t)CTR_GBAR_ABORT})
        launch_fill_u32(w->d.counters + ctr(i), 0u, 1, w->stream);
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    w->sticky = keep;
    return BEPU_OK;
}

bepu_status bepu_get_motion_states_ranges(bepu_world w, const uint32_t* firsts, const uint32_t* counts,
                                          uint32_t nranges, bepu_motion_state* out) {
    if (!w || (nranges && (!firsts || !counts || !out))) return BEPU_E_INVALID_ARG;
    if (upload(w) != BEPU_OK) return BEPU_E_DEVICE;
    std::vector<uint32_t> ids;
    for (uint32_t r = 0; r < nranges; ++r) {
        if ((uint64_t)firsts[r] + counts[r] > w->d.nbodies) return BEPU_E_INVALID_ARG;
        for (uint32_t i = 0; i < counts[r]; ++i) ids.push_back(firsts[r] + i);
    }
    uint32_t k = (uint32_t)ids.size();
    if (check(hipStreamSynchronize(w->stream)) != BEPU_OK) return BEPU_E_DEVICE;
    if (k == 0) return BEPU_OK;
    if (k > w->rb_cap) {
        uint32_t cap = std::max<uint32_t>(k, 2 * w->rb_cap);
        uint32_t* pi = nullptr;
        float4* ps = nullptr;
        if (hipMalloc(&pi, (size_t)cap * sizeof(uint32_t)) != hipSuccess) return BEPU_E_OUT_OF_MEMORY;
        if (hipMalloc(&ps, (size_t)cap * 4 * sizeof(float4)) != hipSuccess) {
            (void)hipFree(pi);
            return BEPU_E_OUT_OF_MEMORY;
        }
        if (w->id_stage) (void)hipFree(w->id_stage);
        if (w->state_stage) (void)hipFree(w->state_stage);
        w->id_stage = pi;
        w->state_stage = ps;
        w->rb_cap = cap;
    }
    if (hipMemcpy(w->id_stage, ids.data(), (size_t)k * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
        return BEPU_E_DEVICE;
    launch_gather_states(w->d, w->id_stage, k, w->state_stage, w->stream);
    bepu_status err = poll_errors(w);
    if (err == BEPU_E_DEVICE && !(w->sticky & STICKY_GBAR_ABORT)) return err;
    std::vector<float4> st((size_t)k * 4);
    if (hipMemcpy(st.data(), w->state_stage, st.size() * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
        return BEPU_E_DEVICE;
    for (uint32_t i = 0; i < k; ++i) {
        const float4 &p = st[4 * i], &q = st[4 * i + 1], &v = st[4 * i + 2], &a = st[4 * i + 3];
        bepu_motion_state& m = out[i];
        m.position[0] = p.x; m.position[1] = p.y; m.position[2] = p.z;
        m.orientation[0] = q.x; m.orientation[1] = q.y; m.orientation[2] = q.z; m.orientation[3] = q.w;
        m.linear_velocity[0] = v.x; m.linear_velocity[1] = v.y; m.linear_velocity[2] = v.z;
        m.angular_velocity[0] = a.x; m.angular_velocity[1] = a.y; m.angular_velocity[2] = a.z;
    }
    if (err == BEPU_OK && ranges_faulted(w, firsts, counts, nranges)) return BEPU_E_NONFINITE;
    return err;
}

bepu_status bepu_get_motion_states_range(bepu_world w, uint32_t first, uint32_t count, bepu_motion_state* out) {
    return bepu_get_motion_states_ranges(w, &first, &count, 1, out);
}

bepu_status bepu_remove_worlds(bepu_world w, const uint32_t* world_ids, uint32_t n) {
    if (!w || (n && !world_ids)) return BEPU_E_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i) {
        bepu_status st = bepu_remove_world(w, world_ids[i]);
        if (st != BEPU_OK) return st;
    }
    return BEPU_OK;
}

bepu_status bepu_set_timing(bepu_world w, int32_t enable) {
    if (!w) return BEPU_E_INVALID_ARG;
    w->timing = enable != 0;
    return BEPU_OK;
}

bepu_status bepu_get_tuning(bepu_world w, bepu_tuning* out) {
    if (!w || !out) return BEPU_E_INVALID_ARG;
    out->narrowphase_block = w->np_block;
    out->convex_lds = w->convex_lds;
    out->convex_side_stream = w->convex_side ? 1u : 0u;
    out->small_island_max = w->d.small_island_max;
    out->step_graph = w->step_graph ? 1u : 0u;
    out->ms_per_step = 0;
    return BEPU_OK;
}

bepu_status bepu_set_tuning(bepu_world w, const bepu_tuning* t) {
    if (!w || !t) return BEPU_E_INVALID_ARG;
    if ((t->narrowphase_block != 64 && t->narrowphase_block != 128 && t->narrowphase_block != 256) ||
        t->convex_lds > 2 || t->convex_side_stream > 1 || (t->small_island_max != 32 && t->small_island_max != 64) ||
        t->step_graph > 1)
        return BEPU_E_INVALID_ARG;
    w->step_graph = t->step_graph != 0;
    w->np_block = t->narrowphase_block;
    w->convex_lds = t->convex_lds;
    w->convex_side = t->convex_side_stream != 0;
    w->d.small_island_max = t->small_island_max;
    return BEPU_OK;
}

// Coordinate descent over the launch configuration on the world's own steps:
// every choice gives bit-identical results, so the trial steps are ordinary
// steps.  Each candidate value of a choice runs `trial_steps` steps in each of
// two rounds (the rounds interleave the candidates, so a workload that changes
// over the steps biases none of them), timed by events on the world's stream.
// Steps at least this long (ms, the tuner's best so far) do not try graph replay.
constexpr float kGraphTrialMaxMs = 0.25f;

bepu_status bepu_autotune(bepu_world w, uint32_t trial_steps, bepu_tuning* out, uint32_t* steps_run) {
    if (!w || trial_steps == 0 || trial_steps > 1000) return BEPU_E_INVALID_ARG;
    bepu_status st = upload(w);
    if (st != BEPU_OK) return st;
    struct Choice {
        uint32_t bepu_tuning::*field;
        std::vector<uint32_t> values;
    };
    // The general convex choices first: they move the most time.
    std::vector<Choice> choices;
    if (w->convex_bodies) {
        if (w->d.hull_max_verts <= kConvexStageVerts) choices.push_back({&bepu_tuning::convex_lds, {2, 0, 1}});
        choices.push_back({&bepu_tuning::convex_side_stream, {0, 1}});
    }
    choices.push_back({&bepu_tuning::narrowphase_block, {64, 128, 256}});
    choices.push_back({&bepu_tuning::small_island_max, {32, 64}});
    // Replaying a captured step saves the host its launches; whether the
    // device runs the graph's launches closer together than the stream's
    // depends on the world (on MI355X a 70-body world runs them 5% apart).
    if (w->d.bp_kind == BEPU_BROADPHASE_GRID) choices.push_back({&bepu_tuning::step_graph, {0, 1}});
    bepu_tuning cur;
    (void)bepu_get_tuning(w, &cur);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return BEPU_E_DEVICE;
    }
    const bool timing = w->timing;
    w->timing = false;
    uint32_t ran = 0;
    float best_ms = 0;
    st = BEPU_OK;
    for (const Choice& c : choices) {
        // Graph replay only saves the host its launches: a step whose device
        // time is long does not wait for them, and there the trials' noise
        // exceeds what replay could gain (measured 2% slower on the batched
        // 4,096-world step of 0.9 ms), so it is tried on short steps only.
        if (c.field == &bepu_tuning::step_graph && best_ms > kGraphTrialMaxMs) continue;
        std::vector<double> total(c.values.size(), 0.0);
        for (int round = 0; round < 2 && st == BEPU_OK; ++round)
            for (size_t i = 0; i < c.values.size() && st == BEPU_OK; ++i) {
                bepu_tuning t = cur;
                t.*(c.field) = c.values[i];
                (void)bepu_set_tuning(w, &t);
                float ms = 0;
                // One untimed step first, then the timed ones queued behind it,
                // so that the device never waits for the host between them.
                st = bepu_step(w, 1);
                if (st == BEPU_OK && hipEventRecord(e0, w->stream) != hipSuccess) st = BEPU_E_DEVICE;
                if (st == BEPU_OK) st = bepu_step(w, (int32_t)trial_steps);
                if (st == BEPU_OK && (hipEventRecord(e1, w->stream) != hipSuccess ||
                                      hipEventSynchronize(e1) != hipSuccess ||
                                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    st = BEPU_E_DEVICE;
                total[i] += ms;
                ran += trial_steps + 1;
            }
        if (st != BEPU_OK) break;
        // The current value is kept unless another is faster by more than 2%:
        // the trial steps see a workload that changes from step to step.
        size_t best = 0;
        for (size_t i = 0; i < total.size(); ++i)
            if (c.values[i] == cur.*(c.field)) best = i;
        for (size_t i = 0; i < total.size(); ++i)
            if (total[i] < 0.98 * total[best]) best = i;
        cur.*(c.field) = c.values[best];
        best_ms = (float)(total[best] / (2.0 * trial_steps));
    }
    (void)bepu_set_tuning(w, &cur);
    w->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cur.ms_per_step = best_ms;
    if (out) *out = cur;
    if (steps_run) *steps_run = ran;
    return st;
}

bepu_status bepu_set_broadphase(bepu_world w, int32_t kind) {
    if (!w || (kind != BEPU_BROADPHASE_GRID && kind != BEPU_BROADPHASE_TREE)) return BEPU_E_INVALID_ARG;
    w->d.bp_kind = (uint32_t)kind;
    return BEPU_OK;
}

// The launches of one step on the world's stream (and its side stream), with
// stage events when `e` is given.
static bepu_status enqueue_step(bepu_world w, hipEvent_t* e) {
    DeviceWorld& d = w->d;
    hipStream_t s = w->stream;
    launch_step_reset(d, s);
    if (e) (void)hipEventRecord(e[0], s);
    // DeactivationManager.Update, then the force update and bounding boxes.
    launch_deactivation(d, w->scan_temp, w->scan_bytes, s);
    launch_body_pre(d, s);
    if (e) (void)hipEventRecord(e[1], s);
    launch_broadphase(d, w->sort_temp, w->sort_bytes, s);
    if (e) (void)hipEventRecord(e[2], s);
    // Steps with stage events run the general convex tests after the
    // closed-form narrowphase, so that each stage's time is its own.
    const bool side = w->convex_bodies && w->convex_side && !e;
    if (d.pair_lists) launch_pair_prep(d, s);
    if (side) {
        // Test the closed-form pairs on the side stream while the general
        // convex pairs (the longer of the two) run here: a launch behind
        // an event starts a few microseconds later than one in order.
        (void)hipEventRecord(w->fork, s);
        (void)hipStreamWaitEvent(w->side, w->fork, 0);
        launch_narrowphase(d, w->np_block, w->side);
        (void)hipEventRecord(w->join, w->side);
        launch_convex(d, w->convex_lds, s);
        (void)hipStreamWaitEvent(s, w->join, 0);
    } else {
        launch_narrowphase(d, w->np_block, s);
    }
    if (e) (void)hipEventRecord(e[3], s);
    if (!side && w->convex_bodies) launch_convex(d, w->convex_lds, s);
    if (e) (void)hipEventRecord(e[4], s);
    if (w->mesh_bodies) launch_mesh(d, s);
    if (e) (void)hipEventRecord(e[5], s);
    // Joints join their bodies' islands; then the islands woken by this
    // step's new constraints, stale frozen pairs, split activation.
    launch_joint_edges(d, s);
    launch_activity(d, s);
    launch_islands(d, w->scan_temp, w->scan_bytes, s);
    if (e) (void)hipEventRecord(e[6], s);
    launch_island_solve(d, s);
    if (e) (void)hipEventRecord(e[7], s);
    launch_global_solver(d, s);
    if (e) (void)hipEventRecord(e[8], s);
    launch_integrate(d, s);
    if (d.ccd) launch_ccd(d, s);
    if (e) (void)hipEventRecord(e[9], s);
    return BEPU_OK;
}

// The host state a captured step depends on besides DeviceWorld (which the
// launches take by value).
static uint64_t graph_config(bepu_world w) {
    return (uint64_t)w->np_block | ((uint64_t)w->convex_lds << 16) | ((uint64_t)w->convex_side << 24) |
           ((uint64_t)w->convex_bodies << 25);
}

// One step through a HIP graph of its launches.  There is one graph per
// parity of the step count (the double-buffered pair arrays swap every step,
// and the launches take their pointers by value); a graph is captured on
// first use and again whenever the world's device description or launch
// configuration changed (bodies added or removed, settings, tuning), the
// executable graph being updated in place where the launches kept their
// shape.  The index of the step is read from device memory, so one graph
// replays every later step of its parity.  The tree broadphase (its radix
// sort), and a failed capture, take ordinary launches.
static bepu_status replay_step(bepu_world w) {
    if (!w->step_graph || w->d.bp_kind != BEPU_BROADPHASE_GRID) return enqueue_step(w, nullptr);
    auto& g = w->graphs[w->steps & 1u];
    const uint64_t cfg = graph_config(w);
    if (!g.exec || g.config != cfg || std::memcmp(&g.key, &w->d, sizeof(DeviceWorld)) != 0) {
        hipGraph_t graph = nullptr;
        bool ok = hipStreamBeginCapture(w->stream, hipStreamCaptureModeRelaxed) == hipSuccess;
        if (ok) {
            const bepu_status est = enqueue_step(w, nullptr);
            ok = hipStreamEndCapture(w->stream, &graph) == hipSuccess && graph && est == BEPU_OK;
        }
        if (ok && g.exec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult why;
            if (hipGraphExecUpdate(g.exec, graph, &bad, &why) != hipSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(g.exec);
                g.exec = nullptr;
            }
        }
        if (ok && !g.exec) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!ok) {
            // Steps of this world are launched one by one from now on.
            (void)hipGetLastError();
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            w->step_graph = false;
            return enqueue_step(w, nullptr);
        }
        std::memcpy((void*)&g.key, &w->d, sizeof(DeviceWorld));
        g.config = cfg;
    }
    return hipGraphLaunch(g.exec, w->stream) == hipSuccess ? BEPU_OK : BEPU_E_DEVICE;
}

bepu_status bepu_step(bepu_world w, int32_t n_steps) {
    if (!w || n_steps < 0) return BEPU_E_INVALID_ARG;
    // A world whose steps overflowed a capacity or aborted a grid barrier
    // takes no more steps until the caller acknowledges it (bepu_clear_error).
    if (w->sticky & (STICKY_OVERFLOW | STICKY_GBAR_ABORT)) return sticky_status(w->sticky);
    bepu_status st = upload(w);
    if (st != BEPU_OK) return st;
    st = push_world_ranges(w);
    if (st != BEPU_OK) return st;
    DeviceWorld& d = w->d;
    hipStream_t s = w->stream;
    // Stage timers: every step of this call when timing is on, none otherwise.
    uint32_t timed = w->timing ? (uint32_t)n_steps : 0u;
    size_t need = (size_t)timed * (ST_COUNT + 1);
    if (need > 0 && hipStreamSynchronize(s) != hipSuccess) return BEPU_E_DEVICE;
    while (w->ev.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return BEPU_E_DEVICE;
        w->ev.push_back(e);
    }
    if (timed) w->timed_steps = timed;
    // Pairs of more than one tester are listed by tester first; a world of
    // boxes only (or spheres only) skips that pass.
    d.pair_lists = (w->convex_bodies || w->mesh_bodies || (w->box_bodies && w->sphere_bodies)) ? 1u : 0u;
    for (int k = 0; k < n_steps; ++k) {
        int slot = (int)timed - (n_steps - k);  // >= 0 for the timed steps
        hipEvent_t* e = slot >= 0 ? &w->ev[(size_t)slot * (ST_COUNT + 1)] : nullptr;
        st = e ? enqueue_step(w, e) : replay_step(w);
        if (st != BEPU_OK) return st;
        // This step's pair states and table become next step's history.
        std::swap(d.pairs_cur, d.pairs_old);
        std::swap(d.hot_cur, d.hot_old);
        std::swap(d.conv_cur, d.conv_old);
        std::swap(d.mtest_cur, d.mtest_old);
        std::swap(d.hash_keys_cur, d.hash_keys_old);
        std::swap(d.hash_vals_cur, d.hash_vals_old);
        w->steps++;
    }
    if (hipGetLastError() != hipSuccess) return BEPU_E_DEVICE;
    return BEPU_OK;
}

// Average device time per stage over the timed steps of the last bepu_step call.
bepu_status bepu_get_stage_times(bepu_world w, float* ms, uint32_t cap, uint32_t* count, const char** names) {
    if (!w) return BEPU_E_INVALID_ARG;
    if (check(hipStreamSynchronize(w->stream)) != BEPU_OK) return BEPU_E_DEVICE;
    if (count) *count = ST_COUNT;
    if (names) *names = kStageNames;
    for (uint32_t i = 0; i < ST_COUNT && i < cap; ++i) {
        double sum = 0;
        for (uint32_t k = 0; k < w->timed_steps; ++k) {
            float t = 0;
            const hipEvent_t* e = &w->ev[(size_t)k * (ST_COUNT + 1)];
            if (hipEventElapsedTime(&t, e[i], e[i + 1]) != hipSuccess) t = 0;
            sum += t;
        }
        ms[i] = w->timed_steps ? (float)(sum / w->timed_steps) : 0.0f;
    }
    return BEPU_OK;
}

static bepu_status read_counters(bepu_world w, uint32_t* c) {
    if (upload(w) != BEPU_OK) return BEPU_E_DEVICE;
    if (check(hipStreamSynchronize(w->stream)) != BEPU_OK) return BEPU_E_DEVICE;
    std::vector<uint32_t> all((size_t)CTR_COUNT * kCtrStride);
    if (hipMemcpy(all.data(), w->d.counters, all.size() * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return BEPU_E_DEVICE;
    for (uint32_t i = 0; i < CTR_COUNT; ++i) c[i] = all[ctr(i)];
    w->sticky |= c[CTR_STICKY] | (c[CTR_OVERFLOW] ? (uint32_t)STICKY_OVERFLOW : 0u) |
                 (c[CTR_GBAR_ABORT] ? (uint32_t)STICKY_GBAR_ABORT : 0u);
    return BEPU_OK;
}

bepu_status bepu_get_motion_states(bepu_world w, bepu_motion_state* out, uint32_t cap, uint32_t* count) {
    if (!w) return BEPU_E_INVALID_ARG;
    if (upload(w) != BEPU_OK) return BEPU_E_DEVICE;
    bepu_status err = poll_errors(w);
    if (err == BEPU_E_DEVICE && !(w->sticky & STICKY_GBAR_ABORT)) return err;
    uint32_t n = w->d.nbodies;
    if (count) *count = n;
    if (!out || cap == 0) return n == 0 ? BEPU_OK : BEPU_E_CAPACITY;
    uint32_t k = std::min(n, cap);
    std::vector<float4> p(k), q(k), v(k), a(k);
    bool ok = hipMemcpy(p.data(), w->d.pos, k * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess &&
              hipMemcpy(q.data(), w->d.quat, k * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess &&
              hipMemcpy(v.data(), w->d.linv, k * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess &&
              hipMemcpy(a.data(), w->d.angv, k * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) return BEPU_E_DEVICE;
    for (uint32_t i = 0; i < k; ++i) {
        bepu_motion_state& m = out[i];
        m.position[0] = p[i].x; m.position[1] = p[i].y; m.position[2] = p[i].z;
        m.orientation[0] = q[i].x; m.orientation[1] = q[i].y; m.orientation[2] = q[i].z; m.orientation[3] = q[i].w;
        m.linear_velocity[0] = v[i].x; m.linear_velocity[1] = v[i].y; m.linear_velocity[2] = v[i].z;
        m.angular_velocity[0] = a[i].x; m.angular_velocity[1] = a[i].y; m.angular_velocity[2] = a[i].z;
    }
    if (n > cap) return BEPU_E_CAPACITY;
    if (err == BEPU_OK) {
        std::vector<uint32_t> bad;
        if (faulted_worlds(w, &bad) != BEPU_OK) return BEPU_E_DEVICE;
        if (!bad.empty()) return BEPU_E_NONFINITE;
    }
    return err;
}

bepu_status bepu_get_pairs(bepu_world w, uint64_t* keys, uint32_t cap, uint32_t* count) {
    if (!w) return BEPU_E_INVALID_ARG;
    uint32_t c[CTR_COUNT];
    if (read_counters(w, c) != BEPU_OK) return BEPU_E_DEVICE;
    if (w->sticky) return sticky_status(w->sticky);
    uint32_t n = std::min(c[CTR_PAIRS], w->d.pair_cap);
    if (w->d.sleeping && w->steps) {
        // Overlaps of two inactive bodies with no pair, and pairs removed as
        // stale, are not pairs.
        std::vector<PairState> ps(n);
        if (n && hipMemcpy(ps.data(), w->d.pairs_old, n * sizeof(PairState), hipMemcpyDeviceToHost) != hipSuccess)
            return BEPU_E_DEVICE;
        std::vector<uint64_t> live;
        for (const PairState& s : ps)
            if (!(s.flags & (PF_DEAD | PF_VOID))) live.push_back(s.key);
        std::sort(live.begin(), live.end());
        if (count) *count = (uint32_t)live.size();
        if (!keys || cap < live.size()) return live.empty() ? BEPU_OK : BEPU_E_CAPACITY;
        std::copy(live.begin(), live.end(), keys);
        return BEPU_OK;
    }
    if (count) *count = n;
    if (!keys || cap < n) return n == 0 ? BEPU_OK : BEPU_E_CAPACITY;
    if (n && hipMemcpy(keys, w->d.pair_keys, n * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess)
        return BEPU_E_DEVICE;
    std::sort(keys, keys + n);
    return BEPU_OK;
}

// The last step's pair states sit in pairs_old after the end-of-step swap.
static bepu_status read_pairs(bepu_world w, std::vector<PairState>& ps, uint32_t* c) {
    if (read_counters(w, c) != BEPU_OK) return BEPU_E_DEVICE;
    uint32_t n = w->steps ? std::min(c[CTR_PAIRS], w->d.pair_cap) : 0;
    ps.resize(n);
    if (n && hipMemcpy(ps.data(), w->d.pairs_old, n * sizeof(PairState), hipMemcpyDeviceToHost) != hipSuccess)
        return BEPU_E_DEVICE;
    // stale and void slots are not pairs
    ps.erase(std::remove_if(ps.begin(), ps.end(), [](const PairState& s) { return (s.flags & (PF_DEAD | PF_VOID)) != 0; }),
             ps.end());
    return BEPU_OK;
}

bepu_status bepu_get_manifolds(bepu_world w, bepu_manifold* out, uint32_t cap, uint32_t* count) {
    if (!w) return BEPU_E_INVALID_ARG;
    std::vector<PairState> ps;
    uint32_t c[CTR_COUNT];
    if (read_pairs(w, ps, c) != BEPU_OK) return BEPU_E_DEVICE;
    std::vector<const PairState*> live;
    for (const PairState& s : ps)
        if (s.ccount > 0) live.push_back(&s);
    std::sort(live.begin(), live.end(), [](const PairState* a, const PairState* b) { return a->key < b->key; });
    uint32_t n = (uint32_t)live.size();
    if (count) *count = n;
    if (!out || cap < n) return n == 0 ? sticky_status(w->sticky) : BEPU_E_CAPACITY;
    for (uint32_t i = 0; i < n; ++i) {
        const PairState& s = *live[i];
        bepu_manifold& m = out[i];
        std::memset(&m, 0, sizeof(m));
        m.key = s.key;
        m.body_a = s.a;
        m.body_b = s.b;
        m.count = s.ccount;
        m.kind = s.kind;
        for (int k = 0; k < s.ccount; ++k) {
            int slot = s.clist[k];
            bepu_contact& o = m.contacts[k];
            o.position[0] = s.cpos[slot].x; o.position[1] = s.cpos[slot].y; o.position[2] = s.cpos[slot].z;
            o.penetration_depth = s.cpos[slot].w;
            o.normal[0] = s.cnrm[slot].x; o.normal[1] = s.cnrm[slot].y; o.normal[2] = s.cnrm[slot].z;
            float idf = s.cnrm[slot].w;
            std::memcpy(&o.id, &idf, 4);
        }
    }
    return sticky_status(w->sticky);
}

bepu_status bepu_get_impulses(bepu_world w, bepu_impulses* out, uint32_t cap, uint32_t* count) {
    if (!w) return BEPU_E_INVALID_ARG;
    std::vector<PairState> ps;
    uint32_t c[CTR_COUNT];
    if (read_pairs(w, ps, c) != BEPU_OK) return BEPU_E_DEVICE;
    std::vector<bepu_impulses> v;
    for (const PairState& s : ps) {
        if (s.ccount <= 0) continue;
        v.emplace_back();
        bepu_impulses& o = v.back();
        std::memset(&o, 0, sizeof(o));
        o.key = s.key;
        o.color = s.color;
        o.count = s.npen;
        for (int k = 0; k < s.npen; ++k) o.normal[k] = s.penacc[k];
        bool convex = s.kind == BEPU_PAIR_BOX_BOX || s.kind == BEPU_PAIR_CONVEX;
        if (convex) {
            o.sliding[0] = s.slacc[0];
            o.sliding[1] = s.slacc[1];
            o.twist = s.twacc;
        } else {
            for (int k = 0; k < s.npen; ++k) o.friction[k] = s.fracc[k];
        }
    }
    std::sort(v.begin(), v.end(), [](const bepu_impulses& a, const bepu_impulses& b) {
        return a.color != b.color ? a.color < b.color : a.key < b.key;
    });
    uint32_t m = (uint32_t)v.size();
    if (count) *count = m;
    if (!out || cap < m) return m == 0 ? sticky_status(w->sticky) : BEPU_E_CAPACITY;
    std::copy(v.begin(), v.end(), out);
    return sticky_status(w->sticky);
}

bepu_status bepu_get_stats(bepu_world w, bepu_step_stats* out) {
    if (!w || !out) return BEPU_E_INVALID_ARG;
    uint32_t c[CTR_COUNT];
    if (read_counters(w, c) != BEPU_OK) return BEPU_E_DEVICE;
    bool any = w->steps > 0;
    uint32_t npairs = any ? std::min(c[CTR_PAIRS], w->d.pair_cap) : 0;
    launch_stats(w->d, npairs, w->stream);
    if (read_counters(w, c) != BEPU_OK) return BEPU_E_DEVICE;
    std::memset(out, 0, sizeof(*out));
    out->bodies = w->d.nbodies;
    out->pairs = npairs;
    out->manifolds = any ? c[CTR_CONS] : 0;
    out->contacts = any ? c[CTR_CONTACTS] : 0;
    out->colors = any ? c[CTR_COLORS] : 0;
    out->steps = w->steps;
    float mp;
    std::memcpy(&mp, &c[CTR_MAXPEN], 4);
    out->max_penetration = any ? mp : 0.0f;
    out->global_manifolds = any ? std::min(c[CTR_GCONS], w->d.pair_cap + w->d.joint_cap) : 0;
    out->convex_pairs[0] = any ? c[CTR_CONVEX] + c[CTR_CONVEX_OTHER0] : 0;
    out->convex_pairs[1] = any ? c[CTR_CONVEX_SHALLOW] + c[CTR_CONVEX_OTHER0 + 1] : 0;
    out->convex_pairs[2] = any ? c[CTR_CONVEX_DEEP] + c[CTR_CONVEX_OTHER0 + 2] : 0;
    out->mesh_pairs = any && w->mesh_bodies ? c[CTR_MESH] : 0;
    out->mesh_testers = any && w->mesh_bodies ? c[CTR_MESH_TESTERS] : 0;
    // Any step since the last bepu_clear_error that overflowed a capacity
    // (BEPU_E_CAPACITY) or whose global solver grid was not resident (BEPU_E_DEVICE).
    return sticky_status(w->sticky);
}

// Space.Remove(entity) (Space.cs:364-482).  The body leaves every stage from
// the next step on; its pairs end with it.  Its id is not reused.
bepu_status bepu_remove(bepu_world w, uint32_t id) {
    if (!w) return BEPU_E_INVALID_ARG;
    uint32_t total = w->uploaded + (uint32_t)w->staged.size();
    if (id >= total || w->absent[id]) return BEPU_E_NOT_FOUND;
    w->absent[id] = 1;
    if (id >= w->uploaded) {
        w->staged[id - w->uploaded].flags = 0u;  // uploaded absent
        return BEPU_OK;
    }
    // The dynamic bodies it had constraints with are activated
    // (DeactivationManager.RemoveSimulationIslandFromMember).
    launch_removal_wake(w->d, id, w->stream);
    launch_fill_u32(w->d.bflags + id, 0u, 1, w->stream);
    // Its joints end; their other connections are activated.
    launch_joint_sweep(w->d, w->stream);
    return hipGetLastError() == hipSuccess ? BEPU_OK : BEPU_E_DEVICE;
}

// new BallSocketJoint(connectionA, connectionB, anchor) added to the space
// (BallSocketJoint.cs:57-63): the bodies' offsets to the anchor in their
// local frames, from their poses after the steps queued so far.  Both bodies
// belong to one sub-world; at least one is dynamic ("Cannot constrain two
// kinematic bodies", BallSocketJoint.cs:281).
bepu_status bepu_add_ball_socket(bepu_world w, uint32_t a, uint32_t b, const float* anchor, uint32_t* id) {
    if (!w || !anchor || a == b || !std::isfinite(anchor[0]) || !std::isfinite(anchor[1]) ||
        !std::isfinite(anchor[2]))
        return BEPU_E_INVALID_ARG;
    const uint32_t total = w->uploaded + (uint32_t)w->staged.size();
    if (a >= total || b >= total || w->absent[a] || w->absent[b]) return BEPU_E_NOT_FOUND;
    if (!w->body_dynamic[a] && !w->body_dynamic[b]) return BEPU_E_INVALID_ARG;
    if (w->body_mesh[a] || w->body_mesh[b]) return BEPU_E_INVALID_ARG;  // joints connect entities
    if (w->body_world[a] != w->body_world[b]) return BEPU_E_INVALID_ARG;
    if (w->d.njoints >= w->d.joint_cap) return BEPU_E_CAPACITY;
    // The joint reads the bodies' poses on the device: upload staged bodies first.
    bepu_status st = upload(w);
    if (st != BEPU_OK) return st;
    const uint32_t wid = w->body_world[a];
    if (w->world_joints.size() <= wid) w->world_joints.resize((size_t)wid + 1, 0);
    const uint32_t jid = w->d.njoints;
    launch_joint_init(w->d, jid, a, b, w->world_joints[wid]++, make_float4(anchor[0], anchor[1], anchor[2], 0),
                      w->stream);
    w->d.njoints++;
    if (id) *id = jid;
    return hipGetLastError() == hipSuccess ? BEPU_OK : BEPU_E_DEVICE;
}

// SpringSettings.Stiffness / Damping (clamped at zero, SpringSettings.cs:61-80;
// both zero is refused, as ComputeErrorReductionAndSoftness throws) and
// Joint.MaxCorrectiveVelocity (Joint.cs:24-38).
bepu_status bepu_set_joint_settings(bepu_world w, uint32_t joint, float stiffness, float damping,
                                    float max_corrective_velocity) {
    if (!w) return BEPU_E_INVALID_ARG;
    if (joint >= w->d.njoints) return BEPU_E_NOT_FOUND;
    if (std::isnan(stiffness) || std::isnan(damping) || std::isnan(max_corrective_velocity)) return BEPU_E_INVALID_ARG;
    stiffness = std::max(0.0f, stiffness);
    damping = std::max(0.0f, damping);
    if (stiffness == 0 && damping == 0) return BEPU_E_INVALID_ARG;
    const float mc = std::max(0.0f, max_corrective_velocity);
    const float mc2 = mc >= FLT_MAX ? FLT_MAX : mc * mc;
    launch_joint_settings(w->d, joint, stiffness, damping, mc, mc2, w->stream);
    return hipGetLastError() == hipSuccess ? BEPU_OK : BEPU_E_DEVICE;
}

// BallSocketJoint.TotalImpulse of every joint, three floats each.
bepu_status bepu_get_joint_impulses(bepu_world w, float* out, uint32_t capacity, uint32_t* count) {
    if (!w) return BEPU_E_INVALID_ARG;
    const uint32_t n = w->d.njoints;
    if (count) *count = n;
    if (!out) return BEPU_OK;
    if (capacity < n) return BEPU_E_CAPACITY;
    bepu_status err = poll_errors(w);
    if (err == BEPU_E_DEVICE && !(w->sticky & STICKY_GBAR_ABORT)) return err;
    std::vector<JointState> js(n);
    if (n && hipMemcpyAsync(js.data(), w->d.joints, n * sizeof(JointState), hipMemcpyDeviceToHost, w->stream) != hipSuccess)
        return BEPU_E_DEVICE;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    for (uint32_t k = 0; k < n; ++k) {
        out[3 * k] = js[k].acc.x;
        out[3 * k + 1] = js[k].acc.y;
        out[3 * k + 2] = js[k].acc.z;
    }
    return err;
}

// Entity state setters for each listed body, the parts `fields` selects:
// Position, Orientation (normalised as the Orientation setter does,
// EntityBase.cs:68-84), LinearVelocity, AngularVelocity; all four is the
// MotionState setter (Entities/EntityBase.cs:218-235).  Takes effect at the
// next step; pair caches, warm-start impulses and stabilization state stay.
static bepu_status upload_state_records(bepu_world w, std::vector<StateRecord>& recs);

bepu_status bepu_set_motion_states(bepu_world w, const uint32_t* ids, const bepu_motion_state* states, uint32_t n,
                                   uint32_t fields) {
    if (!w || (n && (!ids || !states)) || fields == 0 || (fields & ~BEPU_STATE_ALL)) return BEPU_E_INVALID_ARG;
    uint32_t total = w->uploaded + (uint32_t)w->staged.size();
    uint32_t dev = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const bepu_motion_state& m = states[k];
        if (ids[k] >= total || w->absent[ids[k]]) return BEPU_E_NOT_FOUND;
        if (w->body_mesh[ids[k]]) return BEPU_E_INVALID_ARG;  // a static mesh is not an entity
        if (((fields & BEPU_STATE_POSITION) && !finite_n(m.position, 3)) ||
            ((fields & BEPU_STATE_ORIENTATION) && !finite_n(m.orientation, 4)) ||
            ((fields & BEPU_STATE_LINEAR_VELOCITY) && !finite_n(m.linear_velocity, 3)) ||
            ((fields & BEPU_STATE_ANGULAR_VELOCITY) && !finite_n(m.angular_velocity, 3)))
            return BEPU_E_INVALID_ARG;
        if (ids[k] < w->uploaded) dev++;
    }
    std::vector<StateRecord> recs;
    recs.reserve(dev);
    for (uint32_t k = 0; k < n; ++k) {
        const bepu_motion_state& m = states[k];
        q4 q = qnormalize(q4{m.orientation[0], m.orientation[1], m.orientation[2], m.orientation[3]});
        float4 pos = make_float4(m.position[0], m.position[1], m.position[2], 0);
        float4 quat = make_float4(q.x, q.y, q.z, q.w);
        float4 linv = make_float4(m.linear_velocity[0], m.linear_velocity[1], m.linear_velocity[2], 0);
        float4 angv = make_float4(m.angular_velocity[0], m.angular_velocity[1], m.angular_velocity[2], 0);
        if (ids[k] >= w->uploaded) {
            HostBody& b = w->staged[ids[k] - w->uploaded];
            if (fields & BEPU_STATE_POSITION) b.pos = pos;
            if (fields & BEPU_STATE_LINEAR_VELOCITY) b.linv = linv;
            if (fields & BEPU_STATE_ANGULAR_VELOCITY) b.angv = angv;
            if (fields & BEPU_STATE_ORIENTATION) {
                b.quat = quat;
                m3 li{b.li0.x, b.li0.y, b.li0.z, b.li1.x, b.li1.y, b.li1.z, b.li2.x, b.li2.y, b.li2.z};
                m3 I = rotate_inertia(mat_from_quat(q), li);
                b.ii0 = make_float4(I.m11, I.m12, I.m13, b.ii0.w);
                b.ii1 = make_float4(I.m21, I.m22, I.m23, 0);
                b.ii2 = make_float4(I.m31, I.m32, I.m33, 0);
            }
            continue;
        }
        StateRecord r{};
        r.pos = pos;
        r.quat = quat;
        r.linv = linv;
        r.angv = angv;
        r.target = ids[k];
        r.fields = fields;
        recs.push_back(r);
    }
    return upload_state_records(w, recs);
}

// State records to the device, applied in stream order (the records activate
// their bodies; fields 0 writes no state).
static bepu_status upload_state_records(bepu_world w, std::vector<StateRecord>& recs) {
    if (recs.empty()) return BEPU_OK;
    // Queued steps may still read the staging buffer of an earlier call.
    if (hipStreamSynchronize(w->stream) != hipSuccess) return BEPU_E_DEVICE;
    uint32_t k = (uint32_t)recs.size();
    if (k > w->sstage_cap) {
        uint32_t cap = std::max<uint32_t>(k, 2 * w->sstage_cap);
        StateRecord* p = nullptr;
        if (hipMalloc(&p, (size_t)cap * sizeof(StateRecord)) != hipSuccess) return BEPU_E_OUT_OF_MEMORY;
        if (w->sstage) (void)hipFree(w->sstage);
        w->sstage = p;
        w->sstage_cap = cap;
    }
    if (hipMemcpy(w->sstage, recs.data(), (size_t)k * sizeof(StateRecord), hipMemcpyHostToDevice) != hipSuccess)
        return BEPU_E_DEVICE;
    launch_set_states(w->d, w->sstage, k, w->stream);
    return hipGetLastError() == hipSuccess ? BEPU_OK : BEPU_E_DEVICE;
}

// Entity.ActivityInformation.Activate() (SimulationIslandMember.cs:256-277).
bepu_status bepu_activate(bepu_world w, const uint32_t* ids, uint32_t n) {
    if (!w || (n && !ids)) return BEPU_E_INVALID_ARG;
    uint32_t total = w->uploaded + (uint32_t)w->staged.size();
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= total || w->absent[ids[k]]) return BEPU_E_NOT_FOUND;
        if (w->body_mesh[ids[k]]) return BEPU_E_INVALID_ARG;
    }
    std::vector<StateRecord> recs;
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= w->uploaded) continue;  // not yet stepped: active already
        StateRecord r{};
        r.target = ids[k];
        r.fields = 0;
        recs.push_back(r);
    }
    return upload_state_records(w, recs);
}

// Entity.PositionUpdateMode setter (EntityBase.cs:1050-1068): the mode bits of
// each listed body, from the next step on; the body is not activated.
bepu_status bepu_set_position_update_mode(bepu_world w, const uint32_t* ids, uint32_t n, int32_t mode) {
    if (!w || (n && !ids) || mode < BEPU_POSITION_UPDATE_DISCRETE || mode > BEPU_POSITION_UPDATE_CONTINUOUS)
        return BEPU_E_INVALID_ARG;
    uint32_t total = w->uploaded + (uint32_t)w->staged.size();
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= total || w->absent[ids[k]]) return BEPU_E_NOT_FOUND;
        if (w->body_mesh[ids[k]]) return BEPU_E_INVALID_ARG;
    }
    const uint32_t bits = mode == BEPU_POSITION_UPDATE_CONTINUOUS
                              ? (uint32_t)BF_CONTINUOUS
                              : (mode == BEPU_POSITION_UPDATE_PASSIVE ? (uint32_t)BF_PASSIVE : 0u);
    if (bits && n) w->d.ccd = 1;
    std::vector<StateRecord> recs;
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= w->uploaded) {
            HostBody& b = w->staged[ids[k] - w->uploaded];
            b.flags = (b.flags & ~BF_CCD_MODE) | bits;
            continue;
        }
        StateRecord r{};
        r.target = ids[k];
        r.fields = kStateUpdateMode;
        r.pad[0] = bits;
        recs.push_back(r);
    }
    return upload_state_records(w, recs);
}

// Entity.ActivityInformation.IsActive of every body: a dynamic body whose
// island is awake, a kinematic body whose flag is <= 0; 0 for a removed body.
bepu_status bepu_get_activity(bepu_world w, uint8_t* out, uint32_t cap, uint32_t* count) {
    if (!w) return BEPU_E_INVALID_ARG;
    if (upload(w) != BEPU_OK) return BEPU_E_DEVICE;
    bepu_status err = poll_errors(w);
    if (err == BEPU_E_DEVICE && !(w->sticky & STICKY_GBAR_ABORT)) return err;
    const uint32_t n = w->d.nbodies;
    if (count) *count = n;
    if (!out) return BEPU_OK;  // count query
    if (cap < n) return BEPU_E_CAPACITY;
    std::vector<uint32_t> f(n);
    std::vector<float4> st(n);
    if (n && (hipMemcpy(f.data(), w->d.bflags, n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(st.data(), w->d.stab, n * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess))
        return BEPU_E_DEVICE;
    for (uint32_t i = 0; i < n; ++i) {
        const bool present = (f[i] & BF_PRESENT) != 0;
        out[i] = !present ? 0 : ((f[i] & BF_DYNAMIC) ? !(f[i] & BF_ASLEEP) : st[i].y <= 0.0f);
    }
    return err;
}

bepu_status bepu_get_settings(bepu_world w, bepu_settings* out) {
    if (!w || !out) return BEPU_E_INVALID_ARG;
    *out = w->d.s;
    return BEPU_OK;
}

// Solver.IterationLimit (Constraints/Solver.cs:19), ForceUpdater.Gravity
// (OtherSpaceStages/ForceUpdater.cs:19-29), TimeStepSettings.TimeStepDuration
// and the other settings, changed between steps: every later step uses them.
// Values an entity takes at creation (its damping, a hull's margin) keep
// theirs, as the reference's entities do.
bepu_status bepu_set_settings(bepu_world w, const bepu_settings* settings) {
    if (!w || !settings || !settings_valid(*settings)) return BEPU_E_INVALID_ARG;
    w->d.s = *settings;
    clamp_settings(w->d.s);
    return BEPU_OK;
}

// Asynchronous form of bepu_get_motion_states_ranges: queues the gather and
// the copy to pinned host memory behind the steps already queued and returns
// a ticket at once; bepu_collect_motion_states waits for that copy alone.
bepu_status bepu_request_motion_states_ranges(bepu_world w, const uint32_t* firsts, const uint32_t* counts,
                                              uint32_t nranges, uint32_t* ticket) {
    if (!w || !ticket || (nranges && (!firsts || !counts))) return BEPU_E_INVALID_ARG;
    if (upload(w) != BEPU_OK) return BEPU_E_DEVICE;
    uint64_t k64 = 0;
    for (uint32_t r = 0; r < nranges; ++r) {
        if ((uint64_t)firsts[r] + counts[r] > w->d.nbodies) return BEPU_E_INVALID_ARG;
        k64 += counts[r];
    }
    if (k64 > (1u << 30)) return BEPU_E_INVALID_ARG;
    const uint32_t k = (uint32_t)k64;
    const uint32_t t = w->rb_next_ticket;
    auto& rb = w->rb_ring[t % bepu_world_s::kReadbackRing];
    if (rb.busy) return BEPU_E_CAPACITY;  // the ring holds this many uncollected readbacks
    if (!rb.done && hipEventCreateWithFlags(&rb.done, hipEventDisableTiming) != hipSuccess) return BEPU_E_DEVICE;
    if (!rb.err_host && hipHostMalloc((void**)&rb.err_host, 3 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return BEPU_E_OUT_OF_MEMORY;
    if (k > rb.cap) {
        // The slot's previous copies completed when it was collected.
        if (rb.ids_host) (void)hipHostFree(rb.ids_host);
        if (rb.st_host) (void)hipHostFree(rb.st_host);
        if (rb.ids_dev) (void)hipFree(rb.ids_dev);
        if (rb.st_dev) (void)hipFree(rb.st_dev);
        rb.ids_host = nullptr;
        rb.st_host = nullptr;
        rb.ids_dev = nullptr;
        rb.st_dev = nullptr;
        rb.cap = 0;
        uint32_t cap = std::max<uint32_t>(k, 1024);
        if (hipHostMalloc((void**)&rb.ids_host, (size_t)cap * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&rb.st_host, (size_t)cap * 4 * sizeof(float4), hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&rb.ids_dev, (size_t)cap * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(&rb.st_dev, (size_t)cap * 4 * sizeof(float4)) != hipSuccess)
            return BEPU_E_OUT_OF_MEMORY;
        rb.cap = cap;
    }
    uint32_t o = 0;
    for (uint32_t r = 0; r < nranges; ++r)
        for (uint32_t i = 0; i < counts[r]; ++i) rb.ids_host[o++] = firsts[r] + i;
    hipStream_t s = w->stream;
    if (k) {
        if (hipMemcpyAsync(rb.ids_dev, rb.ids_host, (size_t)k * sizeof(uint32_t), hipMemcpyHostToDevice, s) !=
            hipSuccess)
            return BEPU_E_DEVICE;
        launch_gather_states(w->d, rb.ids_dev, k, rb.st_dev, s);
        if (hipMemcpyAsync(rb.st_host, rb.st_dev, (size_t)k * 4 * sizeof(float4), hipMemcpyDeviceToHost, s) !=
            hipSuccess)
            return BEPU_E_DEVICE;
    }
    const uint32_t idx[3] = {CTR_STICKY, CTR_OVERFLOW, CTR_GBAR_ABORT};
    for (int i = 0; i < 3; ++i)
        if (hipMemcpyAsync(rb.err_host + i, w->d.counters + ctr(idx[i]), sizeof(uint32_t), hipMemcpyDeviceToHost,
                           s) != hipSuccess)
            return BEPU_E_DEVICE;
    if (hipEventRecord(rb.done, s) != hipSuccess) return BEPU_E_DEVICE;
    rb.n = k;
    rb.ticket = t;
    rb.busy = true;
    w->rb_next_ticket = t + 1 ? t + 1 : 1;
    *ticket = t;
    return BEPU_OK;
}

// Waits for a requested readback and writes its motion states (in request
// order) to `out`; the status carries the world's error flags as they stood
// when the readback ran (BEPU_E_CAPACITY / BEPU_E_DEVICE, outputs filled).
bepu_status bepu_collect_motion_states(bepu_world w, uint32_t ticket, bepu_motion_state* out) {
    if (!w || ticket == 0) return BEPU_E_INVALID_ARG;
    auto& rb = w->rb_ring[ticket % bepu_world_s::kReadbackRing];
    if (!rb.busy || rb.ticket != ticket) return BEPU_E_NOT_FOUND;
    if (rb.n && !out) return BEPU_E_INVALID_ARG;
    if (hipEventSynchronize(rb.done) != hipSuccess) return BEPU_E_DEVICE;
    for (uint32_t i = 0; i < rb.n; ++i) {
        const float4 &p = rb.st_host[4 * i], &q = rb.st_host[4 * i + 1], &v = rb.st_host[4 * i + 2],
                     &a = rb.st_host[4 * i + 3];
        bepu_motion_state& m = out[i];
        m.position[0] = p.x; m.position[1] = p.y; m.position[2] = p.z;
        m.orientation[0] = q.x; m.orientation[1] = q.y; m.orientation[2] = q.z; m.orientation[3] = q.w;
        m.linear_velocity[0] = v.x; m.linear_velocity[1] = v.y; m.linear_velocity[2] = v.z;
        m.angular_velocity[0] = a.x; m.angular_velocity[1] = a.y; m.angular_velocity[2] = a.z;
    }
    rb.busy = false;
    w->sticky |= rb.err_host[0] | (rb.err_host[1] ? (uint32_t)STICKY_OVERFLOW : 0u) |
                 (rb.err_host[2] ? (uint32_t)STICKY_GBAR_ABORT : 0u);
    return sticky_status(w->sticky);
}

// Sub-worlds evicted by the non-finite guard (a body's position, orientation
// or velocity no longer finite after a step): their bodies left every stage
// from the next step on.
bepu_status bepu_get_faulted_worlds(bepu_world w, uint32_t* ids, uint32_t cap, uint32_t* count) {
    if (!w) return BEPU_E_INVALID_ARG;
    bepu_status st = poll_errors(w);
    if (st == BEPU_E_DEVICE && !(w->sticky & STICKY_GBAR_ABORT)) return st;
    std::vector<uint32_t> bad;
    if (faulted_worlds(w, &bad) != BEPU_OK) return BEPU_E_DEVICE;
    if (count) *count = (uint32_t)bad.size();
    if (!ids) return BEPU_OK;  // count query
    if (cap < bad.size()) return BEPU_E_CAPACITY;
    std::copy(bad.begin(), bad.end(), ids);
    return BEPU_OK;
}

// Whole-world snapshots.  The state is every device array of the world
// (bodies, pair keys and states, tester tables, key tables, islands,
// deactivation labels and turns, joints, counters), the settings and the
// step count; the structure (bodies, removals, sub-worlds, joints, shapes)
// must be the one the snapshot was taken with, and a fingerprint of it is
// checked on restore.
extern "C++" {
namespace {
constexpr char kSnapMagic[8] = {'B', 'E', 'P', 'U', 'S', 'N', 'A', 'P'};
constexpr uint32_t kSnapVersion = 1;
struct SnapHeader {
    char magic[8];
    uint32_t version, narrays;
    uint64_t fingerprint, steps, total;
    uint32_t sticky, device_world_bytes;
    bepu_settings settings;
};

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    return h;
}

template <class T>
uint64_t fnv_vec(uint64_t h, const std::vector<T>& v) {
    const uint64_t n = v.size();
    h = fnv(h, &n, sizeof(n));
    return v.empty() ? h : fnv(h, v.data(), v.size() * sizeof(T));
}

// The world's structure: bodies uploaded and staged, removals, sub-world
// ranges, joints, shapes and the arrays that hold the state.
uint64_t structure_fingerprint(bepu_world w) {
    uint64_t h = 1469598103934665603ull;
    const uint64_t scalars[10] = {w->uploaded, w->staged.size(), w->d.njoints, w->d.nworlds, w->hull_host.size(),
                                  w->mesh_shapes, w->terrain_shapes, w->d.body_cap, w->d.pair_cap, w->d.sleeping};
    h = fnv(h, scalars, sizeof(scalars));
    h = fnv_vec(h, w->absent);
    h = fnv_vec(h, w->world_first);
    h = fnv_vec(h, w->world_count);
    h = fnv_vec(h, w->world_live);
    h = fnv_vec(h, w->state_arrays);
    return h;
}

uint64_t snapshot_bytes(bepu_world w) {
    uint64_t total = sizeof(SnapHeader);
    for (const auto& a : w->state_arrays) total += a.second;
    return total;
}
}  // namespace
}  // extern "C++"

bepu_status bepu_snapshot_size(bepu_world w, uint64_t* bytes) {
    if (!w || !bytes) return BEPU_E_INVALID_ARG;
    *bytes = snapshot_bytes(w);
    return BEPU_OK;
}

bepu_status bepu_save_snapshot(bepu_world w, void* buffer, uint64_t capacity, uint64_t* bytes) {
    if (!w) return BEPU_E_INVALID_ARG;
    const uint64_t total = snapshot_bytes(w);
    if (bytes) *bytes = total;
    if (!buffer || capacity < total) return BEPU_E_INVALID_ARG;
    if (!w->staged.empty()) {
        const bepu_status st = upload(w);  // bodies added since the last step become device state first
        if (st != BEPU_OK) return st;
    }
    if (check(hipStreamSynchronize(w->stream)) != BEPU_OK) return BEPU_E_DEVICE;
    SnapHeader h{};
    std::memcpy(h.magic, kSnapMagic, sizeof(kSnapMagic));
    h.version = kSnapVersion;
    h.narrays = (uint32_t)w->state_arrays.size();
    h.fingerprint = structure_fingerprint(w);
    h.steps = w->steps;
    h.total = total;
    h.sticky = w->sticky;
    h.device_world_bytes = (uint32_t)sizeof(DeviceWorld);
    h.settings = w->d.s;
    char* out = static_cast<char*>(buffer);
    std::memcpy(out, &h, sizeof(h));
    size_t off = sizeof(h);
    for (const auto& a : w->state_arrays) {
        const void* dev = *reinterpret_cast<void* const*>(reinterpret_cast<const char*>(&w->d) + a.first);
        if (hipMemcpy(out + off, dev, a.second, hipMemcpyDeviceToHost) != hipSuccess) return BEPU_E_DEVICE;
        off += a.second;
    }
    return BEPU_OK;
}

bepu_status bepu_restore_snapshot(bepu_world w, const void* buffer, uint64_t size) {
    if (!w || !buffer || size < sizeof(SnapHeader)) return BEPU_E_INVALID_ARG;
    SnapHeader h;
    std::memcpy(&h, buffer, sizeof(h));
    if (std::memcmp(h.magic, kSnapMagic, sizeof(kSnapMagic)) != 0 || h.version != kSnapVersion ||
        h.device_world_bytes != sizeof(DeviceWorld) || h.total != size || h.narrays != w->state_arrays.size())
        return BEPU_E_INVALID_ARG;
    if (!w->staged.empty()) {
        const bepu_status st = upload(w);
        if (st != BEPU_OK) return st;
    }
    if (h.fingerprint != structure_fingerprint(w) || size != snapshot_bytes(w)) return BEPU_E_INVALID_ARG;
    if (check(hipStreamSynchronize(w->stream)) != BEPU_OK) return BEPU_E_DEVICE;
    const char* in = static_cast<const char*>(buffer);
    size_t off = sizeof(h);
    for (const auto& a : w->state_arrays) {
        void* dev = *reinterpret_cast<void**>(reinterpret_cast<char*>(&w->d) + a.first);
        if (hipMemcpy(dev, in + off, a.second, hipMemcpyHostToDevice) != hipSuccess) return BEPU_E_DEVICE;
        off += a.second;
    }
    w->steps = h.steps;
    w->sticky = h.sticky;
    w->d.s = h.settings;
    return BEPU_OK;
}

}  // extern "C"
