// world_host.h -- host-side state of a world handle (bepu_world), shared by
// the C ABI sources.  Internal: not part of the installed interface.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "../../include/bepu_gpu.h"
#include "bodies_host.h"
#include "world.h"

using HostBody = bepu::BodyRecord;

struct bepu_world_s {
    bepu::DeviceWorld d{};
    hipStream_t stream = nullptr;
    // The closed-form pairs are tested on a second stream while the general
    // convex pairs run on the first (both kernels are bound by memory latency
    // and leave most of the device idle alone); fork and join are events.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // Launch configuration (bepu_tuning); small_island_max lives in d.
    uint32_t np_block = 128;
    uint32_t convex_lds = 0;        // 0 global hull reads, 1 LDS staging, 2 chosen per step on the device
    bool convex_side = true;
    bool step_graph = false;        // steps replay a captured HIP graph (bepu_tuning.step_graph)
    // One captured step per parity of the step count (the pair arrays swap
    // every step and the launches hold their pointers), with the device
    // description and launch configuration it was captured with.
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        bepu::DeviceWorld key{};
        uint64_t config = 0;
    };
    StepGraph graphs[2];
    std::vector<HostBody> staged;   // bodies not yet uploaded
    uint32_t uploaded = 0;
    void* sort_temp = nullptr;
    size_t sort_bytes = 0;
    void* scan_temp = nullptr;
    size_t scan_bytes = 0;
    bool timing = false;
    std::vector<hipEvent_t> ev;     // (ST_COUNT + 1) per timed step of the last bepu_step call
    uint32_t timed_steps = 0;
    uint32_t steps = 0;
    std::vector<void*> allocations;
    // Every device array of the world's state, as (offset of its pointer in
    // DeviceWorld, bytes): what a snapshot copies (double-buffered arrays by
    // role, whichever buffer holds it now).
    std::vector<std::pair<size_t, size_t>> state_arrays;
    // Sub-worlds: world k holds bodies [world_first[k], world_first[k] + world_count[k]).
    std::vector<uint32_t> world_first{0}, world_count{0};
    std::vector<uint8_t> world_live{1};
    // Generation of each world id (a new one at every (re)use), and the copy
    // of (first, count, generation) per world the deactivation stage reads:
    // refreshed through a pinned buffer before a step when a range changed.
    std::vector<uint32_t> world_gen{1};
    uint32_t next_gen = 2;
    bool ranges_dirty = true;
    uint32_t* ranges_host = nullptr;
    hipEvent_t ranges_done = nullptr;
    bool ranges_pending = false;
    uint32_t cur_world = 0;                  // world the add calls append to
    std::vector<uint32_t> free_world_ids;
    std::map<uint32_t, uint32_t> free_ranges;  // body ranges of removed worlds: first -> count
    bepu::BodyRecord* stage = nullptr;      // device staging of body records
    uint32_t stage_cap = 0;
    uint32_t* id_stage = nullptr;           // device staging of body ids for readbacks
    float4* state_stage = nullptr;          // 4 float4 per body read back
    uint32_t rb_cap = 0;
    // Hull shapes: vertices back to back in hull_host (mirrored in d.hull_verts).
    struct HullShape {
        uint32_t first, count;
        m3 volume_distribution;
        float minimum_radius;  // without the margin
    };
    std::vector<HullShape> hull_shapes;
    bool convex_bodies = false;             // a capsule or hull was ever added: run k_convex
    bool box_bodies = false, sphere_bodies = false;  // a box / a sphere was ever added
    std::vector<float4> hull_host;
    uint32_t hull_dev_cap = 0;
    // Static mesh shapes: their vertices, indices, tree nodes and records
    // back to back on the host (mirrored in d.mesh_verts, d.mesh_idx,
    // d.mesh_nodes, d.mesh_info); the device copies are reallocated as they grow.
    uint32_t mesh_shapes = 0;
    std::vector<float4> mesh_verts_host, mesh_nodes_host, mesh_info_host;
    std::vector<int32_t> mesh_idx_host;
    // Terrain shapes: heights and records (mirrored in d.terrain_heights, d.terrain_info).
    uint32_t terrain_shapes = 0;
    std::vector<float> terrain_heights_host;
    std::vector<float4> terrain_info_host;
    bool mesh_bodies = false;               // a static mesh or terrain was ever added: run k_mesh
    int device = 0;                         // the HIP device the world lives on
    uint32_t sticky = 0;                    // STICKY_* bits seen by the host since the last bepu_clear_error
    std::vector<uint8_t> absent;            // per body id: removed (bepu_remove, or its world was removed)
    std::vector<uint8_t> body_dynamic;      // per body id: dynamic (mass > 0) when it was made
    std::vector<uint8_t> body_mesh;         // per body id: a static mesh (no state to set, never active)
    std::vector<uint32_t> body_world;       // per body id: its sub-world
    std::vector<uint32_t> world_joints;     // per world id: joints made in it (the next joint's ordinal)
    bepu::StateRecord* sstage = nullptr;    // device staging of motion states written between steps
    uint32_t sstage_cap = 0;
    // Inserted body records go through a ring of pinned host buffers whose
    // asynchronous copies to the device staging are fenced by events, so an
    // insertion never waits for the steps queued before it.
    struct PinnedSlot {
        void* host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    static constexpr int kInsertRing = 3;
    PinnedSlot insert_ring[kInsertRing];
    int insert_next = 0;
    // Asynchronous readbacks of body ranges (bepu_request_motion_states_ranges):
    // a ring of slots, each with its own device and pinned host buffers, the
    // error words copied beside the states, and an event.
    struct Readback {
        uint32_t* ids_host = nullptr;
        uint32_t* ids_dev = nullptr;
        float4* st_dev = nullptr;
        float4* st_host = nullptr;
        uint32_t* err_host = nullptr;  // sticky, overflow and abort words at the readback's place in the stream
        uint32_t cap = 0;
        uint32_t n = 0;
        uint32_t ticket = 0;
        hipEvent_t done = nullptr;
        bool busy = false;
    };
    static constexpr int kReadbackRing = 8;
    Readback rb_ring[kReadbackRing];
    uint32_t rb_next_ticket = 1;
};

namespace bepu {
// The status of the STICKY_* error bits: BEPU_E_DEVICE for a grid barrier
// abort, BEPU_E_CAPACITY for an overflow, else BEPU_OK.
bepu_status sticky_status(uint32_t bits);
// Waits for the world's stream and folds the device's error flags into
// w->sticky; returns sticky_status of the result.  Allocates nothing.
bepu_status poll_errors(bepu_world w);
}  // namespace bepu
