// k_grid.hip -- hierarchical hash-grid broadphase.
//
// Produces the same pair SET as the tree broadphase and the reference's
// DynamicHierarchy (DynamicHierarchy.cs:194-214): closed-interval AABB
// overlap (BoundingBox.cs:58-69) of two bodies, at least one dynamic
// (kinematic-kinematic is NoBroadPhase, CollisionRules.cs:175-178).
//
// Levels l = 0..14 have cubic cells of edge e_l = s0 * 4^l, s0 chosen each
// step so that the largest AABB side of 90% of the bodies fits level 0.  A body lives at the finest level with
// largest side <= 0.99 * e_l, in the cell holding its AABB min corner.  If two
// boxes overlap, their min corners differ by at most the larger of their
// sides per axis, so:
//   * two bodies of one level sit in cells at most one apart per axis;
//   * a body overlaps a body of a coarser level l' only if the latter's cell
//     is within one cell (at level l') of the cell holding the finer body's
//     min corner.
// A body therefore finds all its pairs by probing its own cell and the 13
// "forward" neighbour cells at its level (a same-level pair is reported once:
// through the forward neighbour, or by the smaller sorted index inside one
// cell) and, at every coarser non-empty level of its world, the 8 to 27
// cells from the one below its min corner to the one holding its max corner
// (a cross-level pair is reported by the finer body).  Cell indices are
// computed in double precision, so the one-cell bound holds for coordinates
// up to ~1e13 cell edges from the origin.
//
// Every cell key carries the body's sub-world, so batched worlds never pair.
//
// No sort: a hash table maps each packed cell key to the head of a linked
// list of the bodies in that cell (one atomic exchange per body).  Probes are
// independent of each other; the dependent chain per probe is the table
// slot, the list head and the (short) list.
#include "pairbuf.h"

namespace bepu {

constexpr int kGridLevels = 15;      // level fits in 4 bits and a key is never ~0
constexpr float kFit = 0.99f;
constexpr uint64_t kEmptyCell = ~0ull;
constexpr int kQueryBlock = 256;     // queries per workgroup (one pair-list atomic per workgroup)
constexpr int kQueryBuf = 2048;      // LDS pair buffer; overflow goes straight to the global list

__device__ __forceinline__ int grid_level(float side, float s0) {
    float lim = s0 * kFit;
    int l = 0;
    while (l < kGridLevels - 1 && !(side <= lim)) {
        lim *= 4.0f;
        ++l;
    }
    return l;
}

__device__ __forceinline__ double level_edge(float s0, int l) { return (double)s0 * (double)(1ull << (2 * l)); }

__device__ __forceinline__ int64_t cell_of(float m, double edge) {
    double q = floor((double)m / edge);
    q = fmin(fmax(q, -1e15), 1e15);
    return (int64_t)q;
}

// Packed cell key: level (4 bits), sub-world (20 bits), cell coordinates
// wrapped to 13 bits each.  Cells of one world that alias under the wrap are
// 8192 cells apart, so their bodies fail the box test; bodies of different
// worlds never share a key.
__device__ __forceinline__ uint64_t cell_key(int l, uint32_t world, uint64_t x, uint64_t y, uint64_t z) {
    return ((uint64_t)l << 60) | ((uint64_t)(world & 0xFFFFFu) << 39) | ((x & 0x1FFFull) << 26) |
           ((y & 0x1FFFull) << 13) | (z & 0x1FFFull);
}

__device__ __forceinline__ float max_side(float4 mn, float4 mx) {
    return fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z);
}

constexpr uint32_t kEndOfList = 0xffffffffu;

// Slot of `key` in the cell table, inserting it if absent (~0 when full).
__device__ __forceinline__ uint32_t table_slot(DeviceWorld& w, uint64_t key) {
    uint32_t mask = w.gt_cap - 1, s = (uint32_t)mix64(key) & mask;
    for (uint32_t p = 0; p < w.gt_cap; ++p) {
        unsigned long long prev = atomicCAS((unsigned long long*)&w.gt_key[s], kEmptyCell, key);
        if (prev == kEmptyCell || prev == key) return s;
        s = (s + 1) & mask;
    }
    return ~0u;
}

// Level and cell of every body; the body joins its cell's list (the table
// maps the packed cell key to the list head; heads start empty every step).
__global__ void __launch_bounds__(256) k_grid_insert(DeviceWorld w) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = b < w.nbodies && (w.bflags[b] & BF_PRESENT);
    uint32_t lmask = 0, world = 0;
    if (alive) {
        float4 mn = w.bmin[b], mx = w.bmax[b];
        float s0 = __uint_as_float(w.bounds[6]);
        int l = grid_level(max_side(mn, mx), s0);
        double e = level_edge(s0, l);
        world = w.wid[b];
        uint64_t key = cell_key(l, world, (uint64_t)cell_of(mn.x, e), (uint64_t)cell_of(mn.y, e),
                                (uint64_t)cell_of(mn.z, e));
        w.cell_key[b] = key;
        uint32_t slot = table_slot(w, key);
        if (slot != ~0u) w.gnext[b] = atomicExch(&w.gt_head[slot], b);
        else w.counters[ctr(CTR_OVERFLOW)] = 6;
        lmask = 1u << l;
    }
    // Bodies per world and level, and the first kLevelList of them: lanes with
    // one (world, level) combine into one atomic (a world's bodies have
    // consecutive ids, so a wavefront holds few distinct keys).
    {
        const int lane = __lane_id();
        uint32_t key = alive ? world * 16u + (uint32_t)(31 - __clz(lmask)) : ~0u;
        uint64_t pending = __ballot(alive);
        while (pending) {
            int lead = __ffsll((unsigned long long)pending) - 1;
            uint32_t k0 = __shfl(key, lead);
            uint64_t same = __ballot(key == k0) & pending;
            uint32_t base = 0;
            if (lane == lead) base = atomicAdd(&w.wlev_count[k0], (uint32_t)__popcll(same));
            base = __shfl(base, lead);
            if ((same >> lane) & 1ull) {
                uint32_t idx = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
                if (idx < kLevelList) w.wlev_list[k0 * kLevelList + idx] = b;
            }
            pending &= ~same;
        }
    }
    // Levels used per world: lanes of the first live lane's world combine
    // their masks into one atomic (a world's bodies have consecutive ids).
    uint64_t live = __ballot(alive);
    if (live) {
        int lead = __ffsll((unsigned long long)live) - 1;
        uint32_t w0 = __shfl(world, lead);
        bool same = alive && world == w0;
        uint32_t m = same ? lmask : 0;
        for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off);
        if (__lane_id() == lead) atomicOr(&w.wlevels[w0], m);
        if (alive && !same) atomicOr(&w.wlevels[world], lmask);
    }
}

// List head of `key`, or kEndOfList when no body sits in that cell.
__device__ __forceinline__ uint32_t table_head(const DeviceWorld& w, uint64_t key, uint64_t first, uint32_t slot) {
    uint32_t mask = w.gt_cap - 1, s = slot;
    uint64_t k = first;
    for (uint32_t p = 0; p < w.gt_cap; ++p) {
        if (k == key) return w.gt_head[s];
        if (k == kEmptyCell) return kEndOfList;
        s = (s + 1) & mask;
        k = w.gt_key[s];
    }
    return kEndOfList;
}

// Own cell first, then the 13 neighbours that are "forward" in (z, y, x) order.
__constant__ int8_t c_fwd[14][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},
                                    {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1},
                                    {1, 0, 1},   {-1, 1, 1}, {0, 1, 1},  {1, 1, 1}};

struct GridQuery {
    float4 qmin, qmax;
    uint32_t body;
    bool dyn;
};

// Tests the bodies of one probed cell; `own` restricts to larger body ids
// (the query's own cell at its own level).  The packed key identifies the
// cell exactly up to the 13-bit coordinate wrap, and cells that alias are
// 8192 cells apart, so their bodies fail the box test.
__device__ __forceinline__ void probe_cell(DeviceWorld& w, const GridQuery& q, uint64_t key, uint64_t first,
                                           uint32_t slot, bool own, uint64_t* buf, uint32_t* buf_n) {
    for (uint32_t c = table_head(w, key, first, slot); c != kEndOfList; c = w.gnext[c]) {
        if (own && c <= q.body) continue;
        float4 cmin = w.bmin[c], cmax = w.bmax[c];
        bool hit = (q.dyn || (w.bflags[c] & 1u)) && overlap(q.qmin, q.qmax, cmin, cmax);
        if (hit) buf_push(w, buf, buf_n, kQueryBuf, true, pair_key(q.body, c));
    }
}

__global__ void __launch_bounds__(kQueryBlock) k_grid_query(DeviceWorld w) {
    __shared__ uint64_t buf[kQueryBuf];
    __shared__ uint32_t buf_n, base_slot;
    if (threadIdx.x == 0) buf_n = 0;
    __syncthreads();
    const uint32_t b = blockIdx.x * kQueryBlock + threadIdx.x;
    if (b < w.nbodies && (w.bflags[b] & BF_PRESENT)) {
        GridQuery q;
        q.body = b;
        q.qmin = w.bmin[b];
        q.qmax = w.bmax[b];
        q.dyn = (w.bflags[b] & 1u) != 0;
        const uint64_t qkey = w.cell_key[b];
        const int l = (int)(qkey >> 60);
        const uint32_t world = w.wid[b];
        const uint32_t mask = w.gt_cap - 1;
        // Own level: own cell and forward neighbours (coordinates wrap with the key).
        {
            const uint64_t cx = (qkey >> 26) & 0x1FFFull, cy = (qkey >> 13) & 0x1FFFull, cz = qkey & 0x1FFFull;
            // The 14 table reads are issued together; the keys and slots are
            // recomputed for the probes rather than held in registers.
            auto key_of = [&](int o) {
                return cell_key(l, world, cx + (uint64_t)(int64_t)c_fwd[o][0], cy + (uint64_t)(int64_t)c_fwd[o][1],
                                cz + (uint64_t)(int64_t)c_fwd[o][2]);
            };
            uint64_t first[14];
#pragma unroll
            for (int o = 0; o < 14; ++o) first[o] = w.gt_key[(uint32_t)mix64(key_of(o)) & mask];
#pragma unroll
            for (int o = 0; o < 14; ++o) {
                const uint64_t pk = key_of(o);
                probe_cell(w, q, pk, first[o], (uint32_t)mix64(pk) & mask, o == 0, buf, &buf_n);
            }
        }
        // Coarser non-empty levels: the bodies themselves when the level holds
        // at most kLevelList of the world's bodies (a pile's ground), else the
        // cells a coarser body overlapping this one can sit in.
        uint32_t coarse = w.wlevels[world] & ~((2u << l) - 1u);
        const float s0 = __uint_as_float(w.bounds[6]);
        while (coarse) {
            const int l2 = __ffs(coarse) - 1;
            coarse &= coarse - 1;
            const uint32_t wl = world * 16u + (uint32_t)l2;
            const uint32_t cnt = w.wlev_count[wl];
            if (cnt <= kLevelList) {
                for (uint32_t k = 0; k < cnt; ++k) {
                    uint32_t c = w.wlev_list[wl * kLevelList + k];
                    float4 cmin = w.bmin[c], cmax = w.bmax[c];
                    bool hit = (q.dyn || (w.bflags[c] & 1u)) && overlap(q.qmin, q.qmax, cmin, cmax);
                    if (hit) buf_push(w, buf, &buf_n, kQueryBuf, true, pair_key(q.body, c));
                }
                continue;
            }
            // A coarser body overlapping this one has its min corner in
            // [qmin - edge, qmax] per axis, so its cell lies between the cell
            // below the query's min corner and the cell of its max corner: 2
            // cells per axis, or 3 when the query straddles a cell boundary.
            const double e = level_edge(s0, l2);
            const int64_t x0 = cell_of(q.qmin.x, e), y0 = cell_of(q.qmin.y, e), z0 = cell_of(q.qmin.z, e);
            // (Clamped: a non-finite box must not drive the loop bounds.)
            const int nx = (int)min(max(cell_of(q.qmax.x, e) - x0, (int64_t)0), (int64_t)1) + 2,
                      ny = (int)min(max(cell_of(q.qmax.y, e) - y0, (int64_t)0), (int64_t)1) + 2,
                      nz = (int)min(max(cell_of(q.qmax.z, e) - z0, (int64_t)0), (int64_t)1) + 2;
            for (int dz = 0; dz < nz; ++dz)
                for (int dy = 0; dy < ny; ++dy) {
                    uint64_t pk[3], first[3];
                    uint32_t ps[3];
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) {
                        pk[dx] = cell_key(l2, world, (uint64_t)(x0 - 1 + dx), (uint64_t)(y0 - 1 + dy),
                                          (uint64_t)(z0 - 1 + dz));
                        ps[dx] = (uint32_t)mix64(pk[dx]) & mask;
                        first[dx] = dx < nx ? w.gt_key[ps[dx]] : kEmptyCell;
                    }
#pragma unroll
                    for (int o = 0; o < 3; ++o)
                        if (o < nx) probe_cell(w, q, pk[o], first[o], ps[o], false, buf, &buf_n);
                }
        }
    }
    buf_flush<kQueryBlock>(w, buf, &buf_n, kQueryBuf, &base_slot);
}

void launch_grid(const DeviceWorld& w, hipStream_t st) {
    uint32_t n = w.nbodies;
    hipLaunchKernelGGL(k_grid_insert, dim3((n + 255) / 256), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_grid_query, dim3((n + kQueryBlock - 1) / kQueryBlock), dim3(kQueryBlock), 0, st, w);
}

}  // namespace bepu
