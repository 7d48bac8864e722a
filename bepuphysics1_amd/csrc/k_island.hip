// k_island.hip -- island decomposition of the contact graph and the
// island-resident solver.
//
// Constraints only couple through dynamic bodies, so the connected components
// of the constraint graph over dynamic bodies (simulation islands,
// DeactivationManagement/SimulationIsland.cs) can be solved independently.
// Small and medium islands are each solved by one wavefront / workgroup:
// colouring, prestep, warm start and all iterations run with the island's
// body velocities and inverse inertias in LDS and each constraint's rows in
// registers; colours are separated by workgroup barriers.  Islands too large
// for LDS go to the global path (k_solver.hip).  Both paths colour with the
// same rule and solve colour by colour, so the result does not depend on
// which path an island takes.
#include <hipcub/hipcub.hpp>

#include "solver_rows.h"

namespace bepu {

constexpr int kSmallBlock = 64, kSmallBodies = 64;
constexpr int kMediumBlock = 256, kMediumBodies = 512;

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Union-find with hooking of the larger root under the smaller one, so every
// island's representative is its smallest dynamic body id.
__device__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    while (true) {
        uint32_t p = ld_agent(&parent[x]);
        if (p == x) return x;
        uint32_t g = ld_agent(&parent[p]);
        if (g != p) __hip_atomic_store(&parent[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
    }
}

__global__ void k_uf_init(DeviceWorld w) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < w.nbodies) w.uf_parent[b] = b;
}

__global__ void k_uf_union(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (c >= m) return;
    const PairState& s = w.pairs_cur[w.cons[c]];
    if (!(w.bflags[s.a] & 1u) || !(w.bflags[s.b] & 1u)) return;
    uint32_t a = s.a, b = s.b;
    while (true) {
        a = uf_find(w.uf_parent, a);
        b = uf_find(w.uf_parent, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        uint32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(&w.uf_parent[a], &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

// Final island id of every dynamic body; counts bodies and constraints per island.
__global__ void k_uf_compress(DeviceWorld w) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nbodies) return;
    if (!(w.bflags[b] & 1u)) return;
    uint32_t r = uf_find(w.uf_parent, b);
    w.isl_root[b] = r;
    atomicAdd(&w.isl_bcount[r], 1u);
}

__device__ __forceinline__ uint32_t cons_island(const DeviceWorld& w, const PairState& s) {
    return (w.bflags[s.a] & 1u) ? w.isl_root[s.a] : w.isl_root[s.b];
}

__global__ void k_isl_count_cons(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (c >= m) return;
    atomicAdd(&w.isl_ccount[cons_island(w, w.pairs_cur[w.cons[c]])], 1u);
}

// Scatter constraints and bodies into island-contiguous lists (counting down
// leaves the counts at zero for the next step).
__global__ void k_isl_fill(DeviceWorld w) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (i < m) {
        uint32_t slot = w.cons[i];
        uint32_t r = cons_island(w, w.pairs_cur[slot]);
        w.isl_cons[w.isl_cstart[r] + atomicSub(&w.isl_ccount[r], 1u) - 1] = slot;
    }
    if (i < w.nbodies && (w.bflags[i] & 1u)) {
        uint32_t r = w.isl_root[i];
        uint32_t pos = atomicSub(&w.isl_bcount[r], 1u) - 1;
        w.isl_bodies[w.isl_bstart[r] + pos] = i;
        w.body_local[i] = pos;
    }
}

__device__ __forceinline__ void wave_push(uint32_t* list, uint32_t* counter, bool take, uint32_t value) {
    uint64_t active = __ballot(1);
    uint64_t mask = __ballot(take);
    if (!mask) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)active) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (take) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// Classifies islands by size: small (wavefront), medium (workgroup), large (global path).
__global__ void k_isl_list(DeviceWorld w) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    bool root = r < w.nbodies && (w.bflags[r] & 1u) && w.isl_root[r] == r;
    uint32_t nc = root ? w.isl_cstart[r + 1] - w.isl_cstart[r] : 0;
    uint32_t nb = root ? w.isl_bstart[r + 1] - w.isl_bstart[r] : 0;
    bool small = nc > 0 && nc <= (uint32_t)kSmallBlock && nb <= (uint32_t)kSmallBodies;
    bool medium = nc > 0 && !small && nc <= (uint32_t)kMediumBlock && nb <= (uint32_t)kMediumBodies;
    bool large = nc > 0 && !small && !medium;
    if (r < w.nbodies) w.isl_tier[r] = small ? 1 : (medium ? 2 : (large ? 3 : 0));
    wave_push(w.isl_small, &w.counters[CTR_ISL_SMALL], small, r);
    wave_push(w.isl_medium, &w.counters[CTR_ISL_MEDIUM], medium, r);
    wave_push(w.isl_large, &w.counters[CTR_ISL_LARGE], large, r);
}

// Constraints of large islands form the global path's list.
__global__ void k_isl_large_cons(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    bool take = false;
    uint32_t slot = 0;
    if (c < m) {
        slot = w.cons[c];
        take = w.isl_tier[cons_island(w, w.pairs_cur[slot])] == 3;
    }
    wave_push(w.gcons, &w.counters[CTR_GCONS], take, slot);
}

// ---------------------------------------------------------------------------
// Island solver: one island per workgroup (BLOCK threads, one constraint per
// thread), persistent over the island list of its tier.

template <int BLOCK, int MAXB>
__global__ void __launch_bounds__(BLOCK) k_island(DeviceWorld w, int tier) {
    // Body slots: [0, nb) the island's dynamic bodies (written back at the
    // end), then one read-only slot per constraint that touches a kinematic body.
    __shared__ float4 s_v[MAXB + BLOCK], s_w[MAXB + BLOCK], s_i0[MAXB + BLOCK], s_i1[MAXB + BLOCK],
        s_i2[MAXB + BLOCK];
    __shared__ int s_color[BLOCK];
    __shared__ int s_la[BLOCK], s_lb[BLOCK];
    __shared__ uint64_t s_prio[BLOCK];
    const uint32_t* list = tier == 1 ? w.isl_small : w.isl_medium;
    uint32_t count = w.counters[tier == 1 ? CTR_ISL_SMALL : CTR_ISL_MEDIUM];
    const int tid = threadIdx.x;
    const bepu_settings& st = w.s;
    for (uint32_t it = blockIdx.x; it < count; it += gridDim.x) {
        uint32_t r = list[it];
        uint32_t cs = w.isl_cstart[r], nc = w.isl_cstart[r + 1] - cs;
        uint32_t bs = w.isl_bstart[r], nb = w.isl_bstart[r + 1] - bs;
        for (uint32_t t = tid; t < nb; t += BLOCK) {
            uint32_t b = w.isl_bodies[bs + t];
            s_v[t] = w.linv[b];
            s_w[t] = w.angv[b];
            s_i0[t] = w.ii0[b];
            s_i1[t] = w.ii1[b];
            s_i2[t] = w.ii2[b];
        }
        bool have = (uint32_t)tid < nc;
        uint32_t slot = have ? w.isl_cons[cs + tid] : 0;
        uint32_t ga = 0, gb = 0;
        int la = -1, lb = -1;   // island-local index of a dynamic body, -1 if kinematic
        int sa = 0, sb = 0;     // LDS body slot used by the solve phases
        if (have) {
            const PairState& s = w.pairs_cur[slot];
            ga = s.a;
            gb = s.b;
            if (w.bflags[ga] & 1u) la = (int)w.body_local[ga];
            if (w.bflags[gb] & 1u) lb = (int)w.body_local[gb];
            s_prio[tid] = mix64(s.key);
            // A constraint has at most one kinematic body; it gets slot nb + tid.
            int ext = (int)nb + tid;
            sa = la >= 0 ? la : ext;
            sb = lb >= 0 ? lb : ext;
            if (la < 0 || lb < 0) {
                uint32_t g = la < 0 ? ga : gb;
                s_v[ext] = w.linv[g];
                s_w[ext] = w.angv[g];
                s_i0[ext] = w.ii0[g];
                s_i1[ext] = w.ii1[g];
                s_i2[ext] = w.ii2[g];
            }
        }
        s_la[tid] = la;
        s_lb[tid] = lb;
        s_color[tid] = -1;
        __syncthreads();
        // Colouring: smallest colour unused by higher-priority neighbours
        // (constraints sharing a dynamic body); rounds until every one is set.
        int color = -1;
        uint64_t prio = have ? s_prio[tid] : 0;
        for (int round = 0; round <= (int)nc; ++round) {
            bool need = have && color < 0;
            if (need) {
                uint64_t m0 = 0, m1 = 0;
                bool ready = true;
                for (uint32_t k = 0; k < nc; ++k) {
                    if ((int)k == tid) continue;
                    int ka = s_la[k], kb = s_lb[k];
                    bool nb_ = (ka >= 0 && (ka == la || ka == lb)) || (kb >= 0 && (kb == la || kb == lb));
                    if (!nb_ || s_prio[k] < prio) continue;
                    int ck = __hip_atomic_load(&s_color[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (ck < 0) { ready = false; break; }
                    if (ck < 64) m0 |= 1ull << ck;
                    else m1 |= 1ull << (ck - 64);
                }
                if (ready) {
                    if (~m0) color = __ffsll((unsigned long long)~m0) - 1;
                    else if (~m1) color = 64 + __ffsll((unsigned long long)~m1) - 1;
                    else { color = kMaxColors - 1; w.counters[CTR_OVERFLOW] = 4; }
                }
            }
            __syncthreads();
            if (need && color >= 0) __hip_atomic_store(&s_color[tid], color, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (!__syncthreads_or(have && color < 0)) break;
        }
        int ncol = __syncthreads_or(0) ? 0 : 0;
        {
            // number of colours of this island = max colour + 1
            int mine = have ? color + 1 : 0;
            for (int off = 32; off > 0; off >>= 1) mine = max(mine, __shfl_xor(mine, off));
            if (BLOCK > 64) {
                if ((tid & 63) == 0) s_color[tid >> 6] = mine;
                __syncthreads();
                mine = 0;
                for (int k = 0; k < BLOCK / 64; ++k) mine = max(mine, s_color[k]);
            }
            ncol = mine;
        }
        if (tid == 0) atomicMax(&w.counters[CTR_COLORS], (uint32_t)ncol);
        // Prestep (reads velocities; no body is modified before every thread is done).
        RowConst row;
        RowVar var;
        BodyV A, B;
        const int nbi = (int)nb;
        auto load = [&](int l, BodyV& x) {
            float4 i0 = s_i0[l], i1 = s_i1[l], i2 = s_i2[l];
            x.v = f3_of(s_v[l]);
            x.w = f3_of(s_w[l]);
            x.I = m3{i0.x, i0.y, i0.z, i1.x, i1.y, i1.z, i2.x, i2.y, i2.z};
            x.im = i0.w;
            x.dyn = l < nbi;
        };
        auto store = [&](int l, const BodyV& x) {
            if (l < nbi) {
                s_v[l] = f4_of(x.v, 0.0f);
                s_w[l] = f4_of(x.w, 0.0f);
            }
        };
        if (have) {
            PairState& s = w.pairs_cur[slot];
            load(sa, A);
            load(sb, B);
            prestep_row(st, s, f3_of(w.pos[ga]), f3_of(w.pos[gb]), A, B, row, var);
            s.color = color;
        }
        __syncthreads();
        for (int col = 0; col < ncol; ++col) {
            if (have && color == col) {
                load(sa, A);
                load(sb, B);
                warm_row(row, var, A, B);
                store(sa, A);
                store(sb, B);
            }
            __syncthreads();
        }
        const float mi = st.min_impulse;
        const int mic = st.min_iteration_count;
        for (int iter = 0; iter < st.iteration_limit; ++iter)
            for (int col = 0; col < ncol; ++col) {
                if (have && color == col) {
                    load(sa, A);
                    load(sb, B);
                    solve_row(row, var, A, B, mi, mic);
                    store(sa, A);
                    store(sb, B);
                }
                __syncthreads();
            }
        if (have) writeback_row(w.pairs_cur[slot], var);
        for (uint32_t t = tid; t < nb; t += BLOCK) {
            uint32_t b = w.isl_bodies[bs + t];
            w.linv[b] = s_v[t];
            w.angv[b] = s_w[t];
        }
        __syncthreads();
    }
}

size_t island_scan_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n);
    return bytes;
}

static int persistent_grid(const void* kernel, int block) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, dev);
    int per_cu = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
    if (per_cu < 1) per_cu = 1;
    return prop.multiProcessorCount * per_cu;
}

void launch_islands(const DeviceWorld& w, void* temp, size_t temp_bytes, hipStream_t st) {
    uint32_t n = w.nbodies, cap = w.pair_cap;
    dim3 gb((n + 255) / 256), gc((cap + 255) / 256), b(256);
    hipLaunchKernelGGL(k_uf_init, gb, b, 0, st, w);
    hipLaunchKernelGGL(k_uf_union, gc, b, 0, st, w);
    hipLaunchKernelGGL(k_uf_compress, gb, b, 0, st, w);
    hipLaunchKernelGGL(k_isl_count_cons, gc, b, 0, st, w);
    size_t bytes = temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(temp, bytes, w.isl_ccount, w.isl_cstart, (int)(n + 1), st);
    bytes = temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(temp, bytes, w.isl_bcount, w.isl_bstart, (int)(n + 1), st);
    dim3 gf((std::max(n, cap) + 255) / 256);
    hipLaunchKernelGGL(k_isl_fill, gf, b, 0, st, w);
    hipLaunchKernelGGL(k_isl_list, gb, b, 0, st, w);
    hipLaunchKernelGGL(k_isl_large_cons, gc, b, 0, st, w);
}

void launch_island_solve(const DeviceWorld& w, hipStream_t st) {
    static int g_small = 0, g_medium = 0;
    if (!g_small) g_small = persistent_grid((const void*)k_island<kSmallBlock, kSmallBodies>, kSmallBlock);
    if (!g_medium) g_medium = persistent_grid((const void*)k_island<kMediumBlock, kMediumBodies>, kMediumBlock);
    hipLaunchKernelGGL((k_island<kSmallBlock, kSmallBodies>), dim3(g_small), dim3(kSmallBlock), 0, st, w, 1);
    hipLaunchKernelGGL((k_island<kMediumBlock, kMediumBodies>), dim3(g_medium), dim3(kMediumBlock), 0, st, w, 2);
}

}  // namespace bepu
