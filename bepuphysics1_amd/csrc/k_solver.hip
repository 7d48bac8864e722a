// k_solver.hip -- global solver path for islands too large for one
// workgroup's LDS (the island kernels in k_island.hip take all others).
//
// One persistent kernel, phases separated by grid barriers:
//   1. body -> constraint incidence counts (dynamic bodies only) and
//      priorities mix64(world-local key);
//   2. exclusive scan of the counts (per-workgroup partials, then every
//      workgroup adds the prefix of the partials before it);
//   3. incidence lists;
//   4. colouring rounds: a constraint takes the smallest colour unused by its
//      conflicting neighbours of higher priority -- the sequential greedy
//      colouring in decreasing priority, independent of scheduling and the
//      same rule the island kernels apply;
//   5. constraints bucketed by colour;
//   6. prestep, one constraint per thread;
//   7. warm start + iterations, one grid barrier per colour: no two
//      constraints of a colour share a dynamic body; accumulated impulses are
//      written back to the persistent pair state.
// With no constraint on this path the kernel returns before its first
// barrier (every workgroup reads the same count), so an empty path costs one
// ordinary launch.
//
// The grid barrier is a counter, not a cooperative launch: a cooperative
// launch costs about 20 us on MI355X whatever the grid, paid every step even
// when the path is empty.  The grid is sized from the kernel's occupancy so
// that all its workgroups are resident together on an idle device; the wait
// at a barrier is bounded in wall-clock time, so a grid that is not resident
// (another stream holding compute units) ends the kernel with an error flag
// instead of hanging.
#include "solver_rows.h"

namespace bepu {

constexpr int kSolveBlock = 256;
constexpr int kMaxGrid = 4096;   // workgroups of the solver grid (scan partials)

// Grid barrier over a resident grid.  Arrivals count up in CTR_GBAR (cleared
// by the step reset; the kernel runs once per step); barrier k is passed when
// (k + 1) * gridDim.x workgroups have arrived.  Every wave fences at agent
// scope on both sides, so all writes before the barrier are seen after it.
// Returns false when the kernel must stop: this wait, or another workgroup's,
// ran past `limit` wall-clock ticks.
struct GridBarrier {
    uint32_t* count;
    uint32_t* abort;
    uint32_t* overflow;
    uint64_t limit;
    uint32_t phase;
};

__device__ __forceinline__ bool grid_barrier(GridBarrier& gb, uint32_t* s_ok) {
    __threadfence();
    __syncthreads();
    gb.phase += 1;
    if (threadIdx.x == 0) {
        const uint32_t target = gb.phase * gridDim.x;
        atomicAdd(gb.count, 1u);
        const uint64_t t0 = (uint64_t)wall_clock64();
        uint32_t ok = 1;
        while ((int32_t)(__hip_atomic_load(gb.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
            if (__hip_atomic_load(gb.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if ((uint64_t)wall_clock64() - t0 > gb.limit) {
                atomicExch(gb.abort, 1u);
                atomicExch(gb.overflow, 8u);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *s_ok = ok;
    }
    __syncthreads();
    __threadfence();
    return *s_ok != 0;
}

__device__ __forceinline__ uint32_t gcount(const DeviceWorld& w) {
    return min(w.counters[ctr(CTR_GCONS)], w.pair_cap + w.joint_cap);
}

__device__ __forceinline__ int load_color(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns the colour, or -1 while a higher-priority neighbour is uncoloured.
__device__ int try_color(const DeviceWorld& w, uint32_t c) {
    uint32_t bodies[2];
    cons_ends(w, w.gcons[c], bodies[0], bodies[1]);
    uint64_t p = w.prio[c];
    uint64_t mask0 = 0, mask1 = 0;
    for (int k = 0; k < 2; ++k) {
        uint32_t b = bodies[k];
        if (!(w.bflags[b] & 1u)) continue;
        uint32_t beg = w.inc_offset[b], end = w.inc_offset[b + 1];
        for (uint32_t q = beg; q < end; ++q) {
            uint32_t n = w.inc_list[q];
            if (n == c || w.prio[n] < p) continue;
            int cn = load_color(&w.color[n]);
            if (cn < 0) return -1;
            if (cn < 64) mask0 |= 1ull << cn;
            else mask1 |= 1ull << (cn - 64);
        }
    }
    if (~mask0) return __ffsll((unsigned long long)~mask0) - 1;
    if (~mask1) return 64 + __ffsll((unsigned long long)~mask1) - 1;
    return -2;
}

__device__ __forceinline__ void global_phase(const DeviceWorld& w, uint32_t p, bool warm) {
    uint32_t slot = w.grow_pair[p];
    uint32_t a, b;
    cons_ends(w, slot, a, b);
    BodyV A = load_body_global(w, a), B = load_body_global(w, b);
    const RowConst& r = w.grow_c[p];
    RowVar v = w.grow_v[p];
    if (warm) warm_row(r, v, A, B);
    else solve_row(r, v, A, B, w.s.min_impulse, w.s.min_iteration_count);
    w.grow_v[p] = v;
    if (A.dyn) { w.linv[a] = f4_of(A.v, 0.0f); w.angv[a] = f4_of(A.w, 0.0f); }
    if (B.dyn) { w.linv[b] = f4_of(B.v, 0.0f); w.angv[b] = f4_of(B.w, 0.0f); }
}

// Workgroup-wide inclusive scan of one value per thread (LDS, 256 threads).
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t v, uint32_t* s) {
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < kSolveBlock; off <<= 1) {
        uint32_t x = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    uint32_t r = s[t];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kSolveBlock) k_global(DeviceWorld w, uint64_t wait_limit) {
    __shared__ uint32_t s_scan[kSolveBlock];
    __shared__ uint32_t s_hist[kMaxColors], s_base[kMaxColors], s_start[kMaxColors + 1];
    __shared__ uint32_t s_carry, s_ok;
    const uint32_t m = gcount(w);
    if (m == 0) return;  // uniform across the grid: no barrier is skipped by a subset
    GridBarrier gb{&w.counters[ctr(CTR_GBAR)], &w.counters[ctr(CTR_GBAR_ABORT)], &w.counters[ctr(CTR_OVERFLOW)],
                   wait_limit, 0};
    const uint32_t G = gridDim.x, gsize = G * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = w.nbodies;

    // 1. incidence counts, priorities
    for (uint32_t c = tid; c < m; c += gsize) {
        uint32_t a, b;
        cons_ends(w, w.gcons[c], a, b);
        if (w.bflags[a] & 1u) atomicAdd(&w.inc_count[a], 1u);
        if (w.bflags[b] & 1u) atomicAdd(&w.inc_count[b], 1u);
        w.prio[c] = cons_priority(w, w.gcons[c]);
        w.color[c] = -1;
    }
    if (!grid_barrier(gb, &s_ok)) return;

    // 2. exclusive scan of inc_count[0..n] (inc_count[n] is zero) into inc_offset
    const uint32_t N = n + 1, chunk = (N + G - 1) / G;
    const uint32_t lo = min(N, blockIdx.x * chunk), hi = min(N, lo + chunk);
    {
        uint32_t part = 0;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) part += w.inc_count[i];
        uint32_t tot = block_scan_incl(part, s_scan);
        if (threadIdx.x == kSolveBlock - 1) w.scan_part[blockIdx.x] = tot;
    }
    if (!grid_barrier(gb, &s_ok)) return;
    {
        uint32_t before = 0;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += blockDim.x) before += w.scan_part[b];
        before = block_scan_incl(before, s_scan);
        if (threadIdx.x == kSolveBlock - 1) s_carry = before;
        __syncthreads();
        for (uint32_t tile = lo; tile < hi; tile += kSolveBlock) {
            uint32_t i = tile + threadIdx.x;
            uint32_t v = i < hi ? w.inc_count[i] : 0;
            uint32_t incl = block_scan_incl(v, s_scan);
            if (i < hi) w.inc_offset[i] = s_carry + incl - v;
            __syncthreads();
            if (threadIdx.x == kSolveBlock - 1) s_carry += incl;
            __syncthreads();
        }
    }
    if (!grid_barrier(gb, &s_ok)) return;

    // 3. incidence lists (counting down leaves inc_count at zero for the next step)
    for (uint32_t c = tid; c < m; c += gsize) {
        uint32_t a, b;
        cons_ends(w, w.gcons[c], a, b);
        if (w.bflags[a] & 1u) w.inc_list[w.inc_offset[a] + atomicSub(&w.inc_count[a], 1u) - 1] = c;
        if (w.bflags[b] & 1u) w.inc_list[w.inc_offset[b] + atomicSub(&w.inc_count[b], 1u) - 1] = c;
    }
    if (!grid_barrier(gb, &s_ok)) return;

    // 4. colouring rounds.  Each round colours at least the highest-priority
    // uncoloured constraint, so m + 1 rounds always suffice.  Three rotating
    // counters keep a round's count stable until every thread has read it.
    for (uint32_t round = 0; round <= m; ++round) {
        uint32_t* remaining = &w.counters[ctr(CTR_UNCOLORED0 + (round % 3))];
        if (tid == 0) w.counters[ctr(CTR_UNCOLORED0 + ((round + 1) % 3))] = 0;
        uint32_t left = 0;
        for (uint32_t c = tid; c < m; c += gsize) {
            if (w.color[c] >= 0) continue;
            int col = try_color(w, c);
            if (col == -2) {
                w.counters[ctr(CTR_OVERFLOW)] = 4;
                col = kMaxColors - 1;
            }
            if (col >= 0) __hip_atomic_store(&w.color[c], col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else ++left;
        }
        for (int off = 32; off > 0; off >>= 1) left += __shfl_xor(left, off);
        if (__lane_id() == 0 && left) atomicAdd(remaining, left);
        if (!grid_barrier(gb, &s_ok)) return;
        if (__hip_atomic_load(remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
    }

    // 5. bucket by colour: per-workgroup histogram, one atomic per (workgroup, colour)
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x) s_hist[k] = 0;
    __syncthreads();
    for (uint32_t c = tid; c < m; c += gsize) atomicAdd(&s_hist[w.color[c]], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x)
        if (s_hist[k]) atomicAdd(&w.color_count[k], s_hist[k]);
    if (!grid_barrier(gb, &s_ok)) return;
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int k = 0; k < kMaxColors; ++k) {
            s_start[k] = acc;
            acc += w.color_count[k];
        }
        s_start[kMaxColors] = acc;
    }
    __syncthreads();
    // Rounds of one item per thread: local ranks, one cursor atomic per colour.
    for (uint32_t first = blockIdx.x * blockDim.x; first < m; first += gsize) {
        for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x) s_hist[k] = 0;
        __syncthreads();
        uint32_t c = first + threadIdx.x;
        bool have = c < m;
        int col = have ? w.color[c] : 0;
        uint32_t rank = have ? atomicAdd(&s_hist[col], 1u) : 0;
        __syncthreads();
        for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x)
            if (s_hist[k]) s_base[k] = atomicAdd(&w.color_cursor[k], s_hist[k]);
        __syncthreads();
        if (have) w.ordered[s_start[col] + s_base[col] + rank] = c;
        __syncthreads();
    }
    if (!grid_barrier(gb, &s_ok)) return;
    // Counts and cursors are no longer read: clear them for the next step.
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x) {
            w.color_count[k] = 0;
            w.color_cursor[k] = 0;
        }

    // 6. prestep in solve order
    for (uint32_t p = tid; p < m; p += gsize) {
        uint32_t c = w.ordered[p];
        uint32_t slot = w.gcons[c];
        RowConst r;
        RowVar v;
        if (slot >= w.pair_cap) {
            JointState& J = w.joints[slot - w.pair_cap];
            BodyV A = load_body_global(w, J.a), B = load_body_global(w, J.b);
            const float4 qa = w.quat[J.a], qb = w.quat[J.b];
            prestep_joint(w.s, J, f3_of(w.pos[J.a]), q4{qa.x, qa.y, qa.z, qa.w}, f3_of(w.pos[J.b]),
                          q4{qb.x, qb.y, qb.z, qb.w}, A, B, r, v);
            J.color = w.color[c];
        } else {
            PairState& s = w.pairs_cur[slot];
            BodyV A = load_body_global(w, s.a), B = load_body_global(w, s.b);
            prestep_row(w.s, s, f3_of(w.pos[s.a]), f3_of(w.pos[s.b]), A, B, r, v);
            s.color = w.color[c];
        }
        w.grow_c[p] = r;
        w.grow_v[p] = v;
        w.grow_pair[p] = slot;
    }
    if (!grid_barrier(gb, &s_ok)) return;

    // 7. warm start and iterations, colour by colour
    int ncol = 0;
    for (int k = 0; k < kMaxColors; ++k)
        if (s_start[k + 1] > s_start[k]) ncol = k + 1;
    for (int col = 0; col < ncol; ++col) {
        for (uint32_t p = s_start[col] + tid; p < s_start[col + 1]; p += gsize) global_phase(w, p, true);
        if (!grid_barrier(gb, &s_ok)) return;
    }
    for (int it = 0; it < w.s.iteration_limit; ++it)
        for (int col = 0; col < ncol; ++col) {
            for (uint32_t p = s_start[col] + tid; p < s_start[col + 1]; p += gsize) global_phase(w, p, false);
            if (!grid_barrier(gb, &s_ok)) return;
        }
    for (uint32_t p = tid; p < m; p += gsize) {
        const uint32_t slot = w.grow_pair[p];
        if (slot >= w.pair_cap) writeback_joint(w.joints[slot - w.pair_cap], w.grow_v[p]);
        else writeback_row(w.pairs_cur[slot], w.grow_v[p]);
    }
}

// ---------------------------------------------------------------------------

// Workgroups that fit on the device at once (at most two per compute unit).
static int resident_grid(const void* kernel, int block) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, dev);
    int per_cu = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
    if (per_cu < 1) per_cu = 1;
    if (per_cu > 2) per_cu = 2;
    return std::min(kMaxGrid, prop.multiProcessorCount * per_cu);
}

void launch_global_solver(const DeviceWorld& w, hipStream_t st) {
    static int g = 0;
    static uint64_t limit = 0;
    if (!g) {
        g = resident_grid((const void*)k_global, kSolveBlock);
        int dev = 0, khz = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
        limit = (uint64_t)khz * 1000u;  // one second per barrier wait
    }
    hipLaunchKernelGGL(k_global, dim3(g), dim3(kSolveBlock), 0, st, w, limit);
}

}  // namespace bepu
