// k_solver.hip -- global solver path for islands too large for one
// workgroup's LDS (the island kernels in k_island.hip take all others).
//
// Stages, launched without host synchronisation:
//   1. body -> constraint incidence lists (dynamic bodies only);
//   2. colouring (cooperative kernel, rounds separated by grid barriers): a
//      constraint takes the smallest colour unused by its conflicting
//      neighbours of higher priority mix64(key) -- the sequential greedy
//      colouring in decreasing priority, independent of scheduling and the
//      same rule the island kernels apply;
//   3. constraints bucketed by colour;
//   4. prestep, one thread per constraint;
//   5. warm start + iterations (cooperative kernel, one grid barrier per
//      colour): no two constraints of a colour share a dynamic body;
//   6. accumulated impulses written back to the persistent pair state.
#include <hip/hip_cooperative_groups.h>
#include <hipcub/hipcub.hpp>

#include "solver_rows.h"

namespace cg = cooperative_groups;

namespace bepu {

constexpr int kSolveBlock = 256;

__device__ __forceinline__ uint32_t gcount(const DeviceWorld& w) { return min(w.counters[ctr(CTR_GCONS)], w.pair_cap); }

__global__ void k_inc_count(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= gcount(w)) return;
    const PairState& s = w.pairs_cur[w.gcons[c]];
    if (w.bflags[s.a] & 1u) atomicAdd(&w.inc_count[s.a], 1u);
    if (w.bflags[s.b] & 1u) atomicAdd(&w.inc_count[s.b], 1u);
    w.prio[c] = mix64(s.key);
    w.color[c] = -1;
}

// Fills the lists counting down; leaves inc_count at zero for the next step.
__global__ void k_inc_fill(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= gcount(w)) return;
    const PairState& s = w.pairs_cur[w.gcons[c]];
    if (w.bflags[s.a] & 1u) w.inc_list[w.inc_offset[s.a] + atomicSub(&w.inc_count[s.a], 1u) - 1] = c;
    if (w.bflags[s.b] & 1u) w.inc_list[w.inc_offset[s.b] + atomicSub(&w.inc_count[s.b], 1u) - 1] = c;
}

__device__ __forceinline__ int load_color(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns the colour, or -1 while a higher-priority neighbour is uncoloured.
__device__ int try_color(const DeviceWorld& w, uint32_t c) {
    const PairState& s = w.pairs_cur[w.gcons[c]];
    uint64_t p = w.prio[c];
    uint64_t mask0 = 0, mask1 = 0;
    uint32_t bodies[2] = {s.a, s.b};
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

__global__ void __launch_bounds__(kSolveBlock) k_color(DeviceWorld w) {
    cg::grid_group grid = cg::this_grid();
    uint32_t m = gcount(w);
    if (m == 0) return;  // uniform across the grid: no barrier is skipped by a subset
    uint32_t gsize = gridDim.x * blockDim.x;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    // Each round colours at least the highest-priority uncoloured constraint,
    // so m + 1 rounds always suffice.  Three rotating counters keep a round's
    // count stable until every thread has read it.
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
            if (col >= 0) {
                __hip_atomic_store(&w.color[c], col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                ++left;
            }
        }
        if (left) atomicAdd(remaining, left);
        grid.sync();
        if (__hip_atomic_load(remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
    }
}

// Per-colour histogram with one atomic per (block, colour).
__global__ void __launch_bounds__(kSolveBlock) k_color_hist(DeviceWorld w) {
    __shared__ uint32_t h[kMaxColors];
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x) h[k] = 0;
    __syncthreads();
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < gcount(w)) atomicAdd(&h[w.color[c]], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x)
        if (h[k]) atomicAdd(&w.color_count[k], h[k]);
}

// Colour offsets (one wavefront, two colours per lane); resets the counts
// and cursors for the next step.
__global__ void __launch_bounds__(64) k_color_scan(DeviceWorld w) {
    static_assert(kMaxColors == 128, "two colours per lane");
    if (gcount(w) == 0) return;
    int t = threadIdx.x;
    uint32_t c0 = w.color_count[2 * t], c1 = w.color_count[2 * t + 1];
    uint32_t a = c0 + c1;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t x = __shfl_up(a, off);
        if (t >= off) a += x;
    }
    uint32_t base = a - (c0 + c1);
    w.color_start[2 * t] = base;
    w.color_start[2 * t + 1] = base + c0;
    w.color_count[2 * t] = 0;
    w.color_count[2 * t + 1] = 0;
    w.color_cursor[2 * t] = 0;
    w.color_cursor[2 * t + 1] = 0;
    if (t == 63) w.color_start[kMaxColors] = a;
}

// Bucket by colour: block-local ranks, one atomic per (block, colour).
__global__ void __launch_bounds__(kSolveBlock) k_color_scatter(DeviceWorld w) {
    __shared__ uint32_t h[kMaxColors], base[kMaxColors];
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x) h[k] = 0;
    __syncthreads();
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool have = c < gcount(w);
    int col = have ? w.color[c] : 0;
    uint32_t rank = have ? atomicAdd(&h[col], 1u) : 0;
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x)
        if (h[k]) base[k] = atomicAdd(&w.color_cursor[k], h[k]);
    __syncthreads();
    if (have) w.ordered[w.color_start[col] + base[col] + rank] = c;
}

__global__ void __launch_bounds__(kSolveBlock) k_prestep(DeviceWorld w) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= gcount(w)) return;
    uint32_t c = w.ordered[p];
    uint32_t slot = w.gcons[c];
    PairState& s = w.pairs_cur[slot];
    BodyV A = load_body_global(w, s.a), B = load_body_global(w, s.b);
    RowConst r;
    RowVar v;
    prestep_row(w.s, s, f3_of(w.pos[s.a]), f3_of(w.pos[s.b]), A, B, r, v);
    s.color = w.color[c];
    w.grow_c[p] = r;
    w.grow_v[p] = v;
    w.grow_pair[p] = slot;
}

__device__ __forceinline__ void global_phase(const DeviceWorld& w, uint32_t p, bool warm) {
    uint32_t slot = w.grow_pair[p];
    const PairState& s = w.pairs_cur[slot];
    uint32_t a = s.a, b = s.b;
    BodyV A = load_body_global(w, a), B = load_body_global(w, b);
    const RowConst& r = w.grow_c[p];
    RowVar v = w.grow_v[p];
    if (warm) warm_row(r, v, A, B);
    else solve_row(r, v, A, B, w.s.min_impulse, w.s.min_iteration_count);
    w.grow_v[p] = v;
    if (A.dyn) { w.linv[a] = f4_of(A.v, 0.0f); w.angv[a] = f4_of(A.w, 0.0f); }
    if (B.dyn) { w.linv[b] = f4_of(B.v, 0.0f); w.angv[b] = f4_of(B.w, 0.0f); }
}

__global__ void __launch_bounds__(kSolveBlock) k_solve(DeviceWorld w) {
    cg::grid_group grid = cg::this_grid();
    uint32_t gsize = gridDim.x * blockDim.x;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = gcount(w);
    if (m == 0) return;  // uniform across the grid: no barrier is skipped by a subset
    int ncol = 0;
    for (int k = 0; k < kMaxColors; ++k)
        if (w.color_start[k + 1] > w.color_start[k]) ncol = k + 1;
    for (int col = 0; col < ncol; ++col) {
        uint32_t beg = w.color_start[col], end = w.color_start[col + 1];
        for (uint32_t p = beg + tid; p < end; p += gsize) global_phase(w, p, true);
        grid.sync();
    }
    for (int it = 0; it < w.s.iteration_limit; ++it)
        for (int col = 0; col < ncol; ++col) {
            uint32_t beg = w.color_start[col], end = w.color_start[col + 1];
            for (uint32_t p = beg + tid; p < end; p += gsize) global_phase(w, p, false);
            grid.sync();
        }
    for (uint32_t p = tid; p < m; p += gsize) writeback_row(w.pairs_cur[w.grow_pair[p]], w.grow_v[p]);
}

// ---------------------------------------------------------------------------

static int cooperative_grid(const void* kernel, int block) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, dev);
    int per_cu = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
    if (per_cu < 1) per_cu = 1;
    if (per_cu > 2) per_cu = 2;
    return prop.multiProcessorCount * per_cu;
}

size_t solver_scan_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n);
    return bytes;
}

void launch_global_solver(const DeviceWorld& w, void* temp, size_t temp_bytes, hipStream_t st) {
    static int g_color = 0, g_solve = 0;
    if (!g_color) g_color = cooperative_grid((const void*)k_color, kSolveBlock);
    if (!g_solve) g_solve = cooperative_grid((const void*)k_solve, kSolveBlock);
    uint32_t cap = w.pair_cap;
    dim3 g((cap + kSolveBlock - 1) / kSolveBlock), b(kSolveBlock);
    hipLaunchKernelGGL(k_inc_count, g, b, 0, st, w);
    size_t bytes = temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(temp, bytes, w.inc_count, w.inc_offset, (int)(w.nbodies + 1), st);
    hipLaunchKernelGGL(k_inc_fill, g, b, 0, st, w);
    DeviceWorld wc = w;
    void* args[] = {&wc};
    (void)hipLaunchCooperativeKernel((const void*)k_color, dim3(g_color), dim3(kSolveBlock), args, 0, st);
    hipLaunchKernelGGL(k_color_hist, g, b, 0, st, w);
    hipLaunchKernelGGL(k_color_scan, dim3(1), dim3(64), 0, st, w);
    hipLaunchKernelGGL(k_color_scatter, g, b, 0, st, w);
    hipLaunchKernelGGL(k_prestep, g, b, 0, st, w);
    (void)hipLaunchCooperativeKernel((const void*)k_solve, dim3(g_solve), dim3(kSolveBlock), args, 0, st);
}

}  // namespace bepu
