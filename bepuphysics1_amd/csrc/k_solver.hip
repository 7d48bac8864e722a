// k_solver.hip -- global solver path for islands too large for one
// workgroup's LDS (the island kernels in k_island.hip take all others).
//
// One cooperative kernel, phases separated by grid barriers:
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
// launch.
#include <hip/hip_cooperative_groups.h>

#include "solver_rows.h"

namespace cg = cooperative_groups;

namespace bepu {

constexpr int kSolveBlock = 256;
constexpr int kMaxGrid = 4096;   // workgroups of the cooperative grid (scan partials)

__device__ __forceinline__ uint32_t gcount(const DeviceWorld& w) {
    return min(w.counters[ctr(CTR_GCONS)], w.pair_cap);
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

__global__ void __launch_bounds__(kSolveBlock) k_global(DeviceWorld w) {
    __shared__ uint32_t s_scan[kSolveBlock];
    __shared__ uint32_t s_hist[kMaxColors], s_base[kMaxColors], s_start[kMaxColors + 1];
    __shared__ uint32_t s_carry;
    cg::grid_group grid = cg::this_grid();
    const uint32_t m = gcount(w);
    if (m == 0) return;  // uniform across the grid: no barrier is skipped by a subset
    const uint32_t G = gridDim.x, gsize = G * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = w.nbodies;

    // 1. incidence counts, priorities
    for (uint32_t c = tid; c < m; c += gsize) {
        const PairState& s = w.pairs_cur[w.gcons[c]];
        if (w.bflags[s.a] & 1u) atomicAdd(&w.inc_count[s.a], 1u);
        if (w.bflags[s.b] & 1u) atomicAdd(&w.inc_count[s.b], 1u);
        w.prio[c] = mix64(local_pair_key(w, s.key, s.a));
        w.color[c] = -1;
    }
    grid.sync();

    // 2. exclusive scan of inc_count[0..n] (inc_count[n] is zero) into inc_offset
    const uint32_t N = n + 1, chunk = (N + G - 1) / G;
    const uint32_t lo = min(N, blockIdx.x * chunk), hi = min(N, lo + chunk);
    {
        uint32_t part = 0;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) part += w.inc_count[i];
        uint32_t tot = block_scan_incl(part, s_scan);
        if (threadIdx.x == kSolveBlock - 1) w.scan_part[blockIdx.x] = tot;
    }
    grid.sync();
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
    grid.sync();

    // 3. incidence lists (counting down leaves inc_count at zero for the next step)
    for (uint32_t c = tid; c < m; c += gsize) {
        const PairState& s = w.pairs_cur[w.gcons[c]];
        if (w.bflags[s.a] & 1u) w.inc_list[w.inc_offset[s.a] + atomicSub(&w.inc_count[s.a], 1u) - 1] = c;
        if (w.bflags[s.b] & 1u) w.inc_list[w.inc_offset[s.b] + atomicSub(&w.inc_count[s.b], 1u) - 1] = c;
    }
    grid.sync();

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
        grid.sync();
        if (__hip_atomic_load(remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
    }

    // 5. bucket by colour: per-workgroup histogram, one atomic per (workgroup, colour)
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x) s_hist[k] = 0;
    __syncthreads();
    for (uint32_t c = tid; c < m; c += gsize) atomicAdd(&s_hist[w.color[c]], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x)
        if (s_hist[k]) atomicAdd(&w.color_count[k], s_hist[k]);
    grid.sync();
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
    grid.sync();
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
    grid.sync();

    // 7. warm start and iterations, colour by colour
    int ncol = 0;
    for (int k = 0; k < kMaxColors; ++k)
        if (s_start[k + 1] > s_start[k]) ncol = k + 1;
    for (int col = 0; col < ncol; ++col) {
        for (uint32_t p = s_start[col] + tid; p < s_start[col + 1]; p += gsize) global_phase(w, p, true);
        grid.sync();
    }
    for (int it = 0; it < w.s.iteration_limit; ++it)
        for (int col = 0; col < ncol; ++col) {
            for (uint32_t p = s_start[col] + tid; p < s_start[col + 1]; p += gsize) global_phase(w, p, false);
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
    return std::min(kMaxGrid, prop.multiProcessorCount * per_cu);
}

void launch_global_solver(const DeviceWorld& w, hipStream_t st) {
    static int g = 0;
    if (!g) g = cooperative_grid((const void*)k_global, kSolveBlock);
    DeviceWorld wc = w;
    void* args[] = {&wc};
    (void)hipLaunchCooperativeKernel((const void*)k_global, dim3(g), dim3(kSolveBlock), args, 0, st);
}

}  // namespace bepu
