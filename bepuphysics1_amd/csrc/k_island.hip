// k_island.hip -- island decomposition of the contact graph and the
// island-resident solver.
//
// Constraints only couple through dynamic bodies, so the connected components
// of the constraint graph over dynamic bodies (simulation islands,
// DeactivationManagement/SimulationIsland.cs) can be solved independently.
//   * small islands (<= 64 constraints, or a lower bound the launch
//     configuration sets) are packed several to a wavefront:
//     an island of nc constraints belongs to class k = min(64, 64 / nc) and
//     a bundle holds k islands of one class, 64 / k lanes each;
//   * medium islands (<= 256 constraints) take one 256-thread workgroup;
//   * larger islands go to the global path (k_solver.hip).
// Inside a bundle or workgroup, colouring, prestep, warm start and all
// iterations run with the bodies' velocities and inverse inertias in LDS and
// each constraint's rows in registers; colours are separated by workgroup
// barriers.  Every path colours with the same rule and solves colour by
// colour, so the result does not depend on the path an island takes or on
// which islands share a bundle.
#include <hipcub/hipcub.hpp>

#include "solver_rows.h"
#include "wave.h"

namespace bepu {

constexpr int kPackLanes = 64;                  // one wavefront per bundle
constexpr int kPackDyn = 128;                   // dynamic slots: sum of nb <= 64 + islands <= 128
constexpr int kPackSlots = kPackDyn + kPackLanes;
constexpr int kMediumBlock = 256;
constexpr int kMediumDyn = 264;                 // nb <= nc + 1 <= 257
constexpr int kMediumSlots = kMediumDyn + kMediumBlock;
constexpr uint64_t kConsOne = 1ull << 32;       // constraint count lives in the high word

__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }

// Tier of an island with nc constraints and nb dynamic bodies.  A connected
// island has nb <= nc + 1; the check keeps the slot layout safe regardless.
__device__ __forceinline__ int island_tier(uint32_t nc, uint32_t nb, uint32_t small_max) {
    if (nc == 0) return 0;
    if (nb > nc + 1) return 3;
    if (nc <= min(small_max, (uint32_t)kPackLanes)) return 1;
    if (nc <= (uint32_t)kMediumBlock) return 2;
    return 3;
}

// Islands per bundle for a small island (class index = result - 1).
__device__ __forceinline__ uint32_t pack_class(uint32_t nc) { return min(64u, 64u / nc); }

// The solver's bodies and constraints: dynamic bodies of active islands, and
// pairs with contacts that were not removed, of active islands
// (Solver.UpdateSolverActivity, SolverUpdateable.cs:202-218).
__device__ __forceinline__ bool solver_body(uint32_t f) { return (f & BF_DYNAMIC) && !(f & BF_ASLEEP); }
__device__ __forceinline__ bool solver_constraint(const DeviceWorld& w, const PairHot& s) {
    if (s.ccount <= 0 || (s.flags & (PF_DEAD | PF_VOID))) return false;
    return !((w.bflags[s.a] | w.bflags[s.b]) & BF_ASLEEP);
}

// Island of every dynamic body and every constraint (a pair with contacts);
// counts both per island in one packed word (constraints high, bodies low).
// Without deactivation the islands are also the labels of the next step's
// deactivation stage; with it, the labels already are the components of the
// constraints the solver takes (the union-find forest also joined the pairs
// removed as stale, which would leave islands the solver's tiers take for
// unconnected).
__global__ void __launch_bounds__(256) k_uf_roots(DeviceWorld w) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t npairs = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    bool body = i < w.nbodies && solver_body(w.bflags[i]);
    uint32_t rb = body ? (w.sleeping ? w.comp_root[i] : uf_find(w.uf_parent, i)) : 0;
    if (body) w.isl_root[i] = rb;
    if (body && !w.sleeping) w.comp_root[i] = rb;
    (void)wave_atomic_add<unsigned long long>((unsigned long long*)w.isl_count, rb, body, 1ull);
    bool cons = false;
    uint32_t rc = 0;
    if (i < npairs) {
        const PairHot s = w.hot_cur[i];
        cons = solver_constraint(w, s);
        if (cons) {
            const uint32_t d = (w.bflags[s.a] & 1u) ? s.a : s.b;
            rc = w.sleeping ? w.comp_root[d] : uf_find(w.uf_parent, d);
            w.cons_root[i] = rc;
        }
    }
    (void)wave_atomic_add<unsigned long long>((unsigned long long*)w.isl_count, rc, cons, kConsOne);
    // Joints are constraints of the island of their dynamic connection(s).
    bool joint = false;
    uint32_t rj = 0;
    if (i < w.njoints) {
        const JointState& j = w.joints[i];
        joint = solver_joint(w, j);
        if (joint) {
            const uint32_t d = (w.bflags[j.a] & 1u) ? j.a : j.b;
            rj = w.sleeping ? w.comp_root[d] : uf_find(w.uf_parent, d);
            w.cons_root[w.pair_cap + i] = rj;
        }
    }
    (void)wave_atomic_add<unsigned long long>((unsigned long long*)w.isl_count, rj, joint, kConsOne);
}

// Scatters constraints and bodies into island-contiguous lists (counting down
// leaves the packed counts at zero for the next step), classifies islands
// into tiers, lists the global path's constraints and the medium islands,
// and appends small islands to their packing class's list (one global
// atomic per workgroup and class; the workgroups are large because these
// counters are shared by the whole device).
constexpr int kFillBlock = 1024;
__global__ void __launch_bounds__(kFillBlock) k_isl_fill(DeviceWorld w) {
    __shared__ uint32_t s_tmp[kFillBlock / 64 + 1];
    __shared__ uint32_t s_cls[64], s_base[64];
    if (threadIdx.x < 64) s_cls[threadIdx.x] = 0;
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t npairs = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    unsigned long long* cnt = (unsigned long long*)w.isl_count;

    bool cons = i < npairs && solver_constraint(w, w.hot_cur[i]);
    uint32_t rc = cons ? w.cons_root[i] : 0;
    unsigned long long oc = wave_atomic_add<unsigned long long>(cnt, rc, cons, 0ull - kConsOne);
    bool global = false;
    if (cons) {
        uint64_t s0 = w.isl_start[rc], s1 = w.isl_start[rc + 1];
        w.isl_cons[hi32(s0) + hi32(oc) - 1] = i;
        global = island_tier(hi32(s1) - hi32(s0), lo32(s1) - lo32(s0), w.small_island_max) == 3;
    }
    block_push<kFillBlock>(w.gcons, &w.counters[ctr(CTR_GCONS)], global, i, s_tmp);
    // joints: constraint id pair_cap + j
    const bool joint = i < w.njoints && solver_joint(w, w.joints[i]);
    const uint32_t jid = w.pair_cap + i;
    const uint32_t rj = joint ? w.cons_root[jid] : 0;
    unsigned long long oj = wave_atomic_add<unsigned long long>(cnt, rj, joint, 0ull - kConsOne);
    bool jglobal = false;
    if (joint) {
        uint64_t s0 = w.isl_start[rj], s1 = w.isl_start[rj + 1];
        w.isl_cons[hi32(s0) + hi32(oj) - 1] = jid;
        jglobal = island_tier(hi32(s1) - hi32(s0), lo32(s1) - lo32(s0), w.small_island_max) == 3;
    }
    block_push<kFillBlock>(w.gcons, &w.counters[ctr(CTR_GCONS)], jglobal, jid, s_tmp);

    bool body = i < w.nbodies && solver_body(w.bflags[i]);
    uint32_t rb = body ? w.isl_root[i] : 0;
    unsigned long long ob = wave_atomic_add<unsigned long long>(cnt, rb, body, ~0ull);
    bool medium = false, small = false;
    uint32_t k = 0, rank = 0;
    if (body) {
        uint64_t s0 = w.isl_start[rb], s1 = w.isl_start[rb + 1];
        uint32_t pos = lo32(ob) - 1;
        w.isl_bodies[lo32(s0) + pos] = i;
        w.body_local[i] = pos;
        if (rb == i) {
            uint32_t nc = hi32(s1) - hi32(s0);
            int tier = island_tier(nc, lo32(s1) - lo32(s0), w.small_island_max);
            small = tier == 1;
            medium = tier == 2;
            if (small) {
                k = pack_class(nc) - 1;
                rank = atomicAdd(&s_cls[k], 1u);
            }
        }
    }
    block_push<kFillBlock>(w.isl_medium, &w.counters[ctr(CTR_ISL_MEDIUM)], medium, i, s_tmp);
    if (threadIdx.x < 64) {
        uint32_t c = s_cls[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(&w.counters[ctr(CTR_CLS0 + threadIdx.x)], c) : 0;
    }
    __syncthreads();
    if (small) w.cls_list[(size_t)k * w.cls_cap + s_base[k] + rank] = i;
}

// ---------------------------------------------------------------------------
// Island solver core shared by the packed and the medium kernels.

template <int BLOCK, int NSLOT>
struct IslandLds {
    float4 v[NSLOT], w[NSLOT], i0[NSLOT], i1[NSLOT], i2[NSLOT];
    int da[BLOCK], db[BLOCK];   // dynamic body slot of each lane's constraint, -1 if kinematic
    int color[BLOCK];
    uint64_t prio[BLOCK];
    int red[BLOCK / 64];
};

// Barrier of the lanes solving one island group.  A 64-lane group is one
// wavefront, several of which may share a workgroup while working on
// different bundles, so it must not wait for the other wavefronts: the LDS
// accesses of one wavefront complete in issue order, and the fences keep the
// compiler from moving them across this point.  Larger groups are the whole
// workgroup.
template <int BLOCK>
__device__ __forceinline__ void group_sync() {
    if constexpr (BLOCK == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <int BLOCK, int NSLOT>
__device__ __forceinline__ void lds_body_load(IslandLds<BLOCK, NSLOT>& L, const DeviceWorld& w, int l, uint32_t b) {
    L.v[l] = w.linv[b];
    L.w[l] = w.angv[b];
    L.i0[l] = w.ii0[b];
    L.i1[l] = w.ii1[b];
    L.i2[l] = w.ii2[b];
}

// Solves the constraints held by the lanes of one workgroup.  On entry the
// dynamic bodies are in slots [0, DYN), each lane's kinematic body (if any)
// in slot DYN + lane, and L.da / L.db / L.prio are written for every lane;
// the caller has synchronised.  A lane's neighbours (constraints sharing a
// dynamic body) are among lanes [clo, chi).
template <int BLOCK, int DYN, int NSLOT>
__device__ __forceinline__ void solve_lanes(DeviceWorld& w, IslandLds<BLOCK, NSLOT>& L, bool have, uint32_t slot,
                                            uint32_t ga, uint32_t gb, int sa, int sb, int clo, int chi) {
    const int tid = (int)threadIdx.x % BLOCK;
    const bepu_settings& st = w.s;
    // Colouring: smallest colour unused by higher-priority neighbours; rounds
    // until every constraint is set (each round sets the highest-priority
    // uncoloured constraint of every island at least).
    int color = -1;
    const int la = have ? L.da[tid] : -1, lb = have ? L.db[tid] : -1;
    const uint64_t prio = have ? L.prio[tid] : 0;
    if constexpr (BLOCK == 64) {
        // One wavefront: each lane's higher-priority neighbours as a lane mask,
        // found in one pass; a round is then one ballot.
        uint64_t hp = 0;
        if (have)
            for (int k = clo; k < chi; ++k) {
                if (k == tid) continue;
                int ka = L.da[k], kb = L.db[k];
                bool nb_ = (ka >= 0 && (ka == la || ka == lb)) || (kb >= 0 && (kb == la || kb == lb));
                if (nb_ && L.prio[k] >= prio) hp |= 1ull << k;
            }
        while (true) {
            const bool need = have && color < 0;
            const uint64_t pending = __ballot(need);
            if (!pending) break;
            if (need && (hp & pending) == 0) {
                uint64_t m0 = 0, m1 = 0;
                for (uint64_t m = hp; m; m &= m - 1) {
                    int ck = L.color[__ffsll((unsigned long long)m) - 1];
                    if (ck < 64) m0 |= 1ull << ck;
                    else m1 |= 1ull << (ck - 64);
                }
                if (~m0) color = __ffsll((unsigned long long)~m0) - 1;
                else if (~m1) color = 64 + __ffsll((unsigned long long)~m1) - 1;
                else { color = kMaxColors - 1; w.counters[ctr(CTR_OVERFLOW)] = 4; }
                L.color[tid] = color;
            }
            group_sync<BLOCK>();
        }
    } else {
        // The higher-priority neighbours of each lane as a lane mask over the
        // workgroup, found in one pass; a round then visits only them.
        constexpr int kWords = BLOCK / 64;
        uint64_t hp[kWords];
#pragma unroll
        for (int q = 0; q < kWords; ++q) hp[q] = 0;
        if (have)
#pragma unroll
            for (int q = 0; q < kWords; ++q)
                for (int k = max(clo, q * 64); k < min(chi, q * 64 + 64); ++k) {
                    if (k == tid) continue;
                    int ka = L.da[k], kb = L.db[k];
                    bool nb_ = (ka >= 0 && (ka == la || ka == lb)) || (kb >= 0 && (kb == la || kb == lb));
                    if (nb_ && L.prio[k] >= prio) hp[q] |= 1ull << (k - q * 64);
                }
        for (int round = 0; round <= BLOCK; ++round) {
            bool need = have && color < 0;
            if (need) {
                uint64_t m0 = 0, m1 = 0;
                bool ready = true;
#pragma unroll
                for (int q = 0; q < kWords; ++q)
                    for (uint64_t m = hp[q]; ready && m; m &= m - 1) {
                        int k = q * 64 + __ffsll((unsigned long long)m) - 1;
                        int ck = __hip_atomic_load(&L.color[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (ck < 0) { ready = false; break; }
                        if (ck < 64) m0 |= 1ull << ck;
                        else m1 |= 1ull << (ck - 64);
                    }
                if (ready) {
                    if (~m0) color = __ffsll((unsigned long long)~m0) - 1;
                    else if (~m1) color = 64 + __ffsll((unsigned long long)~m1) - 1;
                    else { color = kMaxColors - 1; w.counters[ctr(CTR_OVERFLOW)] = 4; }
                }
            }
            __syncthreads();
            if (need && color >= 0) __hip_atomic_store(&L.color[tid], color, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (!__syncthreads_or(have && color < 0)) break;
        }
    }
    // Colours used = max colour + 1 over the workgroup.
    int ncol = have ? color + 1 : 0;
    for (int off = 32; off > 0; off >>= 1) ncol = max(ncol, __shfl_xor(ncol, off));
    if (BLOCK > 64) {
        if ((tid & 63) == 0) L.red[tid >> 6] = ncol;
        __syncthreads();
        ncol = 0;
        for (int k = 0; k < BLOCK / 64; ++k) ncol = max(ncol, L.red[k]);
    }

    RowConst row;
    RowVar var;
    BodyV A, B;
    auto load = [&](int l, BodyV& x) {
        float4 i0 = L.i0[l], i1 = L.i1[l], i2 = L.i2[l];
        x.v = f3_of(L.v[l]);
        x.w = f3_of(L.w[l]);
        x.I = m3{i0.x, i0.y, i0.z, i1.x, i1.y, i1.z, i2.x, i2.y, i2.z};
        x.im = i0.w;
        x.dyn = l < DYN;
    };
    auto store = [&](int l, const BodyV& x) {
        if (l < DYN) {
            L.v[l] = f4_of(x.v, 0.0f);
            L.w[l] = f4_of(x.w, 0.0f);
        }
    };
    // Prestep reads velocities only; no body changes before every lane is done.
    const bool joint = have && slot >= w.pair_cap;
    if (joint) {
        JointState& J = w.joints[slot - w.pair_cap];
        load(sa, A);
        load(sb, B);
        const float4 qa = w.quat[ga], qb = w.quat[gb];
        prestep_joint(st, J, f3_of(w.pos[ga]), q4{qa.x, qa.y, qa.z, qa.w}, f3_of(w.pos[gb]), q4{qb.x, qb.y, qb.z, qb.w},
                      A, B, row, var);
        J.color = color;
    } else if (have) {
        // On a register copy of the pair state (one load of all of it); the
        // fields the prestep updates (remembered friction directions) and the
        // colour go back.
        PairState s = w.pairs_cur[slot];
        load(sa, A);
        load(sb, B);
        prestep_row(st, s, f3_of(w.pos[ga]), f3_of(w.pos[gb]), A, B, row, var);
        PairState& g = w.pairs_cur[slot];
#pragma unroll
        for (int k = 0; k < 4; ++k) g.frlin[k] = s.frlin[k];
#pragma unroll
        for (int q = 0; q < 6; ++q) g.sllin[q] = s.sllin[q];
        g.color = color;
    }
    group_sync<BLOCK>();
    for (int col = 0; col < ncol; ++col) {
        if (have && color == col) {
            load(sa, A);
            load(sb, B);
            warm_row(row, var, A, B);
            store(sa, A);
            store(sb, B);
        }
        group_sync<BLOCK>();
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
            group_sync<BLOCK>();
        }
    if (joint) writeback_joint(w.joints[slot - w.pair_cap], var);
    else if (have) writeback_row(w.pairs_cur[slot], var);
}

// Per-lane constraint set-up: global bodies, LDS slots (dynamic: base +
// island-local index; kinematic: DYN + lane) and the neighbour-test entries.
template <int BLOCK, int DYN, int NSLOT>
__device__ __forceinline__ void setup_lane(const DeviceWorld& w, IslandLds<BLOCK, NSLOT>& L, bool have, uint32_t slot,
                                           int dbase, uint32_t& ga, uint32_t& gb, int& sa, int& sb) {
    const int tid = (int)threadIdx.x % BLOCK;
    int la = -1, lb = -1;
    ga = gb = 0;
    sa = sb = 0;
    if (have) {
        cons_ends(w, slot, ga, gb);
        if (w.bflags[ga] & 1u) la = dbase + (int)w.body_local[ga];
        if (w.bflags[gb] & 1u) lb = dbase + (int)w.body_local[gb];
        L.prio[tid] = cons_priority(w, slot);
        // A constraint has at most one kinematic body; it gets slot DYN + lane.
        int ext = DYN + tid;
        sa = la >= 0 ? la : ext;
        sb = lb >= 0 ? lb : ext;
        if (la < 0 || lb < 0) lds_body_load(L, w, ext, la < 0 ? ga : gb);
    }
    L.da[tid] = la;
    L.db[tid] = lb;
    L.color[tid] = -1;
}

using PackLds = IslandLds<kPackLanes, kPackSlots>;
using MediumLds = IslandLds<kMediumBlock, kMediumSlots>;
constexpr int kSolveBlock = kMediumBlock;                 // one medium island or four bundles
constexpr int kSolveWaves = kSolveBlock / kPackLanes;
constexpr size_t kSolveLds = sizeof(MediumLds) > kSolveWaves * sizeof(PackLds) ? sizeof(MediumLds)
                                                                              : kSolveWaves * sizeof(PackLds);

// One bundle of same-class small islands on one wavefront.  s_cnt[t] islands
// are in class t (t + 1 islands per bundle); s_bund is the prefix of the
// classes' bundle counts.
__device__ __forceinline__ void pack_bundle(DeviceWorld& w, PackLds& L, const uint32_t* s_cnt, const uint32_t* s_bund,
                                            uint32_t g) {
    const uint32_t lane = threadIdx.x % kPackLanes;
    // Class of bundle g: the last class whose first bundle is <= g.
    int lo = 0, hi = 64;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_bund[mid] <= g) lo = mid;
        else hi = mid;
    }
    const uint32_t per = (uint32_t)lo + 1;
    const uint32_t lanes = 64u / per;               // >= nc of every island of the class
    const uint32_t first = (g - s_bund[lo]) * per;
    const uint32_t p = lane / lanes, q = lane % lanes;
    const bool isl = p < per && first + p < s_cnt[lo];
    uint32_t r = isl ? w.cls_list[(size_t)lo * w.cls_cap + first + p] : 0;
    uint64_t s0 = isl ? w.isl_start[r] : 0, s1 = isl ? w.isl_start[r + 1] : 0;
    const uint32_t cs = hi32(s0), nc = hi32(s1) - cs, bs = lo32(s0), nb = lo32(s1) - bs;
    const int dbase = (int)(p * (lanes + 1));       // dynamic slots [dbase, dbase + nb), nb <= lanes + 1
    for (uint32_t t = q; isl && t < nb; t += lanes) lds_body_load(L, w, dbase + (int)t, w.isl_bodies[bs + t]);
    const bool have = isl && q < nc;
    const uint32_t slot = have ? w.isl_cons[cs + q] : 0;
    uint32_t ga, gb;
    int sa, sb;
    setup_lane<kPackLanes, kPackDyn, kPackSlots>(w, L, have, slot, dbase, ga, gb, sa, sb);
    group_sync<kPackLanes>();
    const int clo = (int)(p * lanes), chi = clo + (int)(isl ? nc : 0);
    solve_lanes<kPackLanes, kPackDyn, kPackSlots>(w, L, have, slot, ga, gb, sa, sb, clo, chi);
    for (uint32_t t = q; isl && t < nb; t += lanes) {
        uint32_t b = w.isl_bodies[bs + t];
        w.linv[b] = L.v[dbase + (int)t];
        w.angv[b] = L.w[dbase + (int)t];
    }
    group_sync<kPackLanes>();
}

// One medium island on the whole workgroup.
__device__ __forceinline__ void medium_island(DeviceWorld& w, MediumLds& L, uint32_t r) {
    const int tid = threadIdx.x;
    uint64_t s0 = w.isl_start[r], s1 = w.isl_start[r + 1];
    const uint32_t cs = hi32(s0), nc = hi32(s1) - cs, bs = lo32(s0), nb = lo32(s1) - bs;
    for (uint32_t t = tid; t < nb; t += kMediumBlock) lds_body_load(L, w, (int)t, w.isl_bodies[bs + t]);
    const bool have = (uint32_t)tid < nc;
    const uint32_t slot = have ? w.isl_cons[cs + tid] : 0;
    uint32_t ga, gb;
    int sa, sb;
    setup_lane<kMediumBlock, kMediumDyn, kMediumSlots>(w, L, have, slot, 0, ga, gb, sa, sb);
    __syncthreads();
    solve_lanes<kMediumBlock, kMediumDyn, kMediumSlots>(w, L, have, slot, ga, gb, sa, sb, 0, (int)nc);
    for (uint32_t t = tid; t < nb; t += kMediumBlock) {
        uint32_t b = w.isl_bodies[bs + t];
        w.linv[b] = L.v[t];
        w.angv[b] = L.w[t];
    }
    __syncthreads();
}

// Small and medium islands in one persistent launch.  A medium island is a
// long chain of dependent colour steps on one workgroup, so the first
// workgroups (at most half of them) take the medium islands while the
// wavefronts of the others take the bundles of small islands: the two kinds
// overlap instead of the medium islands running alone on an idle device after
// the bundles.  Two wavefronts per SIMD (a few registers spill) beat one:
// every bundle of a large scene is resident at once and they hide each
// other's latency.
__global__ void __launch_bounds__(kSolveBlock) __attribute__((amdgpu_waves_per_eu(2))) k_island_solve(DeviceWorld w) {
    __shared__ __align__(16) unsigned char s_lds[kSolveLds];
    __shared__ uint32_t s_cnt[64], s_bund[65];
    const int tid = threadIdx.x;
    if (tid < 64) {
        // Bundles per class (class t packs t + 1 islands) and their prefix.
        uint32_t c = w.counters[ctr(CTR_CLS0 + tid)];
        uint32_t nb = (c + (uint32_t)tid) / ((uint32_t)tid + 1);
        uint32_t a = nb;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t x = __shfl_up(a, off);
            if (tid >= off) a += x;
        }
        s_cnt[tid] = c;
        s_bund[tid] = a - nb;
        if (tid == 63) s_bund[64] = a;
    }
    __syncthreads();
    const uint32_t nmed = w.counters[ctr(CTR_ISL_MEDIUM)];
    const uint32_t mgroups = min(nmed, gridDim.x / 2);
    if (blockIdx.x < mgroups) {
        MediumLds& L = *reinterpret_cast<MediumLds*>(s_lds);
        for (uint32_t it = blockIdx.x; it < nmed; it += mgroups) medium_island(w, L, w.isl_medium[it]);
        return;
    }
    // Bundle order runs across workgroups first, so consecutive bundles (of
    // the same class, hence similar cost) go to different compute units.
    const uint32_t groups = gridDim.x - mgroups, wave = (uint32_t)tid / kPackLanes;
    PackLds& L = reinterpret_cast<PackLds*>(s_lds)[wave];
    const uint32_t total = s_bund[64];
    for (uint32_t g = wave * groups + (blockIdx.x - mgroups); g < total; g += groups * kSolveWaves)
        pack_bundle(w, L, s_cnt, s_bund, g);
}

size_t island_scan_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr, (int)n);
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
    uint32_t n = w.nbodies, cap = std::max(w.pair_cap, w.njoints);
    dim3 b(256), gm((std::max(n, cap) + 255) / 256);
    hipLaunchKernelGGL(k_uf_roots, gm, b, 0, st, w);
    size_t bytes = temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(temp, bytes, (unsigned long long*)w.isl_count,
                                           (unsigned long long*)w.isl_start, (int)(n + 1), st);
    hipLaunchKernelGGL(k_isl_fill, dim3((std::max(n, cap) + kFillBlock - 1) / kFillBlock), dim3(kFillBlock), 0, st, w);
}

void launch_island_solve(const DeviceWorld& w, hipStream_t st) {
    static int g_solve = 0;
    // At least two workgroups: with medium islands present, at least one takes
    // them and at least one takes the bundles.
    if (!g_solve) g_solve = std::max(2, persistent_grid((const void*)k_island_solve, kSolveBlock));
    hipLaunchKernelGGL(k_island_solve, dim3(g_solve), dim3(kSolveBlock), 0, st, w);
}

}  // namespace bepu
