// k_sleep.hip -- deactivation (sleeping) of simulation islands.
//
// The reference keeps simulation islands incrementally (merges when a
// constraint is added, split attempts when one is removed, a parent chain
// between merged islands) and deactivates a few islands per frame
// (DeactivationManagement/DeactivationManager.cs, SimulationIsland.cs,
// SimulationIslandMember.cs).  Here an island is a connected component of
// the constraints over dynamic bodies, recomputed every step by the union-find
// the solver already builds, and every stage is a data-parallel pass:
//
//   start of the step (DeactivationManager.Update, before the force update):
//     k_deact_wake     islands with a member activated since the last step,
//                      or touched through a constraint by an active
//                      kinematic body, are active (the latter may not
//                      deactivate at their next check);
//     k_deact_members  kinematic flags; candidacy of the members of active
//                      islands; per island: members, candidates, activity;
//     k_deact_turn_m + scan + k_deact_turn
//                      the islands each sub-world looks at this step: in
//                      order of their smallest member after the last island
//                      looked at, while fewer than MaximumDeactivation-
//                      AttemptsPerFrame members were looked at; an active
//                      island of candidates goes to sleep.  k_body_pre
//                      applies the decisions and moves the cursors.
//   after the narrow phase:
//     k_wake_mark / k_wake_apply
//                      an island of this step's constraints with an awake
//                      member or an active kinematic body wakes as a whole
//                      (DeactivationManager.Merge); frozen pairs of a woken
//                      island are stale (NarrowPhase.cs:300-333) and removed;
//     k_live_union / k_live_roots
//                      the components of the constraints left (only when a
//                      pair was removed): the labels of the next step;
//     k_split_wake     a constraint of the last step that is gone, whose two
//                      dynamic bodies are now in different islands, activates
//                      both (DeactivationManager.TryToSplit).
//
// All of it is HBM-bound per-body and per-pair work (a few words each).
#include <hipcub/hipcub.hpp>

#include "wave.h"
#include "world.h"

namespace bepu {

// Lanes sharing the first valid lane's key are combined into one atomic.
__device__ __forceinline__ void wave_atomic_min_u32(uint32_t* arr, uint32_t key, bool valid, uint32_t v) {
    uint64_t vmask = __ballot(valid);
    if (!vmask) return;
    const int lead = __ffsll((unsigned long long)vmask) - 1;
    const uint32_t k0 = __shfl(key, lead);
    const bool same = valid && key == k0;
    uint32_t m = same ? v : 0xffffffffu;
    for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off));
    if (__lane_id() == lead) atomicMin(&arr[k0], m);
    if (valid && !same) atomicMin(&arr[key], v);
}

__device__ __forceinline__ void wave_atomic_or_u32(uint32_t* arr, uint32_t key, bool valid, uint32_t v) {
    uint64_t vmask = __ballot(valid);
    if (!vmask) return;
    const int lead = __ffsll((unsigned long long)vmask) - 1;
    const uint32_t k0 = __shfl(key, lead);
    const bool same = valid && key == k0;
    uint32_t m = same ? v : 0u;
    for (int off = 32; off > 0; off >>= 1) m |= (uint32_t)__shfl_xor((int)m, off);
    if (__lane_id() == lead && m) atomicOr(&arr[k0], m);
    if (valid && !same && v) atomicOr(&arr[key], v);
}

__device__ __forceinline__ float body_velocity_sq(const DeviceWorld& w, uint32_t i) {
    return lensq3(f3_of(w.linv[i])) + lensq3(f3_of(w.angv[i]));
}

// A body of the last step (not one inserted for this step into a reused id).
__device__ __forceinline__ bool carried(const DeviceWorld& w, uint32_t i) {
    return (w.bflags[i] & BF_PRESENT) && w.born[i] != step_index(w);
}

// Islands activated since the last step, and islands an active kinematic
// body touches through a constraint of the last step
// (SimulationIslandMember.cs:105-139: the kinematic member's candidacy update
// sets IsActive and forbids the next deactivation of its connected islands).
__global__ void __launch_bounds__(256) k_deact_wake(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.nbodies) {
        const uint32_t f = w.bflags[i];
        if ((f & BF_DYNAMIC) && (f & BF_WAKE)) w.lab_act[w.comp_root[i]] = 1u;
    }
    // A kinematic connection whose flag becomes active (the candidacy update
    // below) forbids the next deactivation of the island it touches and
    // activates it: through a constraint of the last step, or a joint.
    auto kinematic_rule = [&](uint32_t a, uint32_t b) {
        for (int side = 0; side < 2; ++side) {
            const uint32_t k = side ? b : a, d = side ? a : b;
            const uint32_t fk = w.bflags[k], fd = w.bflags[d];
            if ((fk & BF_DYNAMIC) || !(fd & BF_DYNAMIC)) continue;
            const float next = kinematic_flag_next(w.stab[k].y, body_velocity_sq(w, k), (fk & BF_ALWAYS_ACTIVE) != 0);
            if (next <= 0.0f) {
                w.lab_act[w.comp_root[d]] = 1u;
                atomicOr(&w.bflags[d], (uint32_t)BF_NODEACT);
            }
        }
    };
    const uint32_t np = min(w.counters[ctr(CTR_PAIRS_PREV)], w.pair_cap);
    if (i < np) {
        const PairHot s = w.hot_old[i];
        if (s.ccount > 0 && !(s.flags & (PF_DEAD | PF_VOID)) && carried(w, s.a) && carried(w, s.b))
            kinematic_rule(s.a, s.b);
    }
    if (i < w.njoints) {
        const JointState& j = w.joints[i];
        if (!(j.flags & JF_DEAD) && (w.bflags[j.a] & BF_PRESENT) && (w.bflags[j.b] & BF_PRESENT))
            kinematic_rule(j.a, j.b);
    }
}

// Candidacy of every member (SimulationIslandMember.UpdateDeactivationCandidacy,
// SimulationIslandMember.cs:53-139): kinematic members update their flag;
// dynamic members of active islands run the dynamic branch, the others only
// remember their speed.  Counts members and candidates per island.
__global__ void __launch_bounds__(256) k_deact_members(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bepu_settings& s = w.s;
    bool dyn = false, cand = false, active = false, nodeact = false;
    uint32_t L = 0;
    if (i < w.nbodies) {
        const uint32_t f = w.bflags[i];
        const bool valid = (f & BF_PRESENT) && !w.wfault[w.wid[i]];
        if (valid && !(f & BF_DYNAMIC)) {
            float4 st = w.stab[i];
            const float next = kinematic_flag_next(st.y, body_velocity_sq(w, i), (f & BF_ALWAYS_ACTIVE) != 0);
            if (next != st.y) w.stab[i].y = next;
            const uint32_t nf = next <= 0.0f ? (f | BF_KIN_ACTIVE) : (f & ~(uint32_t)BF_KIN_ACTIVE);
            if (nf != f) w.bflags[i] = nf;
        } else if (valid) {
            dyn = true;
            L = w.comp_root[i];
            active = !(f & BF_ASLEEP) || w.lab_act[L] != 0u;
            nodeact = (f & BF_NODEACT) != 0;
            float4 st = w.stab[i];
            const float velocity = body_velocity_sq(w, i);
            uint32_t nf = f & ~(uint32_t)BF_WAKE;
            if (active) {
                const float lim = s.velocity_lower_limit;
                const float slowing = velocity <= st.x ? 1.0f : 0.0f;
                float tb = st.y;
                if (velocity < lim * lim) tb += s.dt;
                else tb = 0;
                uint32_t c = f & BF_CANDIDATE;
                if (!(f & BF_ALWAYS_ACTIVE)) {
                    if (!c) {
                        if (tb > s.low_velocity_time_minimum && slowing != 0.0f) c = BF_CANDIDATE;
                    } else if (tb <= s.low_velocity_time_minimum) {
                        c = 0;
                        tb = 0;
                    }
                } else {
                    c = 0;
                    tb = 0;
                }
                nf = (nf & ~(uint32_t)(BF_CANDIDATE | BF_ASLEEP)) | c;
                w.stab[i] = make_float4(velocity, tb, slowing, st.w);
            } else {
                w.stab[i].x = velocity;
            }
            cand = (nf & BF_CANDIDATE) != 0;
            if (nf != f) w.bflags[i] = nf;
            if (!(f & BF_ALWAYS_ACTIVE)) w.wsleepy[w.wid[i]] = 1u;
        }
    }
    // wave_atomic_add combines lanes that add the same increment: members and
    // candidates are counted by separate calls.
    (void)wave_atomic_add<unsigned long long>((unsigned long long*)w.lab_count, L, dyn, 1ull);
    (void)wave_atomic_add<unsigned long long>((unsigned long long*)w.lab_count, L, dyn && cand, 1ull << 32);
    wave_atomic_min_u32(w.lab_min, L, dyn, i);
    wave_atomic_or_u32(w.lab_flags, L, dyn, (active ? (uint32_t)LF_ACTIVE : 0u) | (nodeact ? (uint32_t)LF_NODEACT : 0u));
}

__device__ __forceinline__ bool island_rep(const DeviceWorld& w, uint32_t i, uint32_t* label) {
    const uint32_t f = w.bflags[i];
    if (!(f & BF_DYNAMIC) || w.wfault[w.wid[i]]) return false;
    const uint32_t L = w.comp_root[i];
    *label = L;
    return w.lab_min[L] == i;
}

// Members of each island at its smallest member's position (0 elsewhere).
__global__ void __launch_bounds__(256) k_deact_turn_m(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > w.nbodies) return;
    uint32_t m = 0, L = 0;
    if (i < w.nbodies && island_rep(w, i, &L)) m = (uint32_t)w.lab_count[L];
    w.deact_m[i] = m;
}

// DeactivationManager.DeactivateObjects (DeactivationManager.cs:291-320) in
// parallel: the islands of a sub-world in order of their smallest member,
// starting after the last island the previous step looked at and wrapping;
// an island is looked at when the members of the islands before it in that
// order add up to less than the budget.  An island looked at may sleep
// (SimulationIsland.TryToDeactivate, SimulationIsland.cs:92-114): active, no
// deactivation forbidden, every member a candidate.
__global__ void __launch_bounds__(256) k_deact_turn(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nbodies) return;
    uint32_t L = 0;
    if (!island_rep(w, i, &L)) return;
    const uint32_t k = w.wid[i];
    if (!w.wsleepy[k]) return;
    const uint32_t f = w.wrange[3 * k], cnt = w.wrange[3 * k + 1], gen = w.wrange[3 * k + 2];
    if (i < f || i - f >= cnt) return;  // not possible for a live sub-world's body
    uint32_t c = w.wcurgen[k] == gen ? w.wcursor[k] : kNoCursor;
    if (c != kNoCursor && (c < f || c - f >= cnt)) c = kNoCursor;
    const uint32_t* E = w.deact_e;
    uint64_t before;
    uint32_t cyc;
    if (c == kNoCursor) {
        before = E[i] - E[f];
        cyc = i - f;
    } else if (i > c) {
        before = E[i] - E[c + 1];
        cyc = i - c;
    } else {
        before = (uint64_t)(E[f + cnt] - E[c + 1]) + (uint64_t)(E[i] - E[f]);
        cyc = i + cnt - c;
    }
    const int64_t limit = w.s.max_deactivation_attempts;
    if ((int64_t)before >= limit) return;
    const uint64_t cm = w.lab_count[L];
    uint32_t fl = w.lab_flags[L] | LF_CHECKED;
    if (!(fl & LF_NODEACT) && (fl & LF_ACTIVE) && (uint32_t)(cm >> 32) == (uint32_t)cm) fl |= LF_SLEEP;
    w.lab_flags[L] = fl;
    // The last island looked at (the next one is over the budget, or this is
    // the last of the turn) moves the cursor: one writer per sub-world.
    const uint64_t after = before + (uint32_t)cm;
    if ((int64_t)after >= limit || after == (uint64_t)(E[f + cnt] - E[f])) w.wkey[k] = cyc + 1;
}

// After the narrow phase: an island of this step's constraints (frozen ones
// included) is active when a member is awake or one of its constraints holds
// an active kinematic body.  Also moves each sub-world's turn cursor to the
// last island its deactivation stage looked at.
__global__ void __launch_bounds__(256) k_wake_mark(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.nworlds) {
        const uint32_t key = w.wkey[i];
        if (key) {
            const uint32_t f = w.wrange[3 * i], cnt = w.wrange[3 * i + 1], gen = w.wrange[3 * i + 2];
            uint32_t c = w.wcurgen[i] == gen ? w.wcursor[i] : kNoCursor;
            if (c != kNoCursor && (c < f || c - f >= cnt)) c = kNoCursor;
            uint32_t nc = (c == kNoCursor ? f : c) + (key - 1);
            if (nc - f >= cnt) nc -= cnt;
            w.wcursor[i] = nc;
            w.wcurgen[i] = gen;
            w.wkey[i] = 0u;
        }
    }
    if (i < w.nbodies) {
        const uint32_t f = w.bflags[i];
        if ((f & BF_DYNAMIC) && !(f & BF_ASLEEP)) w.act2[uf_find(w.uf_parent, i)] = 1u;
    }
    auto kinematic_wake = [&](uint32_t a, uint32_t b) {
        const uint32_t fa = w.bflags[a], fb = w.bflags[b];
        if (!(fa & BF_DYNAMIC) && (fa & BF_KIN_ACTIVE) && (fb & BF_DYNAMIC)) w.act2[uf_find(w.uf_parent, b)] = 1u;
        if (!(fb & BF_DYNAMIC) && (fb & BF_KIN_ACTIVE) && (fa & BF_DYNAMIC)) w.act2[uf_find(w.uf_parent, a)] = 1u;
    };
    const uint32_t np = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    if (i < np) {
        const PairHot s = w.hot_cur[i];
        if (s.ccount > 0 && !(s.flags & (PF_DEAD | PF_VOID))) kinematic_wake(s.a, s.b);
    }
    if (i < w.njoints && !(w.joints[i].flags & JF_DEAD)) kinematic_wake(w.joints[i].a, w.joints[i].b);
}

// Wakes the members of active islands; a frozen pair with an active body is
// stale and removed (RemoveStaleOverlaps: not updated, one entry active).
__global__ void __launch_bounds__(256) k_wake_apply(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.nbodies) {
        const uint32_t f = w.bflags[i];
        if (f & BF_DYNAMIC) {
            const uint32_t r = uf_find(w.uf_parent, i);
            const bool act = w.act2[r] != 0u;
            if (act && (f & BF_ASLEEP)) w.bflags[i] = f & ~(uint32_t)BF_ASLEEP;
            w.comp_root[i] = r;
            w.uf2[i] = act ? i : r;
        }
    }
    const uint32_t np = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    bool dead = false;
    if (i < np) {
        const PairHot s = w.hot_cur[i];
        if ((s.flags & PF_FROZEN) && !(s.flags & PF_VOID)) {
            const uint32_t fa = w.bflags[s.a], fb = w.bflags[s.b];
            const bool aa = (fa & BF_DYNAMIC) ? w.act2[uf_find(w.uf_parent, s.a)] != 0u : (fa & BF_KIN_ACTIVE) != 0;
            const bool ab = (fb & BF_DYNAMIC) ? w.act2[uf_find(w.uf_parent, s.b)] != 0u : (fb & BF_KIN_ACTIVE) != 0;
            if (aa || ab) {
                w.pairs_cur[i].flags |= PF_DEAD;
                w.hot_cur[i].flags = s.flags | PF_DEAD;
                dead = true;
            }
        }
    }
    const uint64_t m = __ballot(dead);
    if (m && __lane_id() == __ffsll((unsigned long long)m) - 1)
        atomicAdd(&w.counters[ctr(CTR_DEAD)], (uint32_t)__popcll(m));
}

// Components of the constraints left, when a pair was removed: inactive
// islands keep their roots, active ones are joined again over their live
// constraints.
__global__ void __launch_bounds__(256) k_live_union(DeviceWorld w) {
    if (w.counters[ctr(CTR_DEAD)] == 0) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < w.njoints) {
        const JointState& j = w.joints[t];
        if (joint_edge(w, j) && !(w.bflags[j.a] & BF_ASLEEP)) uf_union(w.uf2, j.a, j.b);
    }
    const uint32_t np = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    if (t >= np) return;
    const PairHot s = w.hot_cur[t];
    if (s.ccount <= 0 || (s.flags & (PF_DEAD | PF_VOID))) return;
    const uint32_t fa = w.bflags[s.a], fb = w.bflags[s.b];
    if ((fa & BF_DYNAMIC) && (fb & BF_DYNAMIC) && !(fa & BF_ASLEEP)) uf_union(w.uf2, s.a, s.b);
}

__global__ void __launch_bounds__(256) k_live_roots(DeviceWorld w) {
    if (w.counters[ctr(CTR_DEAD)] == 0) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.nbodies && (w.bflags[i] & BF_DYNAMIC)) w.comp_root[i] = uf_find(w.uf2, i);
}

// SimulationIslandMember.Activate on a dynamic body: the candidate flag is
// cleared, which zeroes the time below the limit.
__device__ __forceinline__ void activate_dynamic(DeviceWorld& w, uint32_t x, uint32_t extra) {
    atomicAnd(&w.bflags[x], ~(uint32_t)BF_CANDIDATE);
    if (extra) atomicOr(&w.bflags[x], extra);
    w.stab[x].y = 0.0f;
}

// Split attempts (DeactivationManager.FlushSplits / TryToSplit,
// DeactivationManager.cs:232-283, 432-555), all flushed at once: a
// constraint of the last step that is gone (no pair, no contacts, or removed
// as stale) between two dynamic bodies now in different islands activates
// both.
__global__ void __launch_bounds__(256) k_split_wake(DeviceWorld w) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t np = min(w.counters[ctr(CTR_PAIRS_PREV)], w.pair_cap);
    if (t >= np) return;
    const PairHot s = w.hot_old[t];
    if (s.ccount <= 0 || (s.flags & (PF_DEAD | PF_VOID))) return;
    if (!carried(w, s.a) || !carried(w, s.b)) return;
    if (!(w.bflags[s.a] & BF_DYNAMIC) || !(w.bflags[s.b] & BF_DYNAMIC)) return;
    if (w.comp_root[s.a] == w.comp_root[s.b]) return;  // still connected
    const uint64_t key = s.a < s.b ? ((uint64_t)s.a << 32 | s.b) : ((uint64_t)s.b << 32 | s.a);
    const int j = hash_find(w.hash_keys_cur, w.hash_vals_cur, w.hash_cap, key);
    if (j >= 0) {
        const PairHot q = w.hot_cur[j];
        if (q.ccount > 0 && !(q.flags & (PF_DEAD | PF_VOID))) return;
    }
    activate_dynamic(w, s.a, 0u);
    activate_dynamic(w, s.b, 0u);
}

// Space.Remove -> RemoveSimulationIslandFromMember (DeactivationManager.cs:
// 565-656): the dynamic bodies the removed body had constraints with are
// activated.  Runs between steps over the last step's pairs.
__global__ void __launch_bounds__(256) k_removal_wake(DeviceWorld w, uint32_t id) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t np = min(w.counters[ctr(CTR_PAIRS)], w.pair_cap);
    if (t >= np) return;
    const PairHot s = w.hot_old[t];
    if (s.ccount <= 0 || (s.flags & (PF_DEAD | PF_VOID)) || (s.a != id && s.b != id)) return;
    const uint32_t o = s.a == id ? s.b : s.a;
    if (w.bflags[o] & BF_DYNAMIC) activate_dynamic(w, o, BF_WAKE);
}

size_t deactivation_scan_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n);
    return bytes;
}

void launch_deactivation(const DeviceWorld& w, void* scan_temp, size_t scan_temp_bytes, hipStream_t st) {
    if (!w.sleeping || w.nbodies == 0) return;
    const uint32_t nb = (w.nbodies + 255) / 256;
    const uint32_t nwake = (std::max<uint32_t>(std::max<uint32_t>(w.nbodies, w.pair_cap), w.njoints) + 255) / 256;
    hipLaunchKernelGGL(k_deact_wake, dim3(nwake), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_deact_members, dim3(nb), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_deact_turn_m, dim3((w.nbodies + 1 + 255) / 256), dim3(256), 0, st, w);
    size_t bytes = scan_temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(scan_temp, bytes, w.deact_m, w.deact_e, (int)(w.nbodies + 1), st);
    hipLaunchKernelGGL(k_deact_turn, dim3(nb), dim3(256), 0, st, w);
}

void launch_activity(const DeviceWorld& w, hipStream_t st) {
    if (!w.sleeping || w.nbodies == 0) return;
    const uint32_t nb = (w.nbodies + 255) / 256;
    const uint32_t np = (w.pair_cap + 255) / 256;
    const uint32_t npj = (std::max<uint32_t>(w.pair_cap, w.njoints) + 255) / 256;
    const uint32_t nboth = (std::max<uint32_t>(w.nbodies, w.pair_cap) + 255) / 256;
    const uint32_t nmark =
        (std::max<uint32_t>(std::max<uint32_t>(w.nbodies, w.pair_cap), std::max<uint32_t>(w.nworlds, w.njoints)) + 255) / 256;
    hipLaunchKernelGGL(k_wake_mark, dim3(nmark), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_wake_apply, dim3(nboth), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_live_union, dim3(npj), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_live_roots, dim3(nb), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_split_wake, dim3(np), dim3(256), 0, st, w);
}

void launch_removal_wake(const DeviceWorld& w, uint32_t id, hipStream_t st) {
    if (!w.sleeping || w.pair_cap == 0) return;
    hipLaunchKernelGGL(k_removal_wake, dim3((w.pair_cap + 255) / 256), dim3(256), 0, st, w, id);
}

}  // namespace bepu
