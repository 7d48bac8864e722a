// k_solver.hip -- graph-coloured sequential-impulse contact solver.
//
// Stages (all launched without host synchronisation):
//   1. body -> constraint incidence lists (dynamic bodies only);
//   2. colouring (cooperative kernel, rounds separated by grid barriers):
//      a constraint takes the smallest colour unused by its conflicting
//      neighbours of higher priority mix64(key); this is the sequential greedy
//      colouring in decreasing priority, so it does not depend on scheduling;
//   3. constraints bucketed by colour (solve order = colour, any order within);
//   4. prestep: Jacobians, effective masses, bias (one thread per constraint);
//   5. warm start + iterations (cooperative kernel, one grid barrier per colour);
//      no two constraints of one colour share a dynamic body, so each colour
//      is a coalesced, race-free kernel phase;
//   6. accumulated impulses written back to the persistent pair state.
//
// Rows follow ContactPenetrationConstraint (ContactPenetrationConstraint.cs:114-321),
// SlidingFrictionTwoAxis (SlidingFrictionTwoAxis.cs:155-482), TwistFrictionConstraint
// (TwistFrictionConstraint.cs:65-204) and ContactFrictionConstraint
// (ContactFrictionConstraint.cs:131-287); the per-row early-out follows
// SolverGroup.SolveUpdateable (SolverGroup.cs:117-150) and Solver.UnsafeSolveIteration
// (Solver.cs:285-314).
#include <hip/hip_cooperative_groups.h>
#include <hipcub/hipcub.hpp>

#include "world.h"

namespace cg = cooperative_groups;

namespace bepu {

constexpr int kSolveBlock = 256;

// Activity word: bit k set = sub-constraint k active in solver (k: 0..3
// penetration, 4 sliding, 5 twist, 6..9 non-convex friction); bit 31 = group
// active.  A second word holds a 3-bit zero-impulse counter per sub-constraint
// (SolverSettings.iterationsAtZeroImpulse), enough for minimum iteration counts <= 6.
constexpr uint32_t kGroupBit = 1u << 31;
__device__ __forceinline__ int zero_count(uint32_t z, int k) { return (z >> (3 * k)) & 7; }
__device__ __forceinline__ uint32_t set_zero(uint32_t z, int k, int v) {
    return (z & ~(7u << (3 * k))) | ((uint32_t)v << (3 * k));
}

__device__ __forceinline__ bool convex_kind(int kind) { return kind == BEPU_PAIR_BOX_BOX || kind == BEPU_PAIR_CONVEX; }

// ---------------------------------------------------------------------------
// 1. incidence

__global__ void k_inc_count(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (c >= m) return;
    const PairState& s = w.pairs_cur[w.cons[c]];
    if (w.bflags[s.a] & 1u) atomicAdd(&w.inc_count[s.a], 1u);
    if (w.bflags[s.b] & 1u) atomicAdd(&w.inc_count[s.b], 1u);
    w.prio[c] = mix64(s.key);
    w.color[c] = -1;
}

// Fills the lists counting down; leaves inc_count at zero for the next step.
__global__ void k_inc_fill(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (c >= m) return;
    const PairState& s = w.pairs_cur[w.cons[c]];
    if (w.bflags[s.a] & 1u) w.inc_list[w.inc_offset[s.a] + atomicSub(&w.inc_count[s.a], 1u) - 1] = c;
    if (w.bflags[s.b] & 1u) w.inc_list[w.inc_offset[s.b] + atomicSub(&w.inc_count[s.b], 1u) - 1] = c;
}

// ---------------------------------------------------------------------------
// 2. colouring

__device__ __forceinline__ int load_color(const int* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }

// Returns the colour, or -1 when a higher-priority neighbour is not coloured yet.
__device__ int try_color(const DeviceWorld& w, uint32_t c) {
    const PairState& s = w.pairs_cur[w.cons[c]];
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
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    uint32_t gsize = gridDim.x * blockDim.x;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    // Each round colours at least the globally highest-priority uncoloured
    // constraint, so m + 1 rounds always suffice.
    for (uint32_t round = 0; round <= m; ++round) {
        uint32_t* remaining = &w.counters[CTR_UNCOLORED0 + (round % 3)];
        if (tid == 0) w.counters[CTR_UNCOLORED0 + ((round + 1) % 3)] = 0;
        uint32_t left = 0;
        for (uint32_t c = tid; c < m; c += gsize) {
            if (w.color[c] >= 0) continue;
            int col = try_color(w, c);
            if (col == -2) {
                w.counters[CTR_OVERFLOW] = 4;
                col = kMaxColors - 1;
            }
            if (col >= 0) {
                __atomic_store_n(&w.color[c], col, __ATOMIC_RELAXED);
                atomicMax(&w.counters[CTR_COLORS], (uint32_t)(col + 1));
            } else {
                ++left;
            }
        }
        if (left) atomicAdd(remaining, left);
        grid.sync();
        if (__atomic_load_n(remaining, __ATOMIC_RELAXED) == 0) {
            if (tid == 0) w.counters[CTR_ROUND] = round + 1;
            break;
        }
    }
}

// ---------------------------------------------------------------------------
// 3. bucket by colour

__global__ void k_color_hist(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (c >= m) return;
    atomicAdd(&w.color_count[w.color[c]], 1u);
}

__global__ void k_color_scan(DeviceWorld w) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t acc = 0;
    for (int k = 0; k < kMaxColors; ++k) {
        w.color_start[k] = acc;
        w.color_cursor[k] = 0;
        acc += w.color_count[k];
        w.color_count[k] = 0;
    }
    w.color_start[kMaxColors] = acc;
}

__global__ void k_color_scatter(DeviceWorld w) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (c >= m) return;
    int col = w.color[c];
    w.ordered[w.color_start[col] + atomicAdd(&w.color_cursor[col], 1u)] = c;
}

// ---------------------------------------------------------------------------
// Body access

struct BodyV {
    f3 v, w;
    m3 I;
    float im;
    bool dyn;
};

__device__ __forceinline__ BodyV load_body(const DeviceWorld& w, uint32_t b) {
    BodyV r;
    r.v = f3_of(w.linv[b]);
    r.w = f3_of(w.angv[b]);
    float4 a = w.ii0[b], c = w.ii1[b], d = w.ii2[b];
    r.I = m3{a.x, a.y, a.z, c.x, c.y, c.z, d.x, d.y, d.z};
    r.im = a.w;
    r.dyn = (w.bflags[b] & 1u) != 0;
    return r;
}

// Entity.ApplyLinearImpulse / ApplyAngularImpulse (EntityBase.cs:754-788)
__device__ __forceinline__ void apply_lin(BodyV& b, f3 l) {
    b.v.x += l.x * b.im;
    b.v.y += l.y * b.im;
    b.v.z += l.z * b.im;
}
__device__ __forceinline__ void apply_ang(BodyV& b, f3 i) {
    b.w.x += i.x * b.I.m11 + i.y * b.I.m21 + i.z * b.I.m31;
    b.w.y += i.x * b.I.m12 + i.y * b.I.m22 + i.z * b.I.m32;
    b.w.z += i.x * b.I.m13 + i.y * b.I.m23 + i.z * b.I.m33;
}

// tX..entry of the effective-mass denominators (ContactPenetrationConstraint.cs:149-170)
__device__ __forceinline__ float eff_entry(f3 j, const BodyV& b) {
    float tX = j.x * b.I.m11 + j.y * b.I.m21 + j.z * b.I.m31;
    float tY = j.x * b.I.m12 + j.y * b.I.m22 + j.z * b.I.m32;
    float tZ = j.x * b.I.m13 + j.y * b.I.m23 + j.z * b.I.m33;
    return tX * j.x + tY * j.y + tZ * j.z + b.im;
}

// ---------------------------------------------------------------------------
// Row storage helpers

struct RowsView {
    const DeviceWorld& w;
    uint32_t cap;
    __device__ float4& pen(int k, int f, uint32_t p) const { return w.rows.pen[((size_t)k * 3 + f) * cap + p]; }
    __device__ float& penacc(int k, uint32_t p) const { return w.rows.penacc[(size_t)k * cap + p]; }
    __device__ float4& sl(int f, uint32_t p) const { return w.rows.sl[(size_t)f * cap + p]; }
    __device__ float4& tw(int f, uint32_t p) const { return w.rows.tw[(size_t)f * cap + p]; }
    __device__ float4& fr(int k, int f, uint32_t p) const { return w.rows.fr[((size_t)k * 3 + f) * cap + p]; }
    __device__ float& fracc(int k, uint32_t p) const { return w.rows.fracc[(size_t)k * cap + p]; }
};

// ContactPenetrationConstraint.RelativeVelocity (ContactPenetrationConstraint.cs:86-104)
__device__ __forceinline__ float pen_relvel(f3 lin, f3 aA, f3 aB, const BodyV& A, const BodyV& B) {
    float l = A.v.x * lin.x + A.v.y * lin.y + A.v.z * lin.z + A.w.x * aA.x + A.w.y * aA.y + A.w.z * aA.z;
    l += -B.v.x * lin.x - B.v.y * lin.y - B.v.z * lin.z + B.w.x * aB.x + B.w.y * aB.y + B.w.z * aB.z;
    return l;
}

// ---------------------------------------------------------------------------
// 4. prestep

__global__ void __launch_bounds__(kSolveBlock) k_prestep(DeviceWorld w) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    if (p >= m) return;
    const bepu_settings& st = w.s;
    RowsView R{w, w.pair_cap};
    uint32_t c = w.ordered[p];
    uint32_t slot = w.cons[c];
    PairState& s = w.pairs_cur[slot];
    BodyV A = load_body(w, s.a), B = load_body(w, s.b);
    f3 pA = f3_of(w.pos[s.a]), pB = f3_of(w.pos[s.b]);
    int npen = s.npen;
    bool convex = convex_kind(s.kind);
    w.rows.bodies[p] = make_int2((int)s.a, (int)s.b);
    w.rows.meta[p] = s.kind | (npen << 4) | ((int)A.dyn << 8) | ((int)B.dyn << 9);
    w.rows.pair[p] = (int)slot;
    uint32_t act = 0;
    bool group = A.dyn || B.dyn;
    if (group) act |= kGroupBit;
    float updateRate = 1 / st.dt;
    // Penetration rows (ContactPenetrationConstraint.Update, :114-233)
    for (int k = 0; k < npen; ++k) {
        int cs = s.pslot[k];
        float4 cp = s.cpos[cs], cn = s.cnrm[cs];
        f3 lin{-cn.x, -cn.y, -cn.z};
        f3 ra = f3_of(cp) - pA;
        f3 aA{(ra.y * lin.z) - (ra.z * lin.y), (ra.z * lin.x) - (ra.x * lin.z), (ra.x * lin.y) - (ra.y * lin.x)};
        f3 rb = f3_of(cp) - pB;
        f3 aB{(lin.y * rb.z) - (lin.z * rb.y), (lin.z * rb.x) - (lin.x * rb.z), (lin.x * rb.y) - (lin.y * rb.x)};
        float eA = A.dyn ? eff_entry(aA, A) : 0.0f;
        float eB = B.dyn ? eff_entry(aB, B) : 0.0f;
        float emi = eA + eB;
        float soft = st.softness * emi * updateRate;
        float v2i = -1 / (soft + emi);
        float bias;
        float depth = cp.w;
        if (depth >= 0) {
            bias = minf_ref(maxf_ref(0, depth - st.allowed_penetration) * st.penetration_recovery_stiffness * updateRate,
                            st.max_penetration_recovery_speed);
            if (s.bounce > 0) {
                float bv = -pen_relvel(lin, aA, aB, A, B);
                if (bv > 0) {
                    float low = st.bounce_velocity_threshold * 0.3f;
                    float frac = clampf_ref((bv - low) / (st.bounce_velocity_threshold - low + 1e-7f), 0, 1);
                    bias = maxf_ref(frac * bv * s.bounce, bias);
                }
            }
        } else {
            bias = depth * updateRate;
        }
        R.pen(k, 0, p) = f4_of(lin, soft);
        R.pen(k, 1, p) = f4_of(aA, bias);
        R.pen(k, 2, p) = f4_of(aB, v2i);
        R.penacc(k, p) = s.penacc[k];
        if (group) act |= 1u << k;
        if (!convex) {
            // ContactFrictionConstraint.Update (ContactFrictionConstraint.cs:180-287)
            f3 vA = cross3(A.w, ra) + A.v;
            f3 vB = cross3(B.w, rb) + B.v;
            f3 rv = vA - vB;
            f3 n = f3_of(cn);
            float nvs = n.x * rv.x + n.y * rv.y + n.z * rv.z;
            rv.x -= nvs * n.x;
            rv.y -= nvs * n.y;
            rv.z -= nvs * n.z;
            float4 fl = s.frlin[k];
            f3 flin = f3_of(fl);
            float fric = 0;
            bool factive = group;
            float len = lensq3(rv);
            if (len > 1e-7f) {
                len = sqrtf(len);
                flin = f3{rv.x / len, rv.y / len, rv.z / len};
                fric = len > st.static_friction_velocity_threshold ? s.kinetic : s.static_f;
            } else if (flin.x != 0 || flin.y != 0 || flin.z != 0) {
                fric = s.static_f;
            } else {
                factive = false;
            }
            s.frlin[k] = f4_of(flin, 0.0f);
            if (factive) {
                f3 fa{(ra.y * flin.z) - (ra.z * flin.y), (ra.z * flin.x) - (ra.x * flin.z), (ra.x * flin.y) - (ra.y * flin.x)};
                f3 fb{(flin.y * rb.z) - (flin.z * rb.y), (flin.z * rb.x) - (flin.x * rb.z), (flin.x * rb.y) - (flin.y * rb.x)};
                float fA = A.dyn ? eff_entry(fa, A) : 0.0f;
                float fB = B.dyn ? eff_entry(fb, B) : 0.0f;
                R.fr(k, 0, p) = f4_of(flin, fric);
                R.fr(k, 1, p) = f4_of(fa, -1 / (fA + fB));
                R.fr(k, 2, p) = f4_of(fb, 0.0f);
                act |= 1u << (6 + k);
            }
            R.fracc(k, p) = s.fracc[k];
        }
    }
    w.rows.slacc[p] = make_float2(s.slacc[0], s.slacc[1]);
    w.rows.twacc[p] = s.twacc;
    if (convex && group && (s.flags & PF_FRICTION_ACTIVE)) {
        // SlidingFrictionTwoAxis.Update (SlidingFrictionTwoAxis.cs:242-440)
        f3 center;
        f3 c0 = f3_of(s.cpos[s.pslot[0]]);
        if (npen == 1) center = c0;
        else if (npen == 2) center = (c0 + f3_of(s.cpos[s.pslot[1]])) * .5f;
        else if (npen == 3) center = (f3_of(s.cpos[s.pslot[2]]) + (c0 + f3_of(s.cpos[s.pslot[1]]))) * .333333333f;
        else center = (f3_of(s.cpos[s.pslot[3]]) + (f3_of(s.cpos[s.pslot[2]]) + (c0 + f3_of(s.cpos[s.pslot[1]])))) * .25f;
        f3 ra = center - pA;
        f3 vA = cross3(A.w, ra) + A.v;
        f3 rb = center - pB;
        f3 vB = cross3(B.w, rb) + B.v;
        f3 rv = vA - vB;
        f3 n = f3_of(s.cnrm[s.pslot[0]]);
        float nvs = n.x * rv.x + n.y * rv.y + n.z * rv.z;
        rv.x -= nvs * n.x;
        rv.y -= nvs * n.y;
        rv.z -= nvs * n.z;
        float L[6];
        for (int q = 0; q < 6; ++q) L[q] = s.sllin[q];
        float fric;
        float len = lensq3(rv);
        if (len > 1e-7f) {
            len = sqrtf(len);
            float inv = 1 / len;
            L[0] = rv.x * inv; L[1] = rv.y * inv; L[2] = rv.z * inv;
            fric = len > st.static_friction_velocity_threshold ? s.kinetic : s.static_f;
        } else {
            fric = s.static_f;
            if (!(L[0] != 0 || L[1] != 0 || L[2] != 0)) {
                f3 ax = cross3(n, f3{1, 0, 0});
                len = lensq3(ax);
                if (len > 1e-7f) {
                    len = sqrtf(len);
                    float inv = 1 / len;
                    L[0] = ax.x * inv; L[1] = ax.y * inv; L[2] = ax.z * inv;
                } else {
                    ax = normalize3(cross3(n, f3{0, 1, 0}));
                    L[0] = ax.x; L[1] = ax.y; L[2] = ax.z;
                }
            }
        }
        L[3] = (L[1] * n.z) - (L[2] * n.y);
        L[4] = (L[2] * n.x) - (L[0] * n.z);
        L[5] = (L[0] * n.y) - (L[1] * n.x);
        for (int q = 0; q < 6; ++q) s.sllin[q] = L[q];
        f3 a0{(ra.y * L[2]) - (ra.z * L[1]), (ra.z * L[0]) - (ra.x * L[2]), (ra.x * L[1]) - (ra.y * L[0])};
        f3 a1{(ra.y * L[5]) - (ra.z * L[4]), (ra.z * L[3]) - (ra.x * L[5]), (ra.x * L[4]) - (ra.y * L[3])};
        f3 b0{(L[1] * rb.z) - (L[2] * rb.y), (L[2] * rb.x) - (L[0] * rb.z), (L[0] * rb.y) - (L[1] * rb.x)};
        f3 b1{(L[4] * rb.z) - (L[5] * rb.y), (L[5] * rb.x) - (L[3] * rb.z), (L[3] * rb.y) - (L[4] * rb.x)};
        // J I^-1 J^T + m^-1 per body (Matrix2x3.Multiply / Transpose, Matrix2x2.Multiply)
        auto entry = [](f3 j0, f3 j1, const BodyV& b, float e[4]) {
            f3 t0{j0.x * b.I.m11 + j0.y * b.I.m21 + j0.z * b.I.m31, j0.x * b.I.m12 + j0.y * b.I.m22 + j0.z * b.I.m32,
                  j0.x * b.I.m13 + j0.y * b.I.m23 + j0.z * b.I.m33};
            f3 t1{j1.x * b.I.m11 + j1.y * b.I.m21 + j1.z * b.I.m31, j1.x * b.I.m12 + j1.y * b.I.m22 + j1.z * b.I.m32,
                  j1.x * b.I.m13 + j1.y * b.I.m23 + j1.z * b.I.m33};
            e[0] = t0.x * j0.x + t0.y * j0.y + t0.z * j0.z;
            e[1] = t0.x * j1.x + t0.y * j1.y + t0.z * j1.z;
            e[2] = t1.x * j0.x + t1.y * j0.y + t1.z * j0.z;
            e[3] = t1.x * j1.x + t1.y * j1.y + t1.z * j1.z;
            e[0] += b.im;
            e[3] += b.im;
        };
        float eA[4] = {0, 0, 0, 0}, eB[4] = {0, 0, 0, 0};
        if (A.dyn) entry(a0, a1, A, eA);
        if (B.dyn) entry(b0, b1, B, eB);
        float m11 = -eA[0] - eB[0], m12 = -eA[1] - eB[1], m21 = -eA[2] - eB[2], m22 = -eA[3] - eB[3];
        float di = 1 / (m11 * m22 - m12 * m21);  // Matrix2x2.Invert (Matrix2x2.cs:160-174)
        R.sl(0, p) = make_float4(L[0], L[1], L[2], fric);
        R.sl(1, p) = make_float4(L[3], L[4], L[5], 0.0f);
        R.sl(2, p) = f4_of(a0, 0.0f);
        R.sl(3, p) = f4_of(a1, 0.0f);
        R.sl(4, p) = f4_of(b0, 0.0f);
        R.sl(5, p) = f4_of(b1, 0.0f);
        R.sl(6, p) = make_float4(m22 * di, -m12 * di, -m21 * di, m11 * di);
        R.sl(7, p) = f4_of(ra, 0.0f);
        R.sl(8, p) = f4_of(rb, 0.0f);
        act |= 1u << 4;
        // TwistFrictionConstraint.Update (TwistFrictionConstraint.cs:114-175)
        float tA = A.dyn ? eff_entry(n, A) : 0.0f;
        float tB = B.dyn ? eff_entry(n, B) : 0.0f;
        float tv2i = -1 / (tA + tB);
        float rav = A.w.x * n.x + A.w.y * n.y + A.w.z * n.z;
        rav -= B.w.x * n.x + B.w.y * n.y + B.w.z * n.z;
        float tf = fabsf(rav) > st.static_friction_velocity_threshold ||
                           fabsf(rv.x) + fabsf(rv.y) + fabsf(rv.z) > st.static_friction_velocity_threshold
                       ? s.kinetic
                       : s.static_f;
        tf *= st.twist_friction_factor;
        float lever[4] = {0, 0, 0, 0};
        for (int k = 0; k < npen; ++k) lever[k] = len3(f3_of(s.cpos[s.pslot[k]]) - center);
        R.tw(0, p) = f4_of(n, tv2i);
        R.tw(1, p) = make_float4(lever[0], lever[1], lever[2], lever[3]);
        R.tw(2, p) = make_float4(tf, 0.0f, 0.0f, 0.0f);
        act |= 1u << 5;
    }
    w.rows.act[p] = act;
    w.rows.zero[p] = 0;
}

// ---------------------------------------------------------------------------
// 5. warm start + iterations

__device__ __forceinline__ void pen_apply(f3 lin, f3 aA, f3 aB, float lambda, BodyV& A, BodyV& B) {
    f3 l = lin * lambda;
    if (A.dyn) {
        apply_lin(A, l);
        apply_ang(A, aA * lambda);
    }
    if (B.dyn) {
        apply_lin(B, -l);
        apply_ang(B, aB * lambda);
    }
}

__device__ __forceinline__ void sl_apply(const float4* S, float lx, float ly, BodyV& A, BodyV& B) {
    f3 lin{lx * S[0].x + ly * S[1].x, lx * S[0].y + ly * S[1].y, lx * S[0].z + ly * S[1].z};
    if (A.dyn) {
        apply_lin(A, lin);
        apply_ang(A, f3{lx * S[2].x + ly * S[3].x, lx * S[2].y + ly * S[3].y, lx * S[2].z + ly * S[3].z});
    }
    if (B.dyn) {
        apply_lin(B, -lin);
        apply_ang(B, f3{lx * S[4].x + ly * S[5].x, lx * S[4].y + ly * S[5].y, lx * S[4].z + ly * S[5].z});
    }
}

__device__ __forceinline__ void tw_apply(f3 ang, float lambda, BodyV& A, BodyV& B) {
    f3 a = ang * lambda;
    if (A.dyn) apply_ang(A, a);
    if (B.dyn) apply_ang(B, -a);
}

__device__ __forceinline__ void store_body(const DeviceWorld& w, uint32_t b, const BodyV& v) {
    w.linv[b] = f4_of(v.v, 0.0f);
    w.angv[b] = f4_of(v.w, 0.0f);
}

// Solver.UnsafeExclusiveUpdate (Solver.cs:277-283)
__device__ void warm_one(const DeviceWorld& w, uint32_t p) {
    uint32_t act = w.rows.act[p];
    if (!(act & kGroupBit)) return;
    RowsView R{w, w.pair_cap};
    int2 bd = w.rows.bodies[p];
    int meta = w.rows.meta[p];
    int npen = (meta >> 4) & 15;
    BodyV A = load_body(w, bd.x), B = load_body(w, bd.y);
    for (int k = 0; k < npen; ++k)
        if (act & (1u << k)) {
            float4 r0 = R.pen(k, 0, p), r1 = R.pen(k, 1, p), r2 = R.pen(k, 2, p);
            float acc = R.penacc(k, p);
            pen_apply(f3_of(r0), f3_of(r1), f3_of(r2), acc, A, B);
        }
    if (convex_kind(meta & 15)) {
        if (act & (1u << 4)) {
            float4 S[6];
            for (int q = 0; q < 6; ++q) S[q] = R.sl(q, p);
            float2 acc = w.rows.slacc[p];
            sl_apply(S, acc.x, acc.y, A, B);
        }
        if (act & (1u << 5)) tw_apply(f3_of(R.tw(0, p)), w.rows.twacc[p], A, B);
    } else {
        for (int k = 0; k < npen; ++k)
            if (act & (1u << (6 + k))) {
                float4 f0 = R.fr(k, 0, p), f1 = R.fr(k, 1, p), f2 = R.fr(k, 2, p);
                pen_apply(f3_of(f0), f3_of(f1), f3_of(f2), R.fracc(k, p), A, B);
            }
    }
    if (A.dyn) store_body(w, bd.x, A);
    if (B.dyn) store_body(w, bd.y, B);
}

// SolverGroup.SolveUpdateable bookkeeping for sub-constraint k after solving.
__device__ __forceinline__ void account(uint32_t& act, uint32_t& zero, int k, float impulse, float min_impulse,
                                        int min_iter, int& active) {
    if (impulse < min_impulse) {
        int z = zero_count(zero, k) + 1;
        if (z > min_iter) act &= ~(1u << k);
        else active++;
        zero = set_zero(zero, k, z > 7 ? 7 : z);
    } else {
        zero = set_zero(zero, k, 0);
        active++;
    }
}

// Solver.UnsafeSolveIteration -> ContactManifoldConstraint.SolveIteration
__device__ void solve_one(const DeviceWorld& w, uint32_t p) {
    uint32_t act = w.rows.act[p];
    if (!(act & kGroupBit)) return;
    uint32_t zero = w.rows.zero[p];
    RowsView R{w, w.pair_cap};
    const float mi = w.s.min_impulse;
    const int mic = w.s.min_iteration_count;
    int2 bd = w.rows.bodies[p];
    int meta = w.rows.meta[p];
    int npen = (meta >> 4) & 15;
    BodyV A = load_body(w, bd.x), B = load_body(w, bd.y);
    int active = 0;
    float pacc[4] = {0, 0, 0, 0};
    for (int k = 0; k < npen; ++k) {
        pacc[k] = R.penacc(k, p);
        if (!(act & (1u << k))) continue;
        float4 r0 = R.pen(k, 0, p), r1 = R.pen(k, 1, p), r2 = R.pen(k, 2, p);
        f3 lin = f3_of(r0), aA = f3_of(r1), aB = f3_of(r2);
        float lambda = (pen_relvel(lin, aA, aB, A, B) - r1.w + r0.w * pacc[k]) * r2.w;
        float prev = pacc[k];
        pacc[k] = maxf_ref(0, pacc[k] + lambda);
        lambda = pacc[k] - prev;
        pen_apply(lin, aA, aB, lambda, A, B);
        R.penacc(k, p) = pacc[k];
        account(act, zero, k, fabsf(lambda), mi, mic, active);
    }
    if (convex_kind(meta & 15)) {
        if (act & (1u << 4)) {
            float4 S[9];
            for (int q = 0; q < 9; ++q) S[q] = R.sl(q, p);
            f3 ra = f3_of(S[7]), rb = f3_of(S[8]);
            float dvx = A.v.x + (A.w.y * ra.z) - (A.w.z * ra.y);
            float dvy = A.v.y + (A.w.z * ra.x) - (A.w.x * ra.z);
            float dvz = A.v.z + (A.w.x * ra.y) - (A.w.y * ra.x);
            dvx += -B.v.x - (B.w.y * rb.z) + (B.w.z * rb.y);
            dvy += -B.v.y - (B.w.z * rb.x) + (B.w.x * rb.z);
            dvz += -B.v.z - (B.w.x * rb.y) + (B.w.y * rb.x);
            float lx = dvx * S[0].x + dvy * S[0].y + dvz * S[0].z;
            float ly = dvx * S[1].x + dvy * S[1].y + dvz * S[1].z;
            float x = lx;
            lx = x * S[6].x + ly * S[6].z;
            ly = x * S[6].y + ly * S[6].w;
            float2 acc = w.rows.slacc[p];
            float px = acc.x, py = acc.y;
            acc.x += lx;
            acc.y += ly;
            float length = acc.x * acc.x + acc.y * acc.y;
            float maxF = 0;
            for (int k = 0; k < npen; ++k) maxF += pacc[k];
            maxF *= S[0].w;
            if (length > maxF * maxF) {
                length = maxF / sqrtf(length);
                acc.x *= length;
                acc.y *= length;
            }
            lx = acc.x - px;
            ly = acc.y - py;
            sl_apply(S, lx, ly, A, B);
            w.rows.slacc[p] = acc;
            account(act, zero, 4, fabsf(lx) + fabsf(ly), mi, mic, active);
        }
        if (act & (1u << 5)) {
            float4 t0 = R.tw(0, p), t1 = R.tw(1, p);
            float fr = R.tw(2, p).x;
            f3 ang = f3_of(t0);
            float lambda = A.w.x * ang.x + A.w.y * ang.y + A.w.z * ang.z;
            lambda -= B.w.x * ang.x + B.w.y * ang.y + B.w.z * ang.z;
            lambda *= t0.w;
            float acc = w.rows.twacc[p];
            float prev = acc;
            float lev[4] = {t1.x, t1.y, t1.z, t1.w};
            float maxF = 0;
            for (int k = 0; k < npen; ++k) maxF += lev[k] * pacc[k];
            maxF *= fr;
            acc = clampf_ref(acc + lambda, -maxF, maxF);
            lambda = acc - prev;
            tw_apply(ang, lambda, A, B);
            w.rows.twacc[p] = acc;
            account(act, zero, 5, fabsf(lambda), mi, mic, active);
        }
    } else {
        for (int k = 0; k < npen; ++k) {
            if (!(act & (1u << (6 + k)))) continue;
            float4 f0 = R.fr(k, 0, p), f1 = R.fr(k, 1, p), f2 = R.fr(k, 2, p);
            f3 lin = f3_of(f0), aA = f3_of(f1), aB = f3_of(f2);
            float vel = 0;
            vel += A.v.x * lin.x + A.v.y * lin.y + A.v.z * lin.z + A.w.x * aA.x + A.w.y * aA.y + A.w.z * aA.z;
            vel += -B.v.x * lin.x - B.v.y * lin.y - B.v.z * lin.z + B.w.x * aB.x + B.w.y * aB.y + B.w.z * aB.z;
            float lambda = vel * f1.w;
            float acc = R.fracc(k, p);
            float prev = acc;
            float maxForce = f0.w * pacc[k];
            acc = clampf_ref(acc + lambda, -maxForce, maxForce);
            lambda = acc - prev;
            pen_apply(lin, aA, aB, lambda, A, B);
            R.fracc(k, p) = acc;
            account(act, zero, 6 + k, fabsf(lambda), mi, mic, active);
        }
    }
    if (active == 0) act &= ~kGroupBit;
    w.rows.act[p] = act;
    w.rows.zero[p] = zero;
    if (A.dyn) store_body(w, bd.x, A);
    if (B.dyn) store_body(w, bd.y, B);
}

__global__ void __launch_bounds__(kSolveBlock) k_solve(DeviceWorld w) {
    cg::grid_group grid = cg::this_grid();
    uint32_t gsize = gridDim.x * blockDim.x;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    int ncol = (int)min(w.counters[CTR_COLORS], (uint32_t)kMaxColors);
    for (int col = 0; col < ncol; ++col) {
        uint32_t beg = w.color_start[col], end = w.color_start[col + 1];
        for (uint32_t p = beg + tid; p < end; p += gsize) warm_one(w, p);
        grid.sync();
    }
    int iters = w.s.iteration_limit;
    for (int it = 0; it < iters; ++it)
        for (int col = 0; col < ncol; ++col) {
            uint32_t beg = w.color_start[col], end = w.color_start[col + 1];
            for (uint32_t p = beg + tid; p < end; p += gsize) solve_one(w, p);
            grid.sync();
        }
    // 6. write the accumulated impulses back to the persistent pair state
    uint32_t m = min(w.counters[CTR_CONS], w.pair_cap);
    RowsView R{w, w.pair_cap};
    for (uint32_t p = tid; p < m; p += gsize) {
        PairState& s = w.pairs_cur[w.rows.pair[p]];
        int npen = s.npen;
        for (int k = 0; k < npen; ++k) s.penacc[k] = R.penacc(k, p);
        if (convex_kind(s.kind)) {
            if (s.flags & PF_FRICTION_ACTIVE) {
                float2 a = w.rows.slacc[p];
                s.slacc[0] = a.x;
                s.slacc[1] = a.y;
                s.twacc = w.rows.twacc[p];
            }
        } else {
            for (int k = 0; k < npen; ++k) s.fracc[k] = R.fracc(k, p);
        }
    }
}

// ---------------------------------------------------------------------------

static int cooperative_grid(const void* kernel, int block) {
    static int cached[2] = {0, 0};
    int idx = kernel == (const void*)k_color ? 0 : 1;
    if (cached[idx]) return cached[idx];
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, dev);
    int per_cu = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
    if (per_cu < 1) per_cu = 1;
    if (per_cu > 2) per_cu = 2;
    cached[idx] = prop.multiProcessorCount * per_cu;
    return cached[idx];
}

size_t solver_scan_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n);
    return bytes;
}

void launch_solver_full(const DeviceWorld& w, void* temp, size_t temp_bytes, hipStream_t st) {
    uint32_t cap = w.pair_cap;
    dim3 g((cap + kSolveBlock - 1) / kSolveBlock), b(kSolveBlock);
    hipLaunchKernelGGL(k_inc_count, g, b, 0, st, w);
    size_t bytes = temp_bytes;
    hipcub::DeviceScan::ExclusiveSum(temp, bytes, w.inc_count, w.inc_offset, (int)(w.nbodies + 1), st);
    hipLaunchKernelGGL(k_inc_fill, g, b, 0, st, w);
    DeviceWorld wc = w;
    void* args[] = {&wc};
    hipLaunchCooperativeKernel((const void*)k_color, dim3(cooperative_grid((const void*)k_color, kSolveBlock)),
                               dim3(kSolveBlock), args, 0, st);
    hipLaunchKernelGGL(k_color_hist, g, b, 0, st, w);
    hipLaunchKernelGGL(k_color_scan, dim3(1), dim3(64), 0, st, w);
    hipLaunchKernelGGL(k_color_scatter, g, b, 0, st, w);
    hipLaunchKernelGGL(k_prestep, g, b, 0, st, w);
    hipLaunchCooperativeKernel((const void*)k_solve, dim3(cooperative_grid((const void*)k_solve, kSolveBlock)),
                               dim3(kSolveBlock), args, 0, st);
}

}  // namespace bepu
