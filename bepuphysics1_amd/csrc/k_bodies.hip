// k_bodies.hip -- per-body stages of the step: deactivation candidacy,
// force update, bounding box (one fused pass), and position integration.
//
// HBM-bound elementwise kernels: one thread per body, float4 SoA loads.
#include "body_shape.h"

namespace bepu {

// Half-octave bin of a non-negative length: bin 2 * (e + 16) + h for
// lengths in [2^e (1 + h/2), 2^e (1.5 + h/2)), clamped to [0, 63].
__device__ __forceinline__ int side_bin(float side) {
    if (!(side > 0.0f)) return 0;
    uint32_t u = __float_as_uint(side);
    int e = (int)((u >> 23) & 0xFFu) - 127;
    int b = 2 * (e + 16) + (int)((u >> 22) & 1u);
    return min(max(b, 0), 63);
}

// Upper end of a half-octave bin.
__device__ __forceinline__ float side_bin_edge(int b) {
    return ldexpf((b & 1) ? 2.0f : 1.5f, b / 2 - 16);
}

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// One pass per body, in the order Space.DoTimeStep runs the stages
// (Space.cs:505-508): SimulationIslandMember.UpdateDeactivationCandidacy
// (SimulationIslandMember.cs:53-95), Entity.UpdateForForces
// (EntityBase.cs:911-977), EntityCollidable.UpdateBoundingBox
// (EntityCollidable.cs:140-148,195-257 with BoxShape.cs:149-174 /
// SphereShape.cs:75-86).  Also folds the AABB centres into the scene bounds
// used to quantise Morton codes.
__global__ void __launch_bounds__(256) k_body_pre(DeviceWorld w) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bepu_settings& s = w.s;
    float cx = 0, cy = 0, cz = 0, side = 0;
    bool valid = false;
    if (i < w.nbodies) {
        w.uf_parent[i] = i;  // union-find forest of this step's islands
        if (w.sleeping) w.act2[i] = 0u;
        valid = (w.bflags[i] & BF_PRESENT) != 0;
        if (valid && w.wfault[w.wid[i]]) {
            // Its world's state became non-finite: every body of that world
            // leaves the simulation, so no other world sees it.
            w.bflags[i] = 0u;
            valid = false;
        }
        if (!valid) {  // a removed world's slot: an empty box overlaps nothing
            w.bmin[i] = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
            w.bmax[i] = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
        }
    }
    bool asleep = false;
    if (valid && w.sleeping) {
        // The island decisions of this step's deactivation stage.
        const uint32_t f = w.bflags[i];
        if (f & BF_DYNAMIC) {
            const uint32_t fl = w.lab_flags[w.comp_root[i]];
            uint32_t nf = f;
            if (fl & LF_CHECKED) nf &= ~(uint32_t)BF_NODEACT;
            if (fl & LF_SLEEP) nf |= BF_ASLEEP;
            if (nf != f) w.bflags[i] = nf;
            asleep = (nf & BF_ASLEEP) != 0;
        }
    }
    if (valid && asleep) {
        // ForceUpdater and BoundingBoxUpdater skip inactive entities
        // (ForceUpdater.cs:75, BoundingBoxUpdater.cs:48): the box stays.
        const float4 mn = w.bmin[i], mx = w.bmax[i];
        cx = 0.5f * (mn.x + mx.x);
        cy = 0.5f * (mn.y + mx.y);
        cz = 0.5f * (mn.z + mx.z);
        side = fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z);
        if (!(side >= 0.0f && side < 3.0e38f)) side = 0.0f;
    } else if (valid) {
        bool dyn = (w.bflags[i] & BF_DYNAMIC) != 0;
        float4 P = w.pos[i];
        float4 Q = w.quat[i];
        float4 V = w.linv[i];
        float4 W = w.angv[i];
        f3 v = f3_of(V), av = f3_of(W);
        q4 q{Q.x, Q.y, Q.z, Q.w};
        if (!dyn && !w.sleeping) {
            // a kinematic member's candidacy (the deactivation stage does it
            // when it runs)
            const uint32_t bf = w.bflags[i];
            const float fl = w.stab[i].y;
            const float next = kinematic_flag_next(fl, lensq3(v) + lensq3(av), (bf & BF_ALWAYS_ACTIVE) != 0);
            if (next != fl) w.stab[i].y = next;
            const uint32_t nf = next <= 0.0f ? (bf | BF_KIN_ACTIVE) : (bf & ~(uint32_t)BF_KIN_ACTIVE);
            if (nf != bf) w.bflags[i] = nf;
        }
        if (dyn) {
            float4 st = w.stab[i];
            float lim = s.velocity_lower_limit;
            float velocity, slowing, tb;
            if (w.sleeping) {
                // candidacy ran in the deactivation stage
                velocity = st.x;
                tb = st.y;
                slowing = st.z;
            } else {
                // candidacy; clearing the candidate flag also zeroes the time
                // below the limit (the IsDeactivationCandidate setter), which an
                // always-active body does every step
                velocity = lensq3(v) + lensq3(av);
                slowing = velocity <= st.x ? 1.0f : 0.0f;
                tb = st.y;
                if (velocity < lim * lim) tb += s.dt;
                else tb = 0;
                const uint32_t bf = w.bflags[i];
                uint32_t cand = bf & BF_CANDIDATE;
                if (!(bf & BF_ALWAYS_ACTIVE)) {
                    if (!cand) {
                        if (tb > s.low_velocity_time_minimum && slowing != 0.0f) cand = BF_CANDIDATE;
                    } else if (tb <= s.low_velocity_time_minimum) {
                        cand = 0;
                        tb = 0;
                    }
                } else {
                    cand = 0;
                    tb = 0;
                }
                const uint32_t nf = ((bf & ~(uint32_t)(BF_CANDIDATE | BF_WAKE)) | cand);
                if (nf != bf) w.bflags[i] = nf;
            }
            // force update: gravity, stabilization boost, damping, world inertia
            f3 gdt = f3{s.gravity[0], s.gravity[1], s.gravity[2]} * s.dt;
            v = gdt + v;
            float2 dm = w.damp[i];
            float lin_boost = 0, ang_boost = 0;
            if (s.use_stabilization && st.w != 0.0f && (slowing != 0.0f || tb > s.low_velocity_time_minimum)) {
                float energy = lensq3(v) + lensq3(av);
                if (energy < lim * lim) {
                    float boost = 1 - (float)(sqrt((double)energy) / (double)(2.0f * lim));
                    float total = dm.y + ang_boost;
                    float rem = 1 - total;
                    ang_boost += boost * rem;
                    total = dm.x + lin_boost;
                    rem = 1 - total;
                    lin_boost += boost * rem;
                }
            }
            float linear = dm.x + lin_boost;
            if (linear > 0) v = v * (float)pow((double)clampf_ref(1 - linear, 0, 1), (double)s.dt);
            float angular = dm.y + ang_boost;
            if (angular > 0) av = av * (float)pow((double)clampf_ref(1 - angular, 0, 1), (double)s.dt);
            m3 R = mat_from_quat(q);
            float4 l0 = w.li0[i], l1 = w.li1[i], l2 = w.li2[i];
            m3 L{l0.x, l0.y, l0.z, l1.x, l1.y, l1.z, l2.x, l2.y, l2.z};
            m3 I = rotate_inertia(R, L);
            float im = w.ii0[i].w;
            w.ii0[i] = make_float4(I.m11, I.m12, I.m13, im);
            w.ii1[i] = make_float4(I.m21, I.m22, I.m23, 0.0f);
            w.ii2[i] = make_float4(I.m31, I.m32, I.m33, 0.0f);
            w.linv[i] = f4_of(v, 0.0f);
            w.angv[i] = f4_of(av, 0.0f);
            w.stab[i] = make_float4(velocity, tb, slowing, st.w);
        }
        // bounding box
        f3 mn, mx;
        shape_aabb(w, w.shape[i], f3_of(P), q, &mn, &mx);
        float dt = s.dt;
        if (dt > 0) {
            if (v.x > 0) mx.x += v.x * dt * 1.0f; else mn.x += v.x * dt * 1.0f;
            if (v.y > 0) mx.y += v.y * dt * 1.0f; else mn.y += v.y * dt * 1.0f;
            if (v.z > 0) mx.z += v.z * dt * 1.0f; else mn.z += v.z * dt * 1.0f;
        }
        w.bmin[i] = f4_of(mn, 0.0f);
        w.bmax[i] = f4_of(mx, 0.0f);
        cx = 0.5f * (mn.x + mx.x);
        cy = 0.5f * (mn.y + mx.y);
        cz = 0.5f * (mn.z + mx.z);
        side = fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z);
        if (!(side >= 0.0f && side < 3.0e38f)) side = 0.0f;  // a non-finite box does not steer the grid size
    }
    // Scene bounds of AABB centres (wave then block reduction; one partial
    // per block, folded by k_bounds_reduce) and the histogram of largest
    // sides (one atomic per block and non-empty bin, bins on separate lines).
    __shared__ uint32_t part[4][6];
    __shared__ uint32_t hist[64];
    if (threadIdx.x < 64) hist[threadIdx.x] = 0;
    __syncthreads();
    if (valid) atomicAdd(&hist[side_bin(side)], 1u);
    uint32_t v6[6] = {valid ? f2ord(cx) : 0xffffffffu, valid ? f2ord(cy) : 0xffffffffu,
                      valid ? f2ord(cz) : 0xffffffffu, valid ? f2ord(cx) : 0u, valid ? f2ord(cy) : 0u,
                      valid ? f2ord(cz) : 0u};
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            v6[k] = min(v6[k], (uint32_t)__shfl_xor((int)v6[k], off));
            v6[3 + k] = max(v6[3 + k], (uint32_t)__shfl_xor((int)v6[3 + k], off));
        }
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; ++k) part[wv][k] = v6[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        uint32_t r = part[0][k];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) r = k < 3 ? min(r, part[q][k]) : max(r, part[q][k]);
        w.bounds_part[blockIdx.x * 6 + k] = r;
    }
    if (threadIdx.x < 64 && hist[threadIdx.x]) atomicAdd(&w.counters[ctr(CTR_SIDE0 + threadIdx.x)], hist[threadIdx.x]);
}

__global__ void __launch_bounds__(256) k_bounds_reduce(DeviceWorld w, uint32_t nparts) {
    __shared__ uint32_t red[6][256];
    __shared__ uint32_t hist[64];
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0, 0, 0};
    for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x)
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], w.bounds_part[p * 6 + k]);
            hi[k] = max(hi[k], w.bounds_part[p * 6 + 3 + k]);
        }
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = lo[k];
        red[3 + k][threadIdx.x] = hi[k];
    }
    if (threadIdx.x < 64) hist[threadIdx.x] = w.counters[ctr(CTR_SIDE0 + threadIdx.x)];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = min(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = max(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) w.bounds[threadIdx.x] = red[threadIdx.x][0];
    if (threadIdx.x == 6) {
        // Grid base cell edge: the largest side of 90% of the bodies fits the
        // finest level (bodies at level 0 have largest side <= 0.99 * edge).
        uint32_t need = w.nbodies - w.nbodies / 10, acc = 0;
        int b = 0;
        for (; b < 63; ++b) {
            acc += hist[b];
            if (acc >= need) break;
        }
        w.bounds[6] = __float_as_uint(side_bin_edge(b) * (1.0f / 0.99f) * 1.0001f);
    }
}

// Entity.PreUpdatePosition (EntityBase.cs:1124-1155), discrete mode.
// PositionUpdater: active entities only (ContinuousPositionUpdater.cs:71).
__global__ void __launch_bounds__(256) k_integrate(DeviceWorld w) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nbodies) return;
    const uint32_t bf = w.bflags[i];
    if (!(bf & BF_PRESENT) || (bf & BF_ASLEEP)) return;
    if (!(bf & BF_DYNAMIC) && w.stab[i].y > 0.0f) return;
    float dt = w.s.dt;
    float4 Q = w.quat[i];
    float4 P = w.pos[i];
    f3 v = f3_of(w.linv[i]);
    f3 av = f3_of(w.angv[i]);
    q4 q{Q.x, Q.y, Q.z, Q.w};
    f3 inc = av * (dt * .5f);
    q4 m{inc.x, inc.y, inc.z, 0};
    m = qmul(m, q);
    q = q4{q.x + m.x, q.y + m.y, q.z + m.z, q.w + m.w};
    q = qnormalize(q);
    w.quat[i] = make_float4(q.x, q.y, q.z, q.w);
    f3 p = f3_of(P);
    if (bf & BF_CCD_MODE) {
        // Passive and continuous bodies move once their pairs' times of
        // impact are known (k_ccd); until then their collidables keep the
        // orientation of the step's start.
        w.qold[i] = Q;
    } else {
        inc = v * dt;
        p = p + inc;
        w.pos[i] = f4_of(p, 0.0f);
    }
    // The non-finite guard (the reference validates states with MathChecker in
    // its debug builds, BEPUutilities/MathChecker.cs:23-50): a body whose state
    // is no longer finite faults its world, which leaves from the next step on.
    const bool finite = isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(q.x) && isfinite(q.y) &&
                        isfinite(q.z) && isfinite(q.w) && isfinite(v.x) && isfinite(v.y) && isfinite(v.z) &&
                        isfinite(av.x) && isfinite(av.y) && isfinite(av.z);
    if (!finite) {
        w.wfault[w.wid[i]] = 1u;
        atomicOr(&w.counters[ctr(CTR_STICKY)], STICKY_NONFINITE);
    }
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Start of a step: clears the counters and this step's pair table.
__global__ void __launch_bounds__(256) k_step_reset(DeviceWorld w) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.hash_cap) w.hash_keys_cur[i] = ~0ull;
    if (i < w.gt_cap) {
        w.gt_key[i] = ~0ull;
        w.gt_head[i] = 0xffffffffu;
    }
    if (i == CTR_STEPS) w.counters[ctr(i)] += 1u;
    if (i < CTR_COUNT && i != CTR_STICKY && i != CTR_PAIRS_PREV && i != CTR_STEPS) {
        // The previous step's overflow and barrier abort stay visible in the
        // sticky word; the thread that clears a flag folds it in.  The pair
        // count is kept as the last step's.
        uint32_t v = w.counters[ctr(i)];
        if (v && i == CTR_OVERFLOW) atomicOr(&w.counters[ctr(CTR_STICKY)], STICKY_OVERFLOW);
        if (v && i == CTR_GBAR_ABORT) atomicOr(&w.counters[ctr(CTR_STICKY)], STICKY_GBAR_ABORT);
        if (i == CTR_PAIRS) w.counters[ctr(CTR_PAIRS_PREV)] = v;
        w.counters[ctr(i)] = 0;
    }
    if (i < w.nworlds) w.wlevels[i] = 0;
    if (i < 16 * w.nworlds) w.wlev_count[i] = 0;
    if (w.sleeping) {
        // per-island sums of the deactivation stage
        if (i < w.nbodies) {
            w.lab_act[i] = 0u;
            w.lab_count[i] = 0ull;
            w.lab_min[i] = 0xffffffffu;
            w.lab_flags[i] = 0u;
        }
        if (i < w.nworlds) w.wsleepy[i] = 0u;
    }
}

void launch_step_reset(const DeviceWorld& w, hipStream_t st) {
    uint32_t n = std::max<uint32_t>(std::max<uint32_t>(w.hash_cap, w.gt_cap), std::max<uint32_t>(CTR_COUNT, 16 * w.nworlds));
    n = std::max<uint32_t>(n, w.nbodies);
    hipLaunchKernelGGL(k_step_reset, dim3((n + 255) / 256), dim3(256), 0, st, w);
}

__global__ void __launch_bounds__(256) k_scatter_bodies(DeviceWorld w, const BodyRecord* recs, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BodyRecord& r = recs[i];
    uint32_t b = r.target;
    w.pos[b] = r.pos;
    w.quat[b] = r.quat;
    w.linv[b] = r.linv;
    w.angv[b] = r.angv;
    w.ii0[b] = r.ii0;
    w.ii1[b] = r.ii1;
    w.ii2[b] = r.ii2;
    w.li0[b] = r.li0;
    w.li1[b] = r.li1;
    w.li2[b] = r.li2;
    w.shape[b] = r.shape;
    w.stab[b] = r.stab;
    w.mat[b] = r.mat;
    w.damp[b] = r.damp;
    w.bflags[b] = r.flags;
    w.comp_root[b] = b;  // its own island until its first step
    w.wid[b] = r.wid;
    w.wbase[b] = r.wbase;
    w.born[b] = r.born;
    w.minrad[b] = r.minrad;
    w.toi[b] = 0x3f800000u;  // 1.0f
}

__global__ void __launch_bounds__(256) k_gather_states(DeviceWorld w, const uint32_t* ids, uint32_t n, float4* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t b = ids[i];
    out[4 * i + 0] = w.pos[b];
    out[4 * i + 1] = w.quat[b];
    out[4 * i + 2] = w.linv[b];
    out[4 * i + 3] = w.angv[b];
}

__global__ void __launch_bounds__(256) k_set_states(DeviceWorld w, const StateRecord* recs, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const StateRecord& r = recs[i];
    uint32_t b = r.target;
    if (r.fields & kStateUpdateMode) {
        w.bflags[b] = (w.bflags[b] & ~BF_CCD_MODE) | r.pad[0];
        return;
    }
    if (r.fields & BEPU_STATE_POSITION) w.pos[b] = r.pos;
    if (r.fields & BEPU_STATE_ORIENTATION) w.quat[b] = r.quat;
    if (r.fields & BEPU_STATE_LINEAR_VELOCITY) w.linv[b] = r.linv;
    if (r.fields & BEPU_STATE_ANGULAR_VELOCITY) w.angv[b] = r.angv;
    // Every setter activates the entity (SimulationIslandMember.Activate,
    // SimulationIslandMember.cs:256-277): a dynamic body's candidate flag and
    // time below the limit clear and its island is active at the next step;
    // a kinematic body's flag becomes -1.
    const uint32_t f = w.bflags[b];
    if (f & BF_DYNAMIC) {
        w.bflags[b] = (f & ~(uint32_t)BF_CANDIDATE) | BF_WAKE;
        w.stab[b].y = 0.0f;
    } else if ((f & BF_PRESENT) && !(f & BF_MESH)) {
        w.stab[b].y = -1.0f;
        w.bflags[b] = f | BF_KIN_ACTIVE;
    }
}

__global__ void __launch_bounds__(256) k_remove_bodies(DeviceWorld w, const uint32_t* ids, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w.bflags[ids[i]] = 0u;
}

void launch_set_states(const DeviceWorld& w, const StateRecord* recs, uint32_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_set_states, dim3((n + 255) / 256), dim3(256), 0, st, w, recs, n);
}

void launch_remove_bodies(const DeviceWorld& w, const uint32_t* ids, uint32_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_remove_bodies, dim3((n + 255) / 256), dim3(256), 0, st, w, ids, n);
}

void launch_scatter_bodies(const DeviceWorld& w, const BodyRecord* recs, uint32_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_scatter_bodies, dim3((n + 255) / 256), dim3(256), 0, st, w, recs, n);
}

void launch_gather_states(const DeviceWorld& w, const uint32_t* ids, uint32_t n, float4* out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_gather_states, dim3((n + 255) / 256), dim3(256), 0, st, w, ids, n, out);
}

void launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}

void launch_body_pre(const DeviceWorld& w, hipStream_t st) {
    if (w.nbodies == 0) return;
    uint32_t blocks = (w.nbodies + 255) / 256;
    hipLaunchKernelGGL(k_body_pre, dim3(blocks), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_bounds_reduce, dim3(1), dim3(256), 0, st, w, blocks);
}

void launch_integrate(const DeviceWorld& w, hipStream_t st) {
    if (w.nbodies == 0) return;
    hipLaunchKernelGGL(k_integrate, dim3((w.nbodies + 255) / 256), dim3(256), 0, st, w);
}

}  // namespace bepu
