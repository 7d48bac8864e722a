// solver_rows.h -- per-constraint math of the contact solver, shared by the
// island kernels and the global fallback path.
//
// A contact manifold constraint is one SolverGroup: c penetration rows plus
// (convex pairs) one two-axis sliding friction row and one twist friction row,
// or (non-convex pairs: sphere-sphere, box-sphere) one friction row per
// contact.  Rows follow ContactPenetrationConstraint (ContactPenetrationConstraint.cs:114-321),
// SlidingFrictionTwoAxis (SlidingFrictionTwoAxis.cs:155-482), TwistFrictionConstraint
// (TwistFrictionConstraint.cs:65-204) and ContactFrictionConstraint
// (ContactFrictionConstraint.cs:131-287) with the reference's operation order;
// the early-out follows SolverGroup.SolveUpdateable (SolverGroup.cs:117-150).
#pragma once
#include "manifold.h"
#include "world.h"

namespace bepu {

// Sub-constraint activity bits: 0..3 penetration, 4 sliding, 5 twist,
// 6..9 non-convex friction; bit 31 = group (isActiveInSolver of the manifold
// constraint).  Zero-impulse counters (SolverSettings.iterationsAtZeroImpulse)
// are 3 bits per sub-constraint, enough for minimum iteration counts <= 6.
constexpr uint32_t kGroupBit = 1u << 31;

BEPU_HD bool is_convex_kind(int kind) { return kind == BEPU_PAIR_BOX_BOX || kind == BEPU_PAIR_CONVEX; }

struct BodyV {
    f3 v, w;
    m3 I;
    float im;
    bool dyn;
};

// Values fixed by the prestep for the whole iteration loop.
struct RowConst {
    int kind, npen;
    f3 plin[4], pA[4], pB[4];
    float psoft[4], pbias[4], pv2i[4];
    union {
        struct {  // convex pairs: sliding and twist friction
            float L[6], sA[6], sB[6], sv2i[4], sfric;
            f3 sra, srb;
            f3 tang;
            float tv2i, tfric, lever[4];
        } c;
        struct {  // non-convex pairs: one friction row per contact
            f3 flin[4], fA[4], fB[4];
            float ffric[4], fv2i[4];
        } n;
        struct {  // joints: the point constraint of a BallSocketJoint
            m3 M;           // massMatrix
            f3 ra, rb;      // worldOffsetA / worldOffsetB
            f3 bias;        // biasVelocity
            float soft;     // softness
        } j;
    };
};

// RowConst.kind of a joint (pair kinds are BEPU_PAIR_*).  A joint's RowVar
// holds its accumulated impulse in pacc[0..2], isActiveInSolver in kGroupBit
// and iterationsAtZeroImpulse in `zero`.
constexpr int kJointKind = 16;

// Values the iterations update.
struct RowVar {
    float pacc[4];
    float facc[4];
    float sacc[2];
    float tacc;
    uint32_t act, zero;
};

// Entity.ApplyLinearImpulse / ApplyAngularImpulse (EntityBase.cs:754-788)
BEPU_D void apply_lin(BodyV& b, f3 l) {
    b.v.x += l.x * b.im;
    b.v.y += l.y * b.im;
    b.v.z += l.z * b.im;
}
BEPU_D void apply_ang(BodyV& b, f3 i) {
    b.w.x += i.x * b.I.m11 + i.y * b.I.m21 + i.z * b.I.m31;
    b.w.y += i.x * b.I.m12 + i.y * b.I.m22 + i.z * b.I.m32;
    b.w.z += i.x * b.I.m13 + i.y * b.I.m23 + i.z * b.I.m33;
}
// Effective-mass entry j I^-1 j + m^-1 (ContactPenetrationConstraint.cs:149-170)
BEPU_D float eff_entry(f3 j, const BodyV& b) {
    float tX = j.x * b.I.m11 + j.y * b.I.m21 + j.z * b.I.m31;
    float tY = j.x * b.I.m12 + j.y * b.I.m22 + j.z * b.I.m32;
    float tZ = j.x * b.I.m13 + j.y * b.I.m23 + j.z * b.I.m33;
    return tX * j.x + tY * j.y + tZ * j.z + b.im;
}
// ContactPenetrationConstraint.RelativeVelocity (ContactPenetrationConstraint.cs:86-104)
BEPU_D float pen_relvel(f3 lin, f3 aA, f3 aB, const BodyV& A, const BodyV& B) {
    float l = A.v.x * lin.x + A.v.y * lin.y + A.v.z * lin.z + A.w.x * aA.x + A.w.y * aA.y + A.w.z * aA.z;
    l += -B.v.x * lin.x - B.v.y * lin.y - B.v.z * lin.z + B.w.x * aB.x + B.w.y * aB.y + B.w.z * aB.z;
    return l;
}
BEPU_D void row_apply(f3 lin, f3 aA, f3 aB, float lambda, BodyV& A, BodyV& B) {
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
BEPU_D void sliding_apply(const RowConst& r, float lx, float ly, BodyV& A, BodyV& B) {
    f3 lin{lx * r.c.L[0] + ly * r.c.L[3], lx * r.c.L[1] + ly * r.c.L[4], lx * r.c.L[2] + ly * r.c.L[5]};
    if (A.dyn) {
        apply_lin(A, lin);
        apply_ang(A, f3{lx * r.c.sA[0] + ly * r.c.sA[3], lx * r.c.sA[1] + ly * r.c.sA[4], lx * r.c.sA[2] + ly * r.c.sA[5]});
    }
    if (B.dyn) {
        apply_lin(B, -lin);
        apply_ang(B, f3{lx * r.c.sB[0] + ly * r.c.sB[3], lx * r.c.sB[1] + ly * r.c.sB[4], lx * r.c.sB[2] + ly * r.c.sB[5]});
    }
}
BEPU_D void twist_apply(f3 ang, float lambda, BodyV& A, BodyV& B) {
    f3 a = ang * lambda;
    if (A.dyn) apply_ang(A, a);
    if (B.dyn) apply_ang(B, -a);
}

// Prestep of one manifold constraint: Solver.UnsafePrestep ->
// UpdateSolverActivity -> group Update (Solver.cs:265-275,
// ContactManifoldConstraint.cs:128-145, SolverGroup.cs:70-76).  Reads the
// pair state, updates its remembered friction directions in place.
BEPU_D void prestep_row(const bepu_settings& st, PairState& s, f3 pA, f3 pB, const BodyV& A, const BodyV& B,
                        RowConst& r, RowVar& v) {
    int npen = s.npen;
    bool convex = is_convex_kind(s.kind);
    r.kind = s.kind;
    r.npen = npen;
    uint32_t act = 0;
    bool group = A.dyn || B.dyn;  // dynamic bodies are always active
    if (group) act |= kGroupBit;
    float updateRate = 1 / st.dt;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v.pacc[k] = s.penacc[k];
        v.facc[k] = s.fracc[k];
        if (k >= npen) continue;
        int cs = s.pslot[k];
        float4 cp = get4(s.cpos, cs), cn = get4(s.cnrm, cs);
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
        r.plin[k] = lin;
        r.pA[k] = aA;
        r.pB[k] = aB;
        r.psoft[k] = soft;
        r.pbias[k] = bias;
        r.pv2i[k] = v2i;
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
            f3 flin = f3_of(s.frlin[k]);
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
            r.n.flin[k] = flin;
            r.n.ffric[k] = fric;
            if (factive) {
                f3 fa{(ra.y * flin.z) - (ra.z * flin.y), (ra.z * flin.x) - (ra.x * flin.z), (ra.x * flin.y) - (ra.y * flin.x)};
                f3 fb{(flin.y * rb.z) - (flin.z * rb.y), (flin.z * rb.x) - (flin.x * rb.z), (flin.x * rb.y) - (flin.y * rb.x)};
                float fA = A.dyn ? eff_entry(fa, A) : 0.0f;
                float fB = B.dyn ? eff_entry(fb, B) : 0.0f;
                r.n.fA[k] = fa;
                r.n.fB[k] = fb;
                r.n.fv2i[k] = -1 / (fA + fB);
                act |= 1u << (6 + k);
            }
        }
    }
    v.sacc[0] = s.slacc[0];
    v.sacc[1] = s.slacc[1];
    v.tacc = s.twacc;
    if (convex && group && (s.flags & PF_FRICTION_ACTIVE)) {
        // SlidingFrictionTwoAxis.Update (SlidingFrictionTwoAxis.cs:242-440)
        f3 center;
        const f3 c0 = f3_of(get4(s.cpos, s.pslot[0])), c1 = f3_of(get4(s.cpos, s.pslot[1])),
                 c2 = f3_of(get4(s.cpos, s.pslot[2])), c3 = f3_of(get4(s.cpos, s.pslot[3]));
        if (npen == 1) center = c0;
        else if (npen == 2) center = (c0 + c1) * .5f;
        else if (npen == 3) center = (c2 + (c0 + c1)) * .333333333f;
        else center = (c3 + (c2 + (c0 + c1))) * .25f;
        f3 ra = center - pA;
        f3 vA = cross3(A.w, ra) + A.v;
        f3 rb = center - pB;
        f3 vB = cross3(B.w, rb) + B.v;
        f3 rv = vA - vB;
        f3 n = f3_of(get4(s.cnrm, s.pslot[0]));
        float nvs = n.x * rv.x + n.y * rv.y + n.z * rv.z;
        rv.x -= nvs * n.x;
        rv.y -= nvs * n.y;
        rv.z -= nvs * n.z;
        float* L = r.c.L;
#pragma unroll
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
#pragma unroll
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
        r.c.sA[0] = a0.x; r.c.sA[1] = a0.y; r.c.sA[2] = a0.z; r.c.sA[3] = a1.x; r.c.sA[4] = a1.y; r.c.sA[5] = a1.z;
        r.c.sB[0] = b0.x; r.c.sB[1] = b0.y; r.c.sB[2] = b0.z; r.c.sB[3] = b1.x; r.c.sB[4] = b1.y; r.c.sB[5] = b1.z;
        r.c.sv2i[0] = m22 * di;
        r.c.sv2i[1] = -m12 * di;
        r.c.sv2i[2] = -m21 * di;
        r.c.sv2i[3] = m11 * di;
        r.c.sfric = fric;
        r.c.sra = ra;
        r.c.srb = rb;
        act |= 1u << 4;
        // TwistFrictionConstraint.Update (TwistFrictionConstraint.cs:114-175)
        float tA = A.dyn ? eff_entry(n, A) : 0.0f;
        float tB = B.dyn ? eff_entry(n, B) : 0.0f;
        r.c.tv2i = -1 / (tA + tB);
        float rav = A.w.x * n.x + A.w.y * n.y + A.w.z * n.z;
        rav -= B.w.x * n.x + B.w.y * n.y + B.w.z * n.z;
        float tf = fabsf(rav) > st.static_friction_velocity_threshold ||
                           fabsf(rv.x) + fabsf(rv.y) + fabsf(rv.z) > st.static_friction_velocity_threshold
                       ? s.kinetic
                       : s.static_f;
        r.c.tfric = tf * st.twist_friction_factor;
        r.c.tang = n;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const f3 ck = k == 0 ? c0 : (k == 1 ? c1 : (k == 2 ? c2 : c3));
            r.c.lever[k] = k < npen ? len3(ck - center) : 0.0f;
        }
        act |= 1u << 5;
    }
    v.act = act;
    v.zero = 0;
}

// Matrix3x3.CreateCrossProduct (BEPUutilities/Matrix3x3.cs:378-389)
BEPU_D m3 cross_matrix(f3 v) { return m3{0, -v.z, v.y, v.z, 0, -v.x, -v.y, v.x, 0}; }
// Matrix3x3.Multiply (Matrix3x3.cs:766-791)
BEPU_D m3 mat_mul(const m3& a, const m3& b) {
    m3 r;
    r.m11 = a.m11 * b.m11 + a.m12 * b.m21 + a.m13 * b.m31;
    r.m12 = a.m11 * b.m12 + a.m12 * b.m22 + a.m13 * b.m32;
    r.m13 = a.m11 * b.m13 + a.m12 * b.m23 + a.m13 * b.m33;
    r.m21 = a.m21 * b.m11 + a.m22 * b.m21 + a.m23 * b.m31;
    r.m22 = a.m21 * b.m12 + a.m22 * b.m22 + a.m23 * b.m32;
    r.m23 = a.m21 * b.m13 + a.m22 * b.m23 + a.m23 * b.m33;
    r.m31 = a.m31 * b.m11 + a.m32 * b.m21 + a.m33 * b.m31;
    r.m32 = a.m31 * b.m12 + a.m32 * b.m22 + a.m33 * b.m32;
    r.m33 = a.m31 * b.m13 + a.m32 * b.m23 + a.m33 * b.m33;
    return r;
}
// Matrix3x3.Subtract (Matrix3x3.cs:973-997)
BEPU_D m3 mat_sub(const m3& a, const m3& b) {
    return m3{a.m11 - b.m11, a.m12 - b.m12, a.m13 - b.m13, a.m21 - b.m21, a.m22 - b.m22,
              a.m23 - b.m23, a.m31 - b.m31, a.m32 - b.m32, a.m33 - b.m33};
}

// Solver.UnsafePrestep -> BallSocketJoint.Update (Solver.cs:265-275,
// BallSocketJoint.cs:229-300): world offsets of the anchors, the softened
// effective mass of the point constraint and the bias toward closing the gap
// between the anchors.  qA / qB give the bodies' orientation matrices
// (Matrix3x3.CreateFromQuaternion, as the entities keep them).
BEPU_D void prestep_joint(const bepu_settings& st, const JointState& J, f3 pA, q4 qA, f3 pB, q4 qB, const BodyV& A,
                          const BodyV& B, RowConst& r, RowVar& v) {
    r.kind = kJointKind;
    r.npen = 0;
    r.j.ra = mat_transform(f3_of(J.la), mat_from_quat(qA));
    r.j.rb = mat_transform(f3_of(J.lb), mat_from_quat(qB));
    // SpringSettings.ComputeErrorReductionAndSoftness (SpringSettings.cs:109-124), update rate 1/dt
    const float dt = st.dt;
    const float rate = 1 / dt;
    const float multiplier = 1 / (dt * J.la.w + J.lb.w);
    const float erp = J.la.w * multiplier;
    r.j.soft = rate * multiplier;
    const m3 rA = cross_matrix(r.j.ra), rB = cross_matrix(r.j.rb);
    m3 k;
    if (A.dyn && B.dyn) {
        const m3 lin{A.im + B.im, 0, 0, 0, A.im + B.im, 0, 0, 0, A.im + B.im};
        const m3 angA = mat_mul(mat_mul(rA, A.I), rA), angB = mat_mul(mat_mul(rB, B.I), rB);
        k = mat_sub(mat_sub(lin, angA), angB);
    } else if (A.dyn) {
        const m3 lin{A.im, 0, 0, 0, A.im, 0, 0, 0, A.im};
        k = mat_sub(lin, mat_mul(mat_mul(rA, A.I), rA));
    } else {
        const m3 lin{B.im, 0, 0, 0, B.im, 0, 0, 0, B.im};
        k = mat_sub(lin, mat_mul(mat_mul(rB, B.I), rB));
    }
    k.m11 += r.j.soft;
    k.m22 += r.j.soft;
    k.m33 += r.j.soft;
    r.j.M = mat_invert(k);
    f3 err = pB + r.j.rb;
    err = err - pA;
    err = err - r.j.ra;
    r.j.bias = err * -erp;
    const float len = lensq3(r.j.bias);
    if (len > J.max_corr_sq) {
        const float m = J.acc.w / sqrtf(len);
        r.j.bias.x *= m;
        r.j.bias.y *= m;
        r.j.bias.z *= m;
    }
    v.pacc[0] = J.acc.x;
    v.pacc[1] = J.acc.y;
    v.pacc[2] = J.acc.z;
    v.pacc[3] = 0;
    v.act = kGroupBit;
    v.zero = 0;
}

// The joint's impulse on its connections (BallSocketJoint.cs:308-328, 364-383).
BEPU_D void joint_apply(const RowConst& r, f3 impulse, BodyV& A, BodyV& B) {
    if (A.dyn) {
        const f3 l = -impulse;
        apply_lin(A, l);
        apply_ang(A, cross3(r.j.ra, l));
    }
    if (B.dyn) {
        apply_lin(B, impulse);
        apply_ang(B, cross3(r.j.rb, impulse));
    }
}

// Solver.UnsafeSolveIteration (Solver.cs:285-314) -> BallSocketJoint.SolveIteration
// (BallSocketJoint.cs:335-388).
BEPU_D void solve_joint(const RowConst& r, RowVar& v, BodyV& A, BodyV& B, float mi, int mic) {
    const f3 aVel = A.v + cross3(A.w, r.j.ra);
    const f3 bVel = B.v + cross3(B.w, r.j.rb);
    f3 lambda;
    lambda.x = aVel.x - bVel.x + r.j.bias.x - r.j.soft * v.pacc[0];
    lambda.y = aVel.y - bVel.y + r.j.bias.y - r.j.soft * v.pacc[1];
    lambda.z = aVel.z - bVel.z + r.j.bias.z - r.j.soft * v.pacc[2];
    lambda = mat_transform(lambda, r.j.M);
    v.pacc[0] += lambda.x;
    v.pacc[1] += lambda.y;
    v.pacc[2] += lambda.z;
    joint_apply(r, lambda, A, B);
    const float impulse = fabsf(lambda.x) + fabsf(lambda.y) + fabsf(lambda.z);
    if (impulse < mi) {
        v.zero++;
        if ((int)v.zero > mic) v.act &= ~kGroupBit;
    } else {
        v.zero = 0;
    }
}

// Solver.UnsafeExclusiveUpdate -> group ExclusiveUpdate (Solver.cs:277-283)
BEPU_D void warm_row(const RowConst& r, const RowVar& v, BodyV& A, BodyV& B) {
    if (!(v.act & kGroupBit)) return;
    if (r.kind == kJointKind) {
        joint_apply(r, f3{v.pacc[0], v.pacc[1], v.pacc[2]}, A, B);
        return;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < r.npen && (v.act & (1u << k))) row_apply(r.plin[k], r.pA[k], r.pB[k], v.pacc[k], A, B);
    if (is_convex_kind(r.kind)) {
        if (v.act & (1u << 4)) sliding_apply(r, v.sacc[0], v.sacc[1], A, B);
        if (v.act & (1u << 5)) twist_apply(r.c.tang, v.tacc, A, B);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < r.npen && (v.act & (1u << (6 + k)))) row_apply(r.n.flin[k], r.n.fA[k], r.n.fB[k], v.facc[k], A, B);
    }
}

// SolverGroup.SolveUpdateable bookkeeping after sub-constraint k solved.
BEPU_D void account(RowVar& v, int k, float impulse, float min_impulse, int min_iter, int& active) {
    if (impulse < min_impulse) {
        int z = (int)((v.zero >> (3 * k)) & 7u) + 1;
        if (z > min_iter) v.act &= ~(1u << k);
        else active++;
        v.zero = (v.zero & ~(7u << (3 * k))) | ((uint32_t)(z > 7 ? 7 : z) << (3 * k));
    } else {
        v.zero &= ~(7u << (3 * k));
        active++;
    }
}

// Solver.UnsafeSolveIteration -> ContactManifoldConstraint.SolveIteration
// (Solver.cs:285-314; ConvexContactManifoldConstraint.cs:199-212;
// NonConvexContactManifoldConstraint.cs:201-212).
BEPU_D void solve_row(const RowConst& r, RowVar& v, BodyV& A, BodyV& B, float mi, int mic) {
    if (!(v.act & kGroupBit)) return;
    if (r.kind == kJointKind) {
        solve_joint(r, v, A, B, mi, mic);
        return;
    }
    int active = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= r.npen || !(v.act & (1u << k))) continue;
        float lambda = (pen_relvel(r.plin[k], r.pA[k], r.pB[k], A, B) - r.pbias[k] + r.psoft[k] * v.pacc[k]) * r.pv2i[k];
        float prev = v.pacc[k];
        v.pacc[k] = maxf_ref(0, v.pacc[k] + lambda);
        lambda = v.pacc[k] - prev;
        row_apply(r.plin[k], r.pA[k], r.pB[k], lambda, A, B);
        account(v, k, fabsf(lambda), mi, mic, active);
    }
    if (is_convex_kind(r.kind)) {
        if (v.act & (1u << 4)) {
            f3 ra = r.c.sra, rb = r.c.srb;
            float dvx = A.v.x + (A.w.y * ra.z) - (A.w.z * ra.y);
            float dvy = A.v.y + (A.w.z * ra.x) - (A.w.x * ra.z);
            float dvz = A.v.z + (A.w.x * ra.y) - (A.w.y * ra.x);
            dvx += -B.v.x - (B.w.y * rb.z) + (B.w.z * rb.y);
            dvy += -B.v.y - (B.w.z * rb.x) + (B.w.x * rb.z);
            dvz += -B.v.z - (B.w.x * rb.y) + (B.w.y * rb.x);
            float lx = dvx * r.c.L[0] + dvy * r.c.L[1] + dvz * r.c.L[2];
            float ly = dvx * r.c.L[3] + dvy * r.c.L[4] + dvz * r.c.L[5];
            float x = lx;
            lx = x * r.c.sv2i[0] + ly * r.c.sv2i[2];
            ly = x * r.c.sv2i[1] + ly * r.c.sv2i[3];
            float px = v.sacc[0], py = v.sacc[1];
            v.sacc[0] += lx;
            v.sacc[1] += ly;
            float length = v.sacc[0] * v.sacc[0] + v.sacc[1] * v.sacc[1];
            float maxF = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < r.npen) maxF += v.pacc[k];
            maxF *= r.c.sfric;
            if (length > maxF * maxF) {
                length = maxF / sqrtf(length);
                v.sacc[0] *= length;
                v.sacc[1] *= length;
            }
            lx = v.sacc[0] - px;
            ly = v.sacc[1] - py;
            sliding_apply(r, lx, ly, A, B);
            account(v, 4, fabsf(lx) + fabsf(ly), mi, mic, active);
        }
        if (v.act & (1u << 5)) {
            f3 ang = r.c.tang;
            float lambda = A.w.x * ang.x + A.w.y * ang.y + A.w.z * ang.z;
            lambda -= B.w.x * ang.x + B.w.y * ang.y + B.w.z * ang.z;
            lambda *= r.c.tv2i;
            float prev = v.tacc;
            float maxF = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < r.npen) maxF += r.c.lever[k] * v.pacc[k];
            maxF *= r.c.tfric;
            v.tacc = clampf_ref(v.tacc + lambda, -maxF, maxF);
            lambda = v.tacc - prev;
            twist_apply(ang, lambda, A, B);
            account(v, 5, fabsf(lambda), mi, mic, active);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= r.npen || !(v.act & (1u << (6 + k)))) continue;
            f3 lin = r.n.flin[k], aA = r.n.fA[k], aB = r.n.fB[k];
            float vel = 0;
            vel += A.v.x * lin.x + A.v.y * lin.y + A.v.z * lin.z + A.w.x * aA.x + A.w.y * aA.y + A.w.z * aA.z;
            vel += -B.v.x * lin.x - B.v.y * lin.y - B.v.z * lin.z + B.w.x * aB.x + B.w.y * aB.y + B.w.z * aB.z;
            float lambda = vel * r.n.fv2i[k];
            float prev = v.facc[k];
            float maxForce = r.n.ffric[k] * v.pacc[k];
            v.facc[k] = clampf_ref(v.facc[k] + lambda, -maxForce, maxForce);
            lambda = v.facc[k] - prev;
            row_apply(lin, aA, aB, lambda, A, B);
            account(v, 6 + k, fabsf(lambda), mi, mic, active);
        }
    }
    if (active == 0) v.act &= ~kGroupBit;
}

// Accumulated impulses back into the persistent pair state (warm start of the next step).
BEPU_D void writeback_row(PairState& s, const RowVar& v) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        s.penacc[k] = v.pacc[k];
        s.fracc[k] = v.facc[k];
    }
    if (is_convex_kind(s.kind) && (s.flags & PF_FRICTION_ACTIVE)) {
        s.slacc[0] = v.sacc[0];
        s.slacc[1] = v.sacc[1];
        s.twacc = v.tacc;
    }
}

// The joint's accumulated impulse back into its state (warm start of the next step).
BEPU_D void writeback_joint(JointState& J, const RowVar& v) {
    J.acc.x = v.pacc[0];
    J.acc.y = v.pacc[1];
    J.acc.z = v.pacc[2];
}

BEPU_D BodyV load_body_global(const DeviceWorld& w, uint32_t b) {
    BodyV r;
    r.v = f3_of(w.linv[b]);
    r.w = f3_of(w.angv[b]);
    float4 a = w.ii0[b], c = w.ii1[b], d = w.ii2[b];
    r.I = m3{a.x, a.y, a.z, c.x, c.y, c.z, d.x, d.y, d.z};
    r.im = a.w;
    r.dyn = (w.bflags[b] & 1u) != 0;
    return r;
}

}  // namespace bepu
