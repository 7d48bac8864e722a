// k_joint.hip -- joints outside the solver: their island edges, set-up,
// spring settings and end.  The solver rows of a joint (prestep, warm start,
// iterations) live beside the contact rows in solver_rows.h, and joints are
// constraints of the islands, the colouring and both solve paths like the
// contact manifold constraints (k_island.hip, k_solver.hip).
//
//   k_joint_edges     a joint between two dynamic bodies unites their islands
//                     in this step's union-find (its SimulationIslandConnection);
//   k_joint_init      new BallSocketJoint(a, b, anchor) (BallSocketJoint.cs:
//                     57-63): each body's offset to the anchor in its local
//                     frame, taken from the body's current pose in stream
//                     order (the OffsetA / OffsetB setters, :108-130);
//   k_joint_settings  SpringSettings.Stiffness / Damping and
//                     Joint.MaxCorrectiveVelocity;
//   k_joint_sweep     a joint of a removed body ends, and its other
//                     connection, if dynamic, is activated (as the removal of
//                     any constraint's body activates the other).
#include <float.h>

#include "wave.h"
#include "world.h"

namespace bepu {

namespace {

__global__ void __launch_bounds__(256) k_joint_edges(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.njoints) return;
    const JointState& j = w.joints[i];
    if (joint_edge(w, j)) uf_union(w.uf_parent, j.a, j.b);
}

__global__ void k_joint_init(DeviceWorld w, uint32_t id, uint32_t a, uint32_t b, uint32_t local, float4 anchor) {
    if (threadIdx.x != 0) return;
    JointState j;
    j.a = a;
    j.b = b;
    j.local = local;
    j.flags = 0;
    const f3 p = f3_of(anchor);
    const float4 qa = w.quat[a], qb = w.quat[b];
    const f3 la = mat_transform_t(p - f3_of(w.pos[a]), mat_from_quat(q4{qa.x, qa.y, qa.z, qa.w}));
    const f3 lb = mat_transform_t(p - f3_of(w.pos[b]), mat_from_quat(q4{qb.x, qb.y, qb.z, qb.w}));
    j.la = f4_of(la, 600000.0f);  // SpringSettings.cs:34-35
    j.lb = f4_of(lb, 90000.0f);
    j.acc = make_float4(0, 0, 0, FLT_MAX);  // Joint.cs:14-15: no corrective velocity limit
    j.max_corr_sq = FLT_MAX;
    j.color = -1;
    j.pad[0] = j.pad[1] = 0;
    w.joints[id] = j;
}

__global__ void k_joint_settings(DeviceWorld w, uint32_t id, float stiffness, float damping, float max_corr,
                                 float max_corr_sq) {
    if (threadIdx.x != 0) return;
    JointState& j = w.joints[id];
    j.la.w = stiffness;
    j.lb.w = damping;
    j.acc.w = max_corr;
    j.max_corr_sq = max_corr_sq;
}

__global__ void __launch_bounds__(256) k_joint_sweep(DeviceWorld w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.njoints) return;
    JointState& j = w.joints[i];
    if (j.flags & JF_DEAD) return;
    const uint32_t fa = w.bflags[j.a], fb = w.bflags[j.b];
    const bool pa = (fa & BF_PRESENT) != 0, pb = (fb & BF_PRESENT) != 0;
    if (pa && pb) return;
    j.flags |= JF_DEAD;
    // SimulationIslandMember.Activate of the remaining connection
    // (SimulationIslandMember.cs:256-277): a dynamic body's candidate flag and
    // time below the limit clear, and its island is active at the next step
    // (as for the removal of a body's contacts, only where bodies may sleep).
    const uint32_t o = pa ? j.a : j.b;
    const uint32_t fo = pa ? fa : fb;
    if (w.sleeping && (pa || pb) && (fo & BF_DYNAMIC)) {
        atomicAnd(&w.bflags[o], ~(uint32_t)BF_CANDIDATE);
        atomicOr(&w.bflags[o], (uint32_t)BF_WAKE);
        w.stab[o].y = 0.0f;
    }
}

}  // namespace

void launch_joint_edges(const DeviceWorld& w, hipStream_t st) {
    if (w.njoints) hipLaunchKernelGGL(k_joint_edges, dim3((w.njoints + 255) / 256), dim3(256), 0, st, w);
}

void launch_joint_init(const DeviceWorld& w, uint32_t id, uint32_t a, uint32_t b, uint32_t local, float4 anchor,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_joint_init, dim3(1), dim3(64), 0, st, w, id, a, b, local, anchor);
}

void launch_joint_settings(const DeviceWorld& w, uint32_t id, float stiffness, float damping, float max_corr,
                           float max_corr_sq, hipStream_t st) {
    hipLaunchKernelGGL(k_joint_settings, dim3(1), dim3(64), 0, st, w, id, stiffness, damping, max_corr, max_corr_sq);
}

void launch_joint_sweep(const DeviceWorld& w, hipStream_t st) {
    if (w.njoints) hipLaunchKernelGGL(k_joint_sweep, dim3((w.njoints + 255) / 256), dim3(256), 0, st, w);
}

}  // namespace bepu
