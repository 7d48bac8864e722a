// manifold.h -- contact list / contact-constraint bookkeeping on a PairState
// held in registers, shared by the narrowphase kernels.
#pragma once
#include "world.h"

namespace bepu {

__device__ __forceinline__ bool convex_kind(int kind) { return kind == BEPU_PAIR_BOX_BOX || kind == BEPU_PAIR_CONVEX; }

// ConvexContactManifoldConstraint / NonConvexContactManifoldConstraint.AddContact
__device__ inline void add_contact(PairState& s, f3 pos, f3 nrm, float depth, int id) {
    int slot = 0;
    while (s.flags & (16u << slot)) ++slot;
    s.flags |= 16u << slot;
    s.cpos[slot] = f4_of(pos, depth);
    s.cnrm[slot] = f4_of(nrm, __int_as_float(id));
    s.clist[s.ccount++] = (int8_t)slot;
    int k = s.npen++;
    s.pslot[k] = (int8_t)slot;
    s.penacc[k] = 0;
    if (convex_kind(s.kind)) {
        if (!(s.flags & PF_FRICTION_ACTIVE)) {
            s.flags |= PF_FRICTION_ACTIVE;
            for (int q = 0; q < 6; ++q) s.sllin[q] = 0;
        }
    } else {
        s.fracc[k] = 0;
        s.frlin[k] = make_float4(0, 0, 0, 0);
    }
}

// ContactManifold.Remove (FastRemoveAt) -> RemoveContact (ordered RemoveAt of the rows)
__device__ inline void remove_contact(PairState& s, int index) {
    int slot = s.clist[index];
    s.ccount--;
    if (index < s.ccount) s.clist[index] = s.clist[s.ccount];
    s.flags &= ~(16u << slot);
    int idx = -1;
    for (int i = 0; i < s.npen; ++i)
        if (s.pslot[i] == slot) { idx = i; break; }
    if (idx < 0) return;
    for (int k = idx; k < s.npen - 1; ++k) {
        s.pslot[k] = s.pslot[k + 1];
        s.penacc[k] = s.penacc[k + 1];
        s.fracc[k] = s.fracc[k + 1];
        s.frlin[k] = s.frlin[k + 1];
    }
    s.npen--;
    if (convex_kind(s.kind) && s.npen == 0) {
        s.flags &= ~PF_FRICTION_ACTIVE;
        s.twacc = 0;
        s.slacc[0] = s.slacc[1] = 0;
    }
}

}  // namespace bepu
