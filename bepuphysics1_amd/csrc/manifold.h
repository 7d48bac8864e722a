// manifold.h -- contact list / contact-constraint bookkeeping on a PairState,
// shared by the narrowphase kernels.  Every run-time index goes through
// selects, so a PairState held in registers stays there.
#pragma once
#include "world.h"

namespace bepu {

__device__ __forceinline__ bool convex_kind(int kind) { return kind == BEPU_PAIR_BOX_BOX || kind == BEPU_PAIR_CONVEX; }

// Run-time indexed access to a 4-entry array through selects.  A pair state
// indexed only this way (and with constants) stays in registers; indexed at run
// time directly, it would be placed in scratch memory.  The four entries pass
// through an empty asm so that the optimiser cannot turn the selects of
// values back into one load from a selected address.
__device__ __forceinline__ void opaque(float& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void opaque(float4& x) {
    opaque(x.x);
    opaque(x.y);
    opaque(x.z);
    opaque(x.w);
}
__device__ __forceinline__ void opaque(int8_t& x) {
    int v = x;
    asm volatile("" : "+v"(v));
    x = (int8_t)v;
}
template <typename T>
__device__ __forceinline__ T get4(const T (&a)[4], int i) {
    T v0 = a[0], v1 = a[1], v2 = a[2], v3 = a[3];
    opaque(v0);
    opaque(v1);
    opaque(v2);
    opaque(v3);
    T r = v0;
    if (i == 1) r = v1;
    if (i == 2) r = v2;
    if (i == 3) r = v3;
    return r;
}
// Every entry is written (with itself or `v`), so no store needs a run-time
// address either.
template <typename T>
__device__ __forceinline__ void set4(T (&a)[4], int i, T v) {
    T v0 = a[0], v1 = a[1], v2 = a[2], v3 = a[3];
    opaque(v0);
    opaque(v1);
    opaque(v2);
    opaque(v3);
    a[0] = i == 0 ? v : v0;
    a[1] = i == 1 ? v : v1;
    a[2] = i == 2 ? v : v2;
    a[3] = i == 3 ? v : v3;
}

// ConvexContactManifoldConstraint / NonConvexContactManifoldConstraint.AddContact
__device__ __forceinline__ void add_contact(PairState& s, f3 pos, f3 nrm, float depth, int id) {
    const int slot = __ffs(~(s.flags >> 4)) - 1;  // the first unused contact slot
    s.flags |= 16u << slot;
    set4(s.cpos, slot, f4_of(pos, depth));
    set4(s.cnrm, slot, f4_of(nrm, __int_as_float(id)));
    set4(s.clist, s.ccount, (int8_t)slot);
    s.ccount++;
    const int k = s.npen++;
    set4(s.pslot, k, (int8_t)slot);
    set4(s.penacc, k, 0.0f);
    if (convex_kind(s.kind)) {
        if (!(s.flags & PF_FRICTION_ACTIVE)) {
            s.flags |= PF_FRICTION_ACTIVE;
#pragma unroll
            for (int q = 0; q < 6; ++q) s.sllin[q] = 0;
        }
    } else {
        set4(s.fracc, k, 0.0f);
        set4(s.frlin, k, make_float4(0, 0, 0, 0));
    }
}

// ContactManifold.Remove (FastRemoveAt) -> RemoveContact (ordered RemoveAt of the rows)
__device__ __forceinline__ void remove_contact(PairState& s, int index) {
    const int slot = get4(s.clist, index);
    s.ccount--;
    if (index < s.ccount) set4(s.clist, index, get4(s.clist, s.ccount));
    s.flags &= ~(16u << slot);
    int idx = -1;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (idx < 0 && i < s.npen && s.pslot[i] == slot) idx = i;
    if (idx < 0) return;
    // Ascending shifts, each reading an entry not yet overwritten.
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k >= idx && k < s.npen - 1) {
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

// The same two operations on a pair state updated in place in global memory
// (the general convex kernel): direct run-time indexing, one dependent access
// per entry instead of reading and writing all four.
__device__ inline void add_contact_inplace(PairState& s, f3 pos, f3 nrm, float depth, int id) {
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

__device__ inline void remove_contact_inplace(PairState& s, int index) {
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
