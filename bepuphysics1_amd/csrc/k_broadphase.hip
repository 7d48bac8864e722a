// k_broadphase.hip -- broadphase stage: dispatch between the two broadphases (the reference swaps its
// broadphase through the Space.BroadPhase setter, Space.cs:110-134).
//   BEPU_BROADPHASE_GRID  hierarchical hash grid (k_grid.hip), the default;
//   BEPU_BROADPHASE_TREE  implicit tree over Morton order (k_bvh.hip).
// Both produce the same pair set.
#include <hipcub/hipcub.hpp>

#include "world.h"

namespace bepu {

void launch_morton(const DeviceWorld& w, hipStream_t st);
void launch_bvh_after_sort(const DeviceWorld& w, hipStream_t st);
void launch_grid(const DeviceWorld& w, hipStream_t st);

size_t broadphase_sort_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 30);
    return bytes;
}

void launch_broadphase(const DeviceWorld& w, void* temp, size_t temp_bytes, hipStream_t st) {
    uint32_t n = w.nbodies;
    if (n < 2) return;
    size_t bytes = temp_bytes;
    if (w.bp_kind == BEPU_BROADPHASE_TREE) {
        launch_morton(w, st);
        (void)hipcub::DeviceRadixSort::SortPairs(temp, bytes, w.morton, w.morton_sorted, w.leaf_ids_tmp, w.leaf_ids,
                                                 (int)n, 0, 30, st);
        launch_bvh_after_sort(w, st);
    } else {
        launch_grid(w, st);
    }
}

}  // namespace bepu
