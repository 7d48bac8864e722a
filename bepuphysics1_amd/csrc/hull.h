// hull.h -- host-side convex hull shape construction (one-time per hull shape).
#pragma once
#include <vector>

#include "dmath.h"

namespace bepu {

struct HullDescription {
    std::vector<f3> vertices;  // surface vertices, recentred on the volume centroid
    m3 volume_distribution;    // about the centroid
    f3 center;                 // volume centroid of the input points' hull
    float volume;
    float minimum_radius;      // InertiaHelper.ComputeMinimumRadius about the centroid (no margin)
};

// ConvexHullShape.ComputeDescription (ConvexHullShape.cs:154-175): the hull
// of `points` (ConvexHullHelper.GetConvexHull), its volume distribution and
// centroid (InertiaHelper.ComputeShapeDistribution), and the unique surface
// vertices recentred on the centroid.  Returns false for a point set without
// volume.
bool compute_hull_description(const f3* points, uint32_t n, HullDescription* out);

}  // namespace bepu
