// mesh_tree.h -- host-side construction of a static mesh shape's bounding
// box tree (one-time per mesh shape).
#pragma once
#include <stdint.h>

#include <vector>

#include "dmath.h"

namespace bepu {

// One tree node: its box and, for an internal node, its children a and b;
// a leaf has a = -1 and b = its triangle (the position of its first index).
// Node 0 is the root.
struct MeshTreeNode {
    f3 bmin, bmax;
    int32_t a, b;
};

// MeshBoundingBoxTree.Reconstruct (DataStructures/MeshBoundingBoxTree.cs:60-69)
// over the world-space triangles of TransformableMeshData (TransformableMeshData.cs:
// 67-72): the triangles are inserted in the order (982451653 i) mod n, each
// walking down from the root to the child whose box volume grows least (B
// on ties) and pairing with the leaf it reaches.  `linear` is the world
// transform's Matrix3x3 (M11..M33, row-major), `translation` its offset.
// Returns the nodes and the tree's depth (nodes on the longest root-leaf path).
void build_mesh_tree(const f3* verts, const int32_t* idx, uint32_t nidx, const m3& linear, f3 translation,
                     std::vector<MeshTreeNode>* nodes, uint32_t* depth);

}  // namespace bepu
