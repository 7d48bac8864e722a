// mesh_tree.cpp -- MeshBoundingBoxTree construction on the host.  Every box
// is computed with the reference's operations (BoundingBox.CreateMerged,
// the leaf margin, the volume products) so the device walks the tree the
// reference builds.
#include "mesh_tree.h"

#include <algorithm>

namespace bepu {

namespace {

// MeshBoundingBoxTree.LeafMargin (MeshBoundingBoxTree.cs:385)
constexpr float kLeafMargin = .001f;

struct Box {
    f3 mn, mx;
};

// BoundingBox.CreateMerged (BEPUutilities/BoundingBox.cs:187-214)
Box merged(const Box& a, const Box& b) {
    Box r;
    r.mn.x = a.mn.x < b.mn.x ? a.mn.x : b.mn.x;
    r.mn.y = a.mn.y < b.mn.y ? a.mn.y : b.mn.y;
    r.mn.z = a.mn.z < b.mn.z ? a.mn.z : b.mn.z;
    r.mx.x = a.mx.x > b.mx.x ? a.mx.x : b.mx.x;
    r.mx.y = a.mx.y > b.mx.y ? a.mx.y : b.mx.y;
    r.mx.z = a.mx.z > b.mx.z ? a.mx.z : b.mx.z;
    return r;
}

float volume(const Box& b) {
    f3 o = b.mx - b.mn;
    return o.x * o.y * o.z;
}

}  // namespace

void build_mesh_tree(const f3* verts, const int32_t* idx, uint32_t nidx, const m3& linear, f3 translation,
                     std::vector<MeshTreeNode>* nodes, uint32_t* depth) {
    nodes->clear();
    *depth = 0;
    const int64_t ntri = (int64_t)(nidx / 3);
    if (ntri == 0) return;
    nodes->reserve((size_t)(2 * ntri));
    std::vector<Box> box;
    box.reserve((size_t)(2 * ntri));
    auto add = [&](const Box& b, int32_t a, int32_t c) {
        nodes->push_back(MeshTreeNode{b.mn, b.mx, a, c});
        box.push_back(b);
        return (int32_t)nodes->size() - 1;
    };
    int32_t root = -1;
    for (int64_t i = 0; i < (int64_t)nidx; i += 3) {
        const int32_t tri = (int32_t)(((982451653LL * (i / 3)) % ntri) * 3);
        // LeafNode: MeshBoundingBoxTreeData.GetBoundingBox of the world-space
        // triangle (AffineTransform.Transform of each vertex), grown by the margin.
        f3 v[3];
        for (int k = 0; k < 3; ++k) v[k] = mat_transform(verts[idx[tri + k]], linear) + translation;
        Box lb;
        lb.mn.x = v[0].x < v[1].x ? v[0].x : v[1].x;
        lb.mn.y = v[0].y < v[1].y ? v[0].y : v[1].y;
        lb.mn.z = v[0].z < v[1].z ? v[0].z : v[1].z;
        lb.mn.x = lb.mn.x < v[2].x ? lb.mn.x : v[2].x;
        lb.mn.y = lb.mn.y < v[2].y ? lb.mn.y : v[2].y;
        lb.mn.z = lb.mn.z < v[2].z ? lb.mn.z : v[2].z;
        lb.mx.x = v[0].x > v[1].x ? v[0].x : v[1].x;
        lb.mx.y = v[0].y > v[1].y ? v[0].y : v[1].y;
        lb.mx.z = v[0].z > v[1].z ? v[0].z : v[1].z;
        lb.mx.x = lb.mx.x > v[2].x ? lb.mx.x : v[2].x;
        lb.mx.y = lb.mx.y > v[2].y ? lb.mx.y : v[2].y;
        lb.mx.z = lb.mx.z > v[2].z ? lb.mx.z : v[2].z;
        lb.mx.x += kLeafMargin;
        lb.mx.y += kLeafMargin;
        lb.mx.z += kLeafMargin;
        lb.mn.x -= kLeafMargin;
        lb.mn.y -= kLeafMargin;
        lb.mn.z -= kLeafMargin;
        const int32_t leaf = add(lb, -1, tri);
        if (root < 0) {
            root = leaf;
            continue;
        }
        if ((*nodes)[root].a < 0) {
            root = add(merged(box[root], lb), root, leaf);
            continue;
        }
        box[root] = merged(lb, box[root]);
        int32_t node = root;
        while (true) {
            const int32_t ca = (*nodes)[node].a, cb = (*nodes)[node].b;
            const Box mA = merged(box[ca], lb), mB = merged(box[cb], lb);
            if (volume(mA) - volume(box[ca]) < volume(mB) - volume(box[cb])) {
                if ((*nodes)[ca].a < 0) {
                    const int32_t in = add(mA, ca, leaf);
                    (*nodes)[node].a = in;
                    break;
                }
                box[ca] = mA;
                node = ca;
            } else {
                if ((*nodes)[cb].a < 0) {
                    const int32_t in = add(mB, leaf, cb);
                    (*nodes)[node].b = in;
                    break;
                }
                box[cb] = mB;
                node = cb;
            }
        }
    }
    for (size_t k = 0; k < nodes->size(); ++k) {
        (*nodes)[k].bmin = box[k].mn;
        (*nodes)[k].bmax = box[k].mx;
    }
    // The root becomes node 0.
    if (root != 0) {
        std::swap((*nodes)[0], (*nodes)[root]);
        for (MeshTreeNode& n : *nodes)
            if (n.a >= 0) {
                if (n.a == 0) n.a = root;
                else if (n.a == root) n.a = 0;
                if (n.b == 0) n.b = root;
                else if (n.b == root) n.b = 0;
            }
    }
    // Depth: nodes on the longest root-to-leaf path.
    std::vector<std::pair<int32_t, uint32_t>> stack{{0, 1u}};
    while (!stack.empty()) {
        auto [k, d] = stack.back();
        stack.pop_back();
        *depth = std::max(*depth, d);
        if ((*nodes)[k].a >= 0) {
            stack.push_back({(*nodes)[k].a, d + 1});
            stack.push_back({(*nodes)[k].b, d + 1});
        }
    }
}

}  // namespace bepu
