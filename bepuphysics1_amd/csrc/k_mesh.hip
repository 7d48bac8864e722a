// k_mesh.hip -- narrowphase of static mesh and terrain pairs:
// StaticMeshContactManifold and TerrainContactManifold over
// TriangleMeshConvexContactManifold.Update
// (CollisionTests/Manifolds/TriangleMeshConvexContactManifold.cs:101-434,
// StaticMeshContactManifold.cs:33-71, TerrainContactManifold.cs:30-100) for
// the pairs k_pair_prep listed.
//
// One wavefront per pair, its state in LDS: the contact refresh, the
// depth-first walk of the mesh's MeshBoundingBoxTree over the convex's box
// (child A's subtree before child B's, MeshBoundingBoxTree.cs:112-122,
// 214-226) or the scan of the terrain's grid under the convex's box in its
// frame, each overlapped
// triangle carried into the convex's local space, pruned by the convex's
// local box and tested by its TriangleConvexPairTester (kept per triangle in
// the pair's tester table) or the TriangleSpherePairTester, the improved
// boundary handling, the removal of stale testers and the contact reduction.
// The manifold's own transform is the identity: the triangles reach world
// space through the mesh shape's (or the terrain's) affine transform.
#include "body_shape.h"
#include "cast.h"
#include "manifold.h"
#include "mesh.h"

namespace bepu {

namespace {

const cvx::RT kMeshTransform{f3{0, 0, 0}, q4{0, 0, 0, 1}};
// CollisionDetectionSettings.nonconvexNormalDotMinimum (CollisionDetectionSettings.cs:52)
constexpr float kNonconvexNormalDotMinimum = .99f;

// An edge or vertex contact queued by AnalyzeCandidate
// (TriangleMeshConvexContactManifold.cs:773-787): the contact in the
// convex's local space, its corrected normal, the edge's (or vertex's) mesh
// vertex indices and whether the tester allows the correction.
struct Boundary {
    MeshContact c;
    f3 corrected;
    int e0, e1;
    bool should_correct;
};

// The per-step scratch of one pair's update.
struct MeshScratch {
    MeshContact cand[kMeshCandidates];
    Boundary edges[kMeshCandidates], verts[kMeshCandidates];
    int bverts[kMeshBlocked];
    int2 bedges[kMeshBlocked];
    int ncand, nedges, nverts, nbv, nbe;
    bool overflow;
};

__device__ __forceinline__ bool has_vertex(const MeshScratch& m, int v) {
    for (int k = 0; k < m.nbv; ++k)
        if (m.bverts[k] == v) return true;
    return false;
}
__device__ __forceinline__ bool has_edge(const MeshScratch& m, int a, int b) {
    for (int k = 0; k < m.nbe; ++k)
        if ((m.bedges[k].x == a && m.bedges[k].y == b) || (m.bedges[k].x == b && m.bedges[k].y == a)) return true;
    return false;
}
__device__ __forceinline__ void block_vertex(MeshScratch& m, int v) {
    if (has_vertex(m, v)) return;
    if (m.nbv < kMeshBlocked) m.bverts[m.nbv++] = v;
    else m.overflow = true;
}
__device__ __forceinline__ void block_edge(MeshScratch& m, int a, int b) {
    if (has_edge(m, a, b)) return;
    if (m.nbe < kMeshBlocked) m.bedges[m.nbe++] = make_int2(a, b);
    else m.overflow = true;
}

// GetNormal (TriangleMeshConvexContactManifold.cs:450-481)
__device__ __forceinline__ f3 corrected_normal(f3 uncorrected, const cvx::TriShape& t, int sided) {
    f3 ab = t.v1 - t.v0, ac = t.v2 - t.v0;
    f3 n;
    if (sided == SIDED_DOUBLE) {
        n = cross3(ab, ac);
        if (dot3(n, uncorrected) < 0) n = -n;
    } else if (sided == SIDED_CLOCKWISE) {
        n = cross3(ac, ab);
    } else {
        n = cross3(ab, ac);
    }
    if (lensq3(n) < 1e-7f) n = uncorrected;
    return n;
}

// AnalyzeCandidate (TriangleMeshConvexContactManifold.cs:483-596): true for a
// face contact, which is always kept; an edge or vertex contact is queued
// (with `corrected`, GetNormal of its triangle) and every other region of the
// triangle blocked.
__device__ __forceinline__ bool analyze(int region, const int ix[3], f3 corrected, bool should_correct, const MeshContact& c,
                        MeshScratch& m) {
    const int A = ix[0], B = ix[1], C = ix[2];
    if (region <= VR_C) {
        Boundary b{c, corrected, region == VR_A ? A : (region == VR_B ? B : C), -1, should_correct};
        if (m.nverts < kMeshCandidates) m.verts[m.nverts++] = b;
        else m.overflow = true;
        block_edge(m, A, B);
        block_edge(m, B, C);
        block_edge(m, A, C);
        if (region != VR_A) block_vertex(m, A);
        if (region != VR_B) block_vertex(m, B);
        if (region != VR_C) block_vertex(m, C);
        return false;
    }
    if (region <= VR_BC) {
        Boundary b{c, corrected, region == VR_BC ? B : A, region == VR_AB ? B : C, should_correct};
        if (m.nedges < kMeshCandidates) m.edges[m.nedges++] = b;
        else m.overflow = true;
        if (region != VR_AB) block_edge(m, A, B);
        if (region != VR_BC) block_edge(m, B, C);
        if (region != VR_AC) block_edge(m, A, C);
        block_vertex(m, A);
        block_vertex(m, B);
        block_vertex(m, C);
        return false;
    }
    block_edge(m, A, B);
    block_edge(m, B, C);
    block_edge(m, A, C);
    block_vertex(m, A);
    block_vertex(m, B);
    block_vertex(m, C);
    return true;
}

// StaticMeshContactManifold's Add (TriangleMeshConvexContactManifold.cs:598-608):
// the contact and its supplement data (local offsets on the convex and the
// mesh, base depth).
__device__ __forceinline__ void add_mesh_contact(PairState& s, ConvexState& e, const MeshContact& c, const cvx::RT& ta) {
    int slot = 0;
    while (s.flags & (16u << slot)) ++slot;
    e.offa[slot] = f4_of(cvx::rt_transform_by_inverse(c.pos, ta), c.depth);
    e.offb[slot] = f4_of(cvx::rt_transform_by_inverse(c.pos, kMeshTransform), 0.0f);
    add_contact_inplace(s, c.pos, c.normal, c.depth, c.id);
}

// AddLocalContact and IsContactUnique (TriangleMeshConvexContactManifold.cs:
// 436-447, 617-674): the candidate goes to world space; a contact of the
// manifold at the same place with a nearly parallel normal takes its data
// instead, a candidate already queued there drops it.
__device__ __forceinline__ void add_local(const DeviceWorld& w, PairState& s, ConvexState& e, MeshScratch& m, MeshContact c,
                          const m3& orientation, const cvx::RT& ta) {
    c.pos = mat_transform(c.pos, orientation) + ta.pos;
    c.normal = mat_transform(c.normal, orientation);
    for (int i = 0; i < s.ccount; i++) {
        const int slot = s.clist[i];
        if (distsq3(f3_of(s.cpos[slot]), c.pos) < w.s.contact_min_separation_sq) {
            float d = dot3(f3_of(s.cnrm[slot]), c.normal);
            if (fabsf(d) >= kNonconvexNormalDotMinimum) {
                s.cnrm[slot] = f4_of(c.normal, s.cnrm[slot].w);
                s.cpos[slot] = f4_of(c.pos, c.depth);
                e.offa[slot] = f4_of(cvx::rt_transform_by_inverse(c.pos, ta), c.depth);
                e.offb[slot] = f4_of(cvx::rt_transform_by_inverse(c.pos, kMeshTransform), 0.0f);
                return;
            }
        }
    }
    for (int i = 0; i < m.ncand; i++) {
        if (distsq3(m.cand[i].pos, c.pos) < w.s.contact_min_separation_sq) {
            float d = dot3(m.cand[i].normal, c.normal);
            if (fabsf(d) >= kNonconvexNormalDotMinimum) return;
        }
    }
    if (m.ncand < kMeshCandidates) m.cand[m.ncand++] = c;
    else m.overflow = true;
}

// ContactReducer.ReduceContacts, general form (CollisionTests/ContactReducer.cs:19-195):
// the manifold's contacts (list order) and the candidates; the list indices
// to remove (ascending, as a bit mask) and the candidates kept (bit mask).
__device__ __forceinline__ void reduce_general(const PairState& s, const MeshScratch& m, uint32_t* remove, uint32_t* keep) {
    const int nc = s.ccount;
    *remove = 0;
    *keep = 0;
    float maxDepth = -FLT_MAX;
    int deepest = -1;
    f3 normal{0, 0, 0};
    for (int i = 0; i < nc; i++) {
        const int slot = s.clist[i];
        normal = normal + f3_of(s.cnrm[slot]);
        if (s.cpos[slot].w > maxDepth) {
            deepest = i;
            maxDepth = s.cpos[slot].w;
        }
    }
    for (int i = 0; i < m.ncand; i++) {
        normal = normal + m.cand[i].normal;
        if (m.cand[i].depth > maxDepth) {
            deepest = nc + i;
            maxDepth = m.cand[i].depth;
        }
    }
    if (lensq3(normal) < 1e-7f) {
        if (nc > 0) normal = f3_of(s.cnrm[s.clist[0]]);
        else normal = m.cand[0].normal;
    }
    auto position = [&](int k) { return k < nc ? f3_of(s.cpos[s.clist[k]]) : m.cand[k - nc].pos; };
    const f3 deepPos = position(deepest);
    float furthestDist = 0;
    int furthest = -1;
    for (int k = 0; k < nc + m.ncand; k++) {
        float d = distsq3(position(k), deepPos);
        if (d > furthestDist) {
            furthestDist = d;
            furthest = k;
        }
    }
    if (furthest == -1) {
        if (nc > 0) {
            for (int i = 1; i < nc; i++) *remove |= 1u << i;
            return;
        }
        *keep = 1u;
        return;
    }
    const f3 xAxis = deepPos - position(furthest);
    const f3 yAxis = cross3(xAxis, normal);
    float minY = FLT_MAX, maxY = -FLT_MAX;
    int iminY = -1, imaxY = -1;
    for (int k = 0; k < nc + m.ncand; k++) {
        float d = dot3(position(k), yAxis);
        if (d < minY) {
            iminY = k;
            minY = d;
        }
        if (d > maxY) {
            imaxY = k;
            maxY = d;
        }
    }
    for (int k = 0; k < nc + m.ncand; k++) {
        const bool extreme = k == deepest || k == furthest || k == iminY || k == imaxY;
        if (k < nc && !extreme) *remove |= 1u << k;
        if (k >= nc && extreme) *keep |= 1u << (k - nc);
    }
}

// One triangle's tester results, kept for the in-order merge.
struct TriOut {
    MeshContact c[2];
    f3 corrected[2];
    int region[2];
    int ix[3];
    int n;               // contacts; -1: not tested (outside the local box, or no tester left)
    int should_correct;
};

// A pair's update in LDS: its state, tester state and tester table (loaded
// from and stored back to the pair slot), the manifold's scratch, the
// chunk's G triangles and their tester results, the tree walk's stack.
template <int G>
struct MeshShared {
    PairState s;
    ConvexState e;
    MeshTester tab[kMeshTesters];
    MeshScratch m;
    TriOut r[G];
    int tri[G];
    int stack[kMeshStack];
    int count;
};

// OR of a per-lane word over the lanes of a group of G.
template <int G>
__device__ __forceinline__ uint32_t group_or(uint32_t v) {
    for (int off = G / 2; off > 0; off >>= 1) v |= (uint32_t)__shfl_xor((int)v, off);
    return v;
}

// The update of one pair (TriangleMeshConvexContactManifold.Update,
// TriangleMeshConvexContactManifold.cs:101-434) by a group of G lanes of the
// wavefront (`gl` the lane in the group, `gb` the group's first lane; an
// inactive group has no pair and only keeps step with the wavefront's
// barriers).  Overlapped triangles are gathered in order, up to one per
// lane; each lane carries its triangle into the convex's local space, prunes
// it by the convex's local box and runs its pair tester (the
// TriangleConvexPairTester state from the pair's table, or the
// TriangleSpherePairTester); the group's first lane then merges the chunk's
// candidates in the order of the triangles (boundary analysis,
// AddLocalContact), so the result is the one the reference's serial loop
// gives.  Every barrier is reached by the whole wavefront.
template <int G>
__device__ __forceinline__ void mesh_pair(const DeviceWorld& w, MeshShared<G>& sh, int gl, int gb, bool active) {
    PairState& s = sh.s;
    ConvexState& e = sh.e;
    MeshScratch& m = sh.m;
    const uint32_t A = active ? s.a : 0u, M = active ? s.b : 0u;
    const float4 PA = active ? w.pos[A] : make_float4(0, 0, 0, 0);
    const float4 QA = active ? w.quat[A] : make_float4(0, 0, 0, 1);
    const cvx::RT ta{f3_of(PA), q4{QA.x, QA.y, QA.z, QA.w}};
    const bool sphere = active && s.kind == BEPU_PAIR_MESH_SPHERE;
    int ntest = active && !sphere ? s.tester : 0;
    if (gl == 0 && active) {
        // ContactRefresher.ContactRefresh (ContactRefresher.cs:18-69)
        int rem[4], nrem = 0;
        for (int k = 0; k < s.ccount; k++) {
            int slot = s.clist[k];
            f3 na = cvx::rt_transform(f3_of(e.offa[slot]), ta);
            f3 nb = cvx::rt_transform(f3_of(e.offb[slot]), kMeshTransform);
            f3 ab = nb - na;
            f3 n = f3_of(s.cnrm[slot]);
            float d = dot3(ab, n);
            f3 t = ab - n * d;
            d = lensq3(t);
            if (d > w.s.contact_invalidation_length_sq) {
                rem[nrem++] = k;
            } else {
                d = dot3(ab, n);
                float depth = e.offa[slot].w - d;
                s.cpos[slot].w = depth;
                if (depth < -w.s.max_contact_distance) rem[nrem++] = k;
                else s.cpos[slot] = f4_of((nb + na) * .5f, depth);
            }
        }
        for (int i = nrem - 1; i >= 0; i--) remove_contact_inplace(s, rem[i]);
        m.ncand = m.nedges = m.nverts = m.nbv = m.nbe = 0;
        m.overflow = false;
    }

    const float4 msh = active ? w.shape[M] : make_float4(0, 0, 0, 0);
    const bool terrain = active && msh.w == (float)BEPU_SHAPE_TERRAIN;
    const int sided = (int)msh.y;
    const bool improve = msh.z != 0.0f;
    // The shape's world transform: the mesh shape's, or the terrain's.
    m3 sm{1, 0, 0, 0, 1, 0, 0, 0, 1};
    f3 st{0, 0, 0};
    MeshInfo mi{};
    TerrainInfo ti{};
    if (terrain) {
        ti = terrain_info(w, __float_as_uint(msh.x));
        sm = ti.m;
        st = ti.t;
    } else if (active) {
        mi = mesh_info(w, __float_as_uint(msh.x));
        sm = mi.m;
        st = mi.t;
    }
    // RigidTransform.Invert (RigidTransform.cs:95-100), AffineTransform.CreateFromRigidTransform
    // (AffineTransform.cs:232-236), AffineTransform.Multiply (AffineTransform.cs:186-194).
    const q4 iq = qconj(ta.q);
    const f3 ip = -qrotate(ta.pos, iq);
    const m3 im = mat_from_quat(iq);
    const m3 lm = mat_mul(sm, im);
    const f3 lt = mat_transform(st, im) + ip;
    f3 lmin{0, 0, 0}, lmax{0, 0, 0};
    if (active) shape_aabb(w, w.shape[A], f3{0, 0, 0}, q4{0, 0, 0, 1}, &lmin, &lmax);
    const m3 orientation = mat_from_quat(ta.q);
    const float4 sha = active ? w.shape[A] : make_float4(0, 0, 0, (float)BEPU_SHAPE_SPHERE);
    TriConvex cv;
    cv.shape = shape_of(w, A, w.hull_verts + __float_as_uint(sha.y), 1);
    cv.min_radius = active ? w.minrad[A] : 0.0f;
    cv.core_scaling = w.s.core_shape_scaling;
    uint32_t updated = 0;
    int guaranteed = 0;
    const uint64_t below = ((1ull << gl) - 1) << gb;  // the group's lanes before this one

    // Tests the group's chunk of `ntri` triangles (sh.tri, in order; 0 when
    // the group has none this pass), one per lane, then merges their
    // candidates in order on the group's first lane.
    auto run_chunk = [&](int ntri) __attribute__((always_inline)) {
        TriOut& r = sh.r[gl];
        cvx::TriShape t;
        int ix[3] = {-1, -1, -1};
        bool pass = false;
        if (gl < ntri) {
            const int tri = sh.tri[gl];
            if (terrain) {
                // ConfigureLocalTriangle (TerrainContactManifold.cs:76-100):
                // TerrainVertexIndices.ToSequentialIndex for the indices.
                int col[3], row[3];
                terrain_local_indices(ti, tri, col, row);
                f3 v[3];
                for (int k = 0; k < 3; ++k) {
                    ix[k] = row[k] * (int)ti.width + col[k];
                    v[k] = affine_transform(f3{(float)col[k], terrain_height(w, ti, col[k], row[k]), (float)row[k]},
                                            lm, lt);
                }
                t.v0 = v[0];
                t.v1 = v[1];
                t.v2 = v[2];
            } else {
                // ConfigureLocalTriangle (StaticMeshContactManifold.cs:51-66): a
                // leaf holds the triangle's first index position.
                ix[0] = w.mesh_idx[mi.ibase + tri];
                ix[1] = w.mesh_idx[mi.ibase + tri + 1];
                ix[2] = w.mesh_idx[mi.ibase + tri + 2];
                t.v0 = affine_transform(f3_of(w.mesh_verts[mi.vbase + ix[0]]), lm, lt);
                t.v1 = affine_transform(f3_of(w.mesh_verts[mi.vbase + ix[1]]), lm, lt);
                t.v2 = affine_transform(f3_of(w.mesh_verts[mi.vbase + ix[2]]), lm, lt);
            }
            t.margin = 0;
            // Toolbox.GetTriangleBoundingBox against the convex's local box
            // (TriangleMeshConvexContactManifold.cs:143-150).
            f3 tmin, tmax;
            tri_axis(t.v0.x, t.v1.x, t.v2.x, &tmin.x, &tmax.x);
            tri_axis(t.v0.y, t.v1.y, t.v2.y, &tmin.y, &tmax.y);
            tri_axis(t.v0.z, t.v1.z, t.v2.z, &tmin.z, &tmax.z);
            pass = !(tmin.x > lmax.x || tmin.y > lmax.y || tmin.z > lmax.z) &&
                   !(lmin.x > tmax.x || lmin.y > tmax.y || lmin.z > tmax.z);
            r.ix[0] = ix[0];
            r.ix[1] = ix[1];
            r.ix[2] = ix[2];
        }
        MeshContact out[2];
        int n = -1, region_sphere = VR_ABC;
        bool should_correct = false;
        // The pair tester of each triangle (TriangleIndices), found in the
        // table or appended in triangle order.
        const bool tconv = pass && !sphere;
        int found = -1;
        if (tconv)
            for (int q = 0; q < ntest; ++q)
                if (sh.tab[q].ia == ix[0] && sh.tab[q].ib == ix[1] && sh.tab[q].ic == ix[2]) {
                    found = q;
                    break;
                }
        __syncthreads();  // the chunk's indices are in LDS
        // A triangle repeated within the chunk (the same TriangleIndices)
        // must meet the tester its earlier occurrence updated.
        bool dup = false;
        if (tconv)
            for (int q = 0; q < gl && !dup; ++q)
                dup = sh.r[q].ix[0] == ix[0] && sh.r[q].ix[1] == ix[1] && sh.r[q].ix[2] == ix[2];
        if (__ballot(dup) == 0) {
            const uint64_t fresh = __ballot(tconv && found < 0);
            int ts = found;
            if (tconv && found < 0) {
                const int rank = __popcll(fresh & below);
                if (ntest + rank < kMeshTesters) {
                    ts = ntest + rank;
                    sh.tab[ts] = MeshTester{ix[0], ix[1], ix[2], TT_PLANE, 0, 0.0f, 0.0f, 0.0f};
                } else {
                    m.overflow = true;  // the table is full: this triangle is not tested
                }
            }
            const uint64_t gmask = (G == 64 ? ~0ull : ((1ull << G) - 1)) << gb;
            ntest = min(ntest + __popcll(fresh & gmask), kMeshTesters);
            updated |= group_or<G>(ts >= 0 ? 1u << ts : 0u);
            if (ts >= 0) {
                MeshTester tst = sh.tab[ts];
                n = triangle_convex(cv, t, sided, tst, out);
                sh.tab[ts] = tst;
                should_correct = tst.state == TT_DEEP;
            }
        } else {
            // With a repeat in any group of the wavefront, every group takes
            // its chunk's triangles one at a time, in order.
            for (int q = 0; q < G; ++q) {
                int ts = -1, nt = ntest;
                if (gl == q && tconv) {
                    for (int z = 0; z < ntest; ++z)
                        if (sh.tab[z].ia == ix[0] && sh.tab[z].ib == ix[1] && sh.tab[z].ic == ix[2]) {
                            ts = z;
                            break;
                        }
                    if (ts < 0) {
                        if (ntest < kMeshTesters) {
                            ts = ntest;
                            nt = ntest + 1;
                            sh.tab[ts] = MeshTester{ix[0], ix[1], ix[2], TT_PLANE, 0, 0.0f, 0.0f, 0.0f};
                        } else {
                            m.overflow = true;
                        }
                    }
                    if (ts >= 0) {
                        MeshTester tst = sh.tab[ts];
                        n = triangle_convex(cv, t, sided, tst, out);
                        sh.tab[ts] = tst;
                        should_correct = tst.state == TT_DEEP;
                    }
                }
                ntest = __shfl(nt, gb + q);
                ts = __shfl(ts, gb + q);
                if (ts >= 0) updated |= 1u << ts;
                __syncthreads();
            }
        }
        if (sphere && pass) n = triangle_sphere(cv.shape.margin, t, sided, &out[0], &region_sphere);
        if (gl < ntri) {
            r.n = n;
            r.should_correct = should_correct ? 1 : 0;
            for (int j = 0; j < n; ++j) {
                r.c[j] = out[j];
                if (improve) {
                    r.region[j] = sphere ? region_sphere : triangle_convex_region(t, out[j].normal);
                    r.corrected[j] = corrected_normal(out[j].normal, t, sided);
                }
            }
        }
        __syncthreads();
        if (gl == 0) {
            for (int q = 0; q < ntri; ++q) {
                const int nq = sh.r[q].n;
                for (int j = 0; j < nq; j++) {
                    const MeshContact c = sh.r[q].c[j];
                    if (improve) {
                        const int qix[3] = {sh.r[q].ix[0], sh.r[q].ix[1], sh.r[q].ix[2]};
                        if (analyze(sh.r[q].region[j], qix, sh.r[q].corrected[j], sh.r[q].should_correct != 0, c,
                                    m)) {
                            guaranteed++;
                            add_local(w, s, e, m, c, orientation, ta);
                        }
                    } else {
                        add_local(w, s, e, m, c, orientation, ta);
                    }
                }
            }
        }
        __syncthreads();
    };

    // Terrain pairs: TerrainContactManifold.FindOverlappingTriangles
    // (TerrainContactManifold.cs:30-64) and TerrainShape.GetOverlaps
    // (TerrainShape.cs:497-543).
    int minX = 0, minY = 0, ny = 1, nquads = 0, width = 0;
    f3 qmin{0, 0, 0}, qmax{0, 0, 0};
    if (terrain) {
        // ConvexShape.GetLocalBoundingBox (ConvexShape.cs:225-281) of the
        // convex in the terrain's frame (AffineTransform.Invert,
        // AffineTransform.cs:173-178; Multiply(RigidTransform,
        // AffineTransform), :202-211; EntityShape.TransformLocalExtremePoints,
        // EntityShape.cs:58-63), grown by its velocity over the step.
        const m3 inv = mat_invert(ti.m);
        const f3 inv_t = -mat_transform(ti.t, inv);
        const m3 bm = mat_mul(orientation, inv);
        const f3 bt = mat_transform(ta.pos, inv) + inv_t;
        const f3 right = cvx::local_extreme_with_margin(cv.shape, f3{bm.m11, bm.m21, bm.m31});
        const f3 left = cvx::local_extreme_with_margin(cv.shape, f3{-bm.m11, -bm.m21, -bm.m31});
        const f3 up = cvx::local_extreme_with_margin(cv.shape, f3{bm.m12, bm.m22, bm.m32});
        const f3 down = cvx::local_extreme_with_margin(cv.shape, f3{-bm.m12, -bm.m22, -bm.m32});
        const f3 backward = cvx::local_extreme_with_margin(cv.shape, f3{bm.m13, bm.m23, bm.m33});
        const f3 forward = cvx::local_extreme_with_margin(cv.shape, f3{-bm.m13, -bm.m23, -bm.m33});
        const f3 pos{right.x * bm.m11 + right.y * bm.m21 + right.z * bm.m31,
                     up.x * bm.m12 + up.y * bm.m22 + up.z * bm.m32,
                     backward.x * bm.m13 + backward.y * bm.m23 + backward.z * bm.m33};
        const f3 neg{left.x * bm.m11 + left.y * bm.m21 + left.z * bm.m31,
                     down.x * bm.m12 + down.y * bm.m22 + down.z * bm.m32,
                     forward.x * bm.m13 + forward.y * bm.m23 + forward.z * bm.m33};
        qmax = f3{bt.x + pos.x, bt.y + pos.y, bt.z + pos.z};
        qmin = f3{bt.x + neg.x, bt.y + neg.y, bt.z + neg.z};
        f3 tv = mat_transform(f3_of(w.linv[A]), mat_invert(ti.m));
        tv = tv * w.s.dt;
        if (tv.x > 0) qmax.x += tv.x; else qmin.x += tv.x;
        if (tv.y > 0) qmax.y += tv.y; else qmin.y += tv.y;
        if (tv.z > 0) qmax.z += tv.z; else qmin.z += tv.z;
        if (gl == 0) {
            // Kept for the time of impact, which walks the same triangles.
            e.sep = f4_of(qmin, 1.0f);
            e.dir = f4_of(qmax, 0.0f);
        }
        width = (int)ti.width;
        minX = max((int)qmin.x, 0);
        minY = max((int)qmin.z, 0);
        const int maxX = min((int)qmax.x, width - 2), maxY = min((int)qmax.z, (int)ti.length - 2);
        ny = maxY - minY + 1;
        nquads = (maxX >= minX && maxY >= minY) ? (maxX - minX + 1) * ny : 0;
    }
    // Static mesh pairs: MeshBoundingBoxTree.GetOverlaps over the convex's
    // box, depth first (child A's subtree before child B's), walked by the
    // group's first lane up to a chunk of leaves at a time.
    const bool mesh = active && !terrain;
    const f3 bmin = mesh ? f3_of(w.bmin[A]) : f3{0, 0, 0}, bmax = mesh ? f3_of(w.bmax[A]) : f3{0, 0, 0};
    const float4* nodes = w.mesh_nodes + 2 * (size_t)mi.nbase;
    auto overlaps = [&](int node) {
        const float4 lo = nodes[2 * node], hi = nodes[2 * node + 1];
        if (bmin.x > hi.x || bmin.y > hi.y || bmin.z > hi.z) return false;
        if (lo.x > bmax.x || lo.y > bmax.y || lo.z > bmax.z) return false;
        return true;
    };
    int top = 0;
    if (gl == 0 && mesh && mi.ntri > 0 && overlaps(0)) sh.stack[top++] = 0;
    const uint64_t gmask = (G == 64 ? ~0ull : ((1ull << G) - 1)) << gb;
    // One pass per chunk: a terrain's next G / 2 covered cells (columns
    // outer, rows inner) give the triangles of those quads whose heights
    // reach the box; a mesh's walk gives its next G leaves.
    for (int q0 = 0;; q0 += G / 2) {
        bool hit = false;
        int quad = 0;
        const int qq = q0 + gl;
        if (terrain && gl < G / 2 && qq < nquads) {
            const int i = minX + qq / ny, j = minY + qq % ny;
            const float y1 = terrain_height(w, ti, i, j), y2 = terrain_height(w, ti, i + 1, j),
                        y3 = terrain_height(w, ti, i, j + 1), y4 = terrain_height(w, ti, i + 1, j + 1);
            float highest = y1, lowest = y1;
            if (y2 > highest) highest = y2; else if (y2 < lowest) lowest = y2;
            if (y3 > highest) highest = y3; else if (y3 < lowest) lowest = y3;
            if (y4 > highest) highest = y4; else if (y4 < lowest) lowest = y4;
            hit = !(qmax.y < lowest || qmin.y > highest);
            quad = (i + j * width) << 1;
        }
        const uint64_t hits = __ballot(hit);
        if (hit) {
            const int rank = __popcll(hits & below);
            sh.tri[2 * rank] = quad;
            sh.tri[2 * rank + 1] = quad | 1;
        }
        if (gl == 0) {
            int cnt = 0;
            while (top > 0 && cnt < G) {
                const int node = sh.stack[--top];
                const int a = __float_as_int(nodes[2 * node].w), b = __float_as_int(nodes[2 * node + 1].w);
                if (a >= 0) {
                    if (overlaps(b)) {
                        if (top < kMeshStack) sh.stack[top++] = b;
                        else m.overflow = true;
                    }
                    if (overlaps(a)) {
                        if (top < kMeshStack) sh.stack[top++] = a;
                        else m.overflow = true;
                    }
                    continue;
                }
                sh.tri[cnt++] = b;
            }
            sh.count = terrain ? 2 * __popcll(hits & gmask) : cnt;
        }
        const bool more = terrain ? q0 + G / 2 < nquads : __shfl(top, gb) > 0;
        __syncthreads();
        const int cnt = sh.count;
        __syncthreads();
        if (__ballot(cnt > 0 || more) == 0) break;
        run_chunk(cnt);
    }

    if (gl != 0 || !active) return;
    if (improve) {
        // Edge contacts fighting each other in one plane get the triangle
        // normal (TriangleMeshConvexContactManifold.cs:241-316).
        if (m.nverts == 0 && guaranteed == 0 && m.nedges > 1) {
            bool same_plane = true, against = false;
            const f3 first = m.edges[0].c.normal;
            m.edges[0].corrected = normalize3(m.edges[0].corrected);
            float d = dot3(first, m.edges[0].corrected);
            if (fabsf(d) > .01f) {
                same_plane = false;
            } else {
                for (int i = 1; i < m.nedges; i++) {
                    d = dot3(m.edges[i].c.normal, first);
                    if (d < 0) against = true;
                    d = dot3(m.edges[i].c.normal, m.edges[0].corrected);
                    if (fabsf(d) > .01f) {
                        same_plane = false;
                        break;
                    }
                }
            }
            if (same_plane && against) {
                m.edges[0].c.normal = m.edges[0].corrected;
                m.edges[0].should_correct = true;
                for (int i = 1; i < m.nedges; i++) {
                    m.edges[i].corrected = normalize3(m.edges[i].corrected);
                    d = dot3(m.edges[i].corrected, m.edges[i].c.normal);
                    if ((double)d < .01) {
                        m.edges[i].c.normal = m.edges[i].corrected;
                        m.edges[i].should_correct = true;
                    }
                }
            }
        }
        // Unblocked boundary contacts as they are; blocked ones with their
        // normal corrected when allowed (TriangleMeshConvexContactManifold.cs:322-376).
        for (int pass = 0; pass < 2; ++pass) {
            const int nb = pass == 0 ? m.nedges : m.nverts;
            for (int i = 0; i < nb; ++i) {
                Boundary& bc = pass == 0 ? m.edges[i] : m.verts[i];
                const bool blocked = pass == 0 ? has_edge(m, bc.e0, bc.e1) : has_vertex(m, bc.e0);
                if (!blocked) {
                    add_local(w, s, e, m, bc.c, orientation, ta);
                } else if (bc.should_correct || guaranteed == 0) {
                    bc.corrected = normalize3(bc.corrected);
                    float d = dot3(bc.corrected, bc.c.normal);
                    bc.c.normal = bc.corrected;
                    bc.c.depth *= maxf_ref(0, d);
                    add_local(w, s, e, m, bc.c, orientation, ta);
                }
            }
        }
    }

    // Stale pair testers leave the table, the last entry taking a removed
    // one's place (TriangleMeshConvexContactManifold.cs:387-401).
    for (int i = ntest - 1; i >= 0; --i) {
        if (!(updated & (1u << i))) {
            sh.tab[i] = sh.tab[ntest - 1];
            if (updated & (1u << (ntest - 1))) updated |= 1u << i;
            --ntest;
        }
    }
    if (!sphere) s.tester = ntest;

    if (s.ccount + m.ncand > 4) {
        uint32_t remove, keep;
        reduce_general(s, m, &remove, &keep);
        for (int i = 3; i >= 0; i--)
            if (remove & (1u << i)) remove_contact_inplace(s, i);
        for (int i = m.ncand - 1; i >= 0; i--)
            if (keep & (1u << i)) add_mesh_contact(s, e, m.cand[i], ta);
    } else {
        for (int i = 0; i < m.ncand; i++) add_mesh_contact(s, e, m.cand[i], ta);
    }
    if (m.overflow) w.counters[ctr(CTR_OVERFLOW)] = 5;
}

static_assert(sizeof(PairState) % 16 == 0 && sizeof(ConvexState) % 16 == 0 && sizeof(MeshTester) % 16 == 0,
              "pair records are copied as float4s");

// Lanes per pair: four pairs per wavefront.
constexpr int kMeshGroup = 16;

// One group of kMeshGroup lanes per listed pair, 64 / kMeshGroup pairs per
// single-wavefront workgroup (grid-stride over the list).
__global__ void __launch_bounds__(64) k_mesh(DeviceWorld w) {
    constexpr int G = kMeshGroup, P = 64 / G;
    __shared__ MeshShared<G> shared[P];
    const int lane = threadIdx.x, g = lane / G, gl = lane % G, gb = g * G;
    MeshShared<G>& sh = shared[g];
    const uint32_t n = w.counters[ctr(CTR_MESH)];
    constexpr int kPair4 = sizeof(PairState) / 16, kConv4 = sizeof(ConvexState) / 16,
                  kTester4 = sizeof(MeshTester) / 16;
    for (uint32_t base = blockIdx.x * P; base < n; base += gridDim.x * P) {
        const bool active = base + g < n;
        const uint2 wk = active ? w.mesh_work[base + g] : make_uint2(0, 0);
        float4* pd = reinterpret_cast<float4*>(&sh.s);
        float4* cd = reinterpret_cast<float4*>(&sh.e);
        float4* td = reinterpret_cast<float4*>(sh.tab);
        if (active) {
            // k_pair_prep carried last step's pair state, supplement data and
            // tester table over; a new pair starts from zero.
            const float4* ps = reinterpret_cast<const float4*>(&w.pairs_cur[wk.x]);
            for (int i = gl; i < kPair4; i += G) pd[i] = ps[i];
            const float4* cs = reinterpret_cast<const float4*>(&w.conv_cur[wk.x]);
            for (int i = gl; i < kConv4; i += G) cd[i] = wk.y == ~0u ? make_float4(0, 0, 0, 0) : cs[i];
            const int nt = w.pairs_cur[wk.x].tester * kTester4;
            const float4* ts = reinterpret_cast<const float4*>(w.mtest_cur + (size_t)wk.x * kMeshTesters);
            for (int i = gl; i < nt; i += G) td[i] = ts[i];
        }
        __syncthreads();
        mesh_pair<G>(w, sh, gl, gb, active);
        __syncthreads();
        if (active) {
            float4* pw = reinterpret_cast<float4*>(&w.pairs_cur[wk.x]);
            for (int i = gl; i < kPair4; i += G) pw[i] = pd[i];
            float4* cw = reinterpret_cast<float4*>(&w.conv_cur[wk.x]);
            for (int i = gl; i < kConv4; i += G) cw[i] = cd[i];
            const int nw = sh.s.tester * kTester4;
            float4* tw = reinterpret_cast<float4*>(w.mtest_cur + (size_t)wk.x * kMeshTesters);
            for (int i = gl; i < nw; i += G) tw[i] = td[i];
            if (gl == 0) {
                w.hot_cur[wk.x] = hot_of(sh.s);
                if (!hash_insert(w.hash_keys_cur, w.hash_vals_cur, w.hash_cap, w.pair_keys[wk.x], wk.x))
                    w.counters[ctr(CTR_OVERFLOW)] = 3;
            }
        }
        // The mesh is not dynamic: a mesh pair joins no islands.
        __syncthreads();
    }
}

}  // namespace

void launch_mesh(const DeviceWorld& w, hipStream_t st) {
    if (!w.meshes || w.pair_cap == 0) return;
    constexpr uint32_t P = 64 / kMeshGroup;
    const uint32_t grid = std::min<uint32_t>(16384, (w.pair_cap + P - 1) / P);
    hipLaunchKernelGGL(k_mesh, dim3(grid), dim3(64), 0, st, w);
}

}  // namespace bepu
