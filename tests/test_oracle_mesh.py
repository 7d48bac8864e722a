"""Pins the oracle's static mesh path (StaticMeshContactManifold over
TriangleMeshConvexContactManifold, TriangleConvexPairTester,
TriangleSpherePairTester, MeshBoundingBoxTree) against what the geometry
fixes: the tree's overlap sets against brute force and a hand-derived
five-triangle build, closest points on a triangle against a dense search,
plane contacts of a box over a triangle (depth, position and normal in
closed form, posed and rotated), deep contacts, Voronoi regions of contact
normals, sphere-triangle contacts, the escape counter, one-sided meshes, and
whole simulations on a mesh ground against the same bodies on a box ground.
The reference ships no golden vectors for this path.  CPU only.
"""
import math

import numpy as np
import pytest

from bepuphysics1_amd import scenes

BOX = {"kind": 0, "half": [0.5, 0.5, 0.5], "margin": 0.04}
VR_A, VR_B, VR_C, VR_AB, VR_AC, VR_BC, VR_ABC = range(7)


def rot(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def brute_overlaps(verts, idx, xf, lo, hi):
    m, t = xf if xf is not None else (np.eye(3), np.zeros(3))
    out = set()
    for k in range(0, len(idx), 3):
        # AffineTransform.Transform with Matrix3x3 rows: p' = p M + t
        tri = np.array([verts[idx[k + j]] @ np.asarray(m, np.float64) + t for j in range(3)])
        mn, mx = tri.min(0) - 0.001, tri.max(0) + 0.001
        if np.all(mn <= hi + 1e-5) and np.all(lo <= mx + 1e-5):
            if np.all(mn <= hi - 1e-5) and np.all(lo <= mx - 1e-5):
                out.add(k)
            else:
                out.add(-1 - k)  # on the boundary: either answer is fine
    return out


def test_tree_overlaps_equal_brute_force(oracle):
    rng = np.random.default_rng(3)
    V, I, _ = scenes.heightfield_mesh(cells=12, size=24.0, amplitude=2.0, seed=5)
    R = rot([0.3, 1.0, -0.2], 0.7)
    xf = (R.astype(np.float32), np.array([1.5, -2.0, 0.5], np.float32))
    w = oracle.World()
    sid = w.create_mesh_shape(V, I, xf)
    boxes, links = w.mesh_tree(sid)
    leaves = sorted(int(b) for a, b in links if a < 0)
    assert leaves == list(range(0, len(I), 3))  # every triangle once
    assert len(links) == 2 * (len(I) // 3) - 1
    for _ in range(60):
        c = rng.uniform(-14, 14, 3)
        h = rng.uniform(0.2, 4.0, 3)
        got = w.mesh_overlaps(sid, c - h, c + h)
        assert len(got) == len(set(got.tolist()))
        want = brute_overlaps(V, I, (R, xf[1]), c - h, c + h)
        sure = {k for k in want if k >= 0}
        maybe = {-1 - k for k in want if k < 0}
        assert sure <= set(got.tolist()) <= sure | maybe


def test_tree_build_and_order_hand_derived(oracle):
    # Five unit triangles at x = 0, 10, 18, 30, 40 (triangle k at index 3k).
    # MeshBoundingBoxTree.Reconstruct inserts (982451653 * i) mod 5 = 0, 3, 1,
    # 4, 2; each insertion walks to the child whose volume grows least.  By
    # hand: {{T0, {T2, T1}}, {T4, T3}}, so a query over everything visits
    # 0, 2, 1, 4, 3 (first-index positions 0, 6, 3, 12, 9).
    xs = [0.0, 10.0, 18.0, 30.0, 40.0]
    V, I = [], []
    for k, x in enumerate(xs):
        V += [[x, 0, 0], [x + 1, 0, 0], [x, 0.5, 1]]
        I += [3 * k, 3 * k + 1, 3 * k + 2]
    w = oracle.World()
    sid = w.create_mesh_shape(np.array(V, np.float32), np.array(I, np.int32))
    got = w.mesh_overlaps(sid, [-5, -5, -5], [50, 5, 5])
    assert got.tolist() == [0, 6, 3, 12, 9]
    # A box over triangles 1 and 2 only.
    assert w.mesh_overlaps(sid, [9, -1, -1], [19.5, 1, 2]).tolist() == [6, 3]
    boxes, links = w.mesh_tree(sid)
    # The root box spans every triangle plus the leaf margin.
    np.testing.assert_allclose(boxes[0], [-0.001, -0.001, -0.001, 41.001, 0.501, 1.001], atol=1e-6)


def test_closest_point_on_triangle(oracle):
    rng = np.random.default_rng(1)
    uv = rng.uniform(size=(20000, 2))
    uv = np.where(uv.sum(1, keepdims=True) > 1, 1 - uv, uv)
    for _ in range(40):
        tri = rng.normal(size=(3, 3)).astype(np.float32)
        p = rng.normal(size=3).astype(np.float32) * 2
        region, q = oracle.closest_point_triangle(tri, p)
        pts = tri[0] + uv[:, :1] * (tri[1] - tri[0]) + uv[:, 1:] * (tri[2] - tri[0])
        best = np.min(np.linalg.norm(pts - p, axis=1))
        assert abs(np.linalg.norm(q - p) - best) < 2e-2
        # q lies on the triangle: barycentric coordinates in [0, 1]
        M = np.stack([tri[1] - tri[0], tri[2] - tri[0]], 1)
        bc, *_ = np.linalg.lstsq(M, q - tri[0], rcond=None)
        assert bc.min() > -1e-4 and bc.sum() < 1 + 1e-4
        # the region names the feature q lies on
        on_vertex = [np.linalg.norm(q - tri[k]) < 1e-5 for k in range(3)]
        if region in (VR_A, VR_B, VR_C):
            assert on_vertex[region]
        elif region == VR_ABC:
            assert bc.min() > -1e-5


@pytest.mark.parametrize("angle", [0.0, 0.4, 1.3])
@pytest.mark.parametrize("Y", [0.49, 0.46, 0.41])
def test_plane_contact_closed_form(oracle, Y, angle):
    # A unit box (margin 0.04) at the origin over a large triangle in the plane
    # y = -Y, both turned by R: the plane test's contact is the projection of
    # the box core's extreme point onto the plane, depth 0.5 - Y, normal along
    # the plane normal pointing away from the box.
    R = rot([0.2, 0.1, 1.0], angle)
    tri = np.array([[-6, -Y, -6], [6, -Y, -6], [0, -Y, 6]]) @ R.T
    n = np.array([0, -1.0, 0]) @ R.T
    t = oracle.TriangleTester()
    box = dict(BOX)
    shape = box
    # rotate the box with the frame: a box's support is axis-aligned, so turn
    # the triangle only when the box is axis-aligned (angle 0) ...
    if angle != 0.0:
        # ... else use a hull of the box's corners, rotated by R
        corners = np.array([[x, y, z] for x in (-.46, .46) for y in (-.46, .46) for z in (-.46, .46)]) @ R.T
        shape = {"kind": 2, "verts": corners.astype(np.float32), "margin": 0.04}
    c = t.candidates(shape, tri.astype(np.float32), 0, min_radius=0.5)
    assert t.state == 0 and len(c) == 1
    assert abs(c[0, 6] - (0.5 - Y)) < 1e-5
    np.testing.assert_allclose(c[0, 3:6], n, atol=2e-6)
    # the position lies on the plane below a corner or face point of the core
    assert abs(np.dot(c[0, :3], n) - Y) < 1e-5


@pytest.mark.parametrize("order", [(0, 1, 2), (2, 0, 1), (0, 2, 1), (1, 2, 0)])
def test_plane_test_leaves_for_points_outside_each_edge(oracle, order):
    # The box core's extreme point (its bottom centre) projects just outside
    # the edge PQ of a triangle the box overlaps: the plane test hands over
    # to the GJK states whichever labels (A, B, C) the edge carries; over the
    # face it stays a plane contact.
    P, Q, R = [0.1, -0.45, -1.0], [0.1, -0.45, 1.0], [3.0, -0.45, 0.0]
    tri = np.array([(P, Q, R)[k] for k in order], np.float32)
    t = oracle.TriangleTester()
    t.candidates(BOX, tri, 0, min_radius=0.5)
    assert t.state != 0
    t = oracle.TriangleTester()
    t.candidates(BOX, tri - np.array([0.2, 0, 0], np.float32), 0, min_radius=0.5)
    assert t.state == 0


def test_deep_contact_depth(oracle):
    # Sunk by 0.3 (deeper than the margins): the plane test checks its depth
    # with GJK/MPR and adds the inner-sphere contact.
    tri = np.array([[-6, -0.2, -6], [6, -0.2, -6], [0, -0.2, 6]], np.float32)
    t = oracle.TriangleTester()
    c = t.candidates(BOX, tri, 0, min_radius=0.5)
    assert len(c) == 2
    for k in range(2):
        np.testing.assert_allclose(c[k, 3:6], [0, -1, 0], atol=1e-5)
        assert abs(c[k, 6] - 0.3) < 1e-4
    # A tester already Deep: MPR's depth and normal.
    t.st[0] = 3
    c = t.candidates(BOX, tri, 0, min_radius=0.5)
    assert abs(c[0, 6] - 0.3) < 1e-3 and c[0, 7] == -1
    np.testing.assert_allclose(c[0, 3:6], [0, -1, 0], atol=1e-3)


def test_contact_normal_regions(oracle):
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    # a contact normal points from the convex to the triangle: outward
    # directions of the triangle's features, negated
    cases = [([0, 1, 0], VR_ABC), ([0, -1, 0], VR_ABC), ([0, 0, -1], VR_AB), ([-1, 0, 0], VR_AC),
             ([1, 0, 1], VR_BC), ([-1, 0, -1], VR_A), ([1, 0.1, -0.3], VR_B), ([-0.3, 0.1, 1], VR_C)]
    for out_dir, region in cases:
        d = np.asarray(out_dir, np.float32)
        d /= np.linalg.norm(d)
        assert oracle.triangle_region(tri, -d) == region, out_dir


def test_sphere_triangle_closed_forms(oracle):
    r = 0.5
    tri = np.array([[-1, -0.4, -1], [1, -0.4, -1], [0, -0.4, 1]], np.float32)
    c, reg = oracle.triangle_sphere(r, tri)
    assert reg == VR_ABC
    np.testing.assert_allclose(c[0], [0, -0.4, 0, 0, -1, 0, 0.1, 0], atol=1e-6)
    # near the edge AB (z = -1 side) from outside: the closest point is on the edge
    tri2 = tri + np.array([0, 0.1, 1.2], np.float32)  # edge AB at z = 0.2, y = -0.3
    c, reg = oracle.triangle_sphere(r, tri2)
    assert reg == VR_AB
    d = np.array([0, -0.3, 0.2])
    np.testing.assert_allclose(c[0, :3], d, atol=1e-6)
    np.testing.assert_allclose(c[0, 3:6], d / np.linalg.norm(d), atol=1e-6)
    assert abs(c[0, 6] - (r - np.linalg.norm(d))) < 1e-6
    # out of reach
    c, _ = oracle.triangle_sphere(r, tri - np.array([0, 0.2, 0], np.float32))
    assert len(c) == 0
    # one-sided: the sphere behind a counterclockwise triangle gets nothing
    back = tri[[0, 2, 1]]
    assert len(oracle.triangle_sphere(r, back, sidedness=1)[0]) == 1
    assert len(oracle.triangle_sphere(r, back, sidedness=2)[0]) == 0


def test_escape_returns_to_plane_after_ten_separated_tests(oracle):
    # A triangle far off to the side: the plane test's extreme point is
    # outside the face, so the tester goes Separated; after ten separated
    # tests it tries the plane test again (TryToEscape).
    tri = np.array([[3, -0.6, -1], [5, -0.6, -1], [4, -0.6, 1]], np.float32)
    t = oracle.TriangleTester()
    states = []
    for _ in range(12):
        t.candidates(BOX, tri, 0, min_radius=0.5)
        states.append(t.state)
    assert states[:9] == [1] * 9 and states[9] == 0 and states[10] == 1


def _run(ground, shape, dims, pos, steps, vel=(0, 0, 0), sidedness=0, improve=True, cells=10, quat=(0, 0, 0, 1)):
    from oracle import refcpu
    w = refcpu.World()
    if ground == "mesh":
        V, I, _ = scenes.heightfield_mesh(cells=cells, size=20.0)
        w.add_static_mesh(w.create_mesh_shape(V, I), sidedness, improve)
    else:
        w.add_body(0, [20, 1, 20], 0, [0, -0.5, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0], [1, 0.8, 0])
    w.add_body(shape, dims, 1.0, list(pos) + list(quat) + list(vel) + [0, 0, 0], [1, 0.8, 0])
    w.step(steps)
    return w


@pytest.mark.parametrize("shape,dims,h", [(0, [1, 1, 1], 0.5), (1, [0.5, 0, 0], 0.5), (3, [1.0, 0.3, 0], 0.8)])
def test_bodies_rest_on_a_mesh_as_on_a_box_ground(oracle, shape, dims, h):
    a = _run("mesh", shape, dims, (0.3, h + 0.05, 0.2), 240).motion_states()[1]
    b = _run("box", shape, dims, (0.3, h + 0.05, 0.2), 240).motion_states()[1]
    # both sink to the allowed penetration (0.01) below the surface
    assert abs(a[1] - b[1]) < 2e-3 and abs(a[1] - (h - 0.01)) < 3e-3
    assert np.abs(a[7:13]).max() < 0.01


def test_sliding_and_rolling_on_a_mesh_as_on_a_box_ground(oracle):
    # A box sliding at 4 m/s stops after about v0^2 / (2 mu g), mu = 0.8 x 0.8
    # (kinetic friction of both materials, product blend), a little sooner for
    # the damping and the static friction of the last 0.2 m/s; a sphere
    # rolling at 3 m/s behaves as on the box ground.
    a = _run("mesh", 0, [1, 1, 1], (0.3, 0.5, 0.2), 90, vel=(4, 0, 0)).motion_states()[1]
    b = _run("box", 0, [1, 1, 1], (0.3, 0.5, 0.2), 90, vel=(4, 0, 0)).motion_states()[1]
    stop = 0.3 + 16 / (2 * 0.64 * 9.81)
    assert -0.07 < a[0] - stop < 0 and abs(a[0] - b[0]) < 0.01
    assert np.abs(a[7:10]).max() < 0.01
    a = _run("mesh", 1, [0.5, 0, 0], (0.3, 0.5, 0.2), 120, vel=(3, 0, 0)).motion_states()[1]
    b = _run("box", 1, [0.5, 0, 0], (0.3, 0.5, 0.2), 120, vel=(3, 0, 0)).motion_states()[1]
    np.testing.assert_allclose(a[:13], b[:13], atol=1e-4)


def test_one_sided_mesh_lets_bodies_through_from_behind(oracle):
    # The heightfield's triangles have AB x AC pointing up.  The testers keep
    # contacts of a Clockwise triangle on that side (TriangleSpherePairTester.cs:
    # 42-51: a Counterclockwise triangle drops them when the sphere's centre
    # is on the AB x AC side), so a sphere dropped from above rests on a
    # Clockwise mesh and falls through a Counterclockwise one.
    up = _run("mesh", 1, [0.5, 0, 0], (0.3, 1.5, 0.2), 120, sidedness=1).motion_states()[1]
    down = _run("mesh", 1, [0.5, 0, 0], (0.3, 1.5, 0.2), 120, sidedness=2).motion_states()[1]
    assert abs(up[1] - 0.49) < 3e-3
    assert down[1] < -1.0


def test_box_on_one_sided_mesh(oracle):
    # The plane test of a one-sided triangle looks for the convex's extreme
    # point on the side the triangle keeps: a box dropped on a Clockwise mesh
    # rests on it, on a Counterclockwise one it falls through.
    up = _run("mesh", 0, [1, 1, 1], (0.3, 0.6, 0.2), 120, sidedness=1).motion_states()[1]
    down = _run("mesh", 0, [1, 1, 1], (0.3, 0.6, 0.2), 120, sidedness=2).motion_states()[1]
    assert abs(up[1] - 0.49) < 3e-3
    assert down[1] < -1.0


def test_sphere_on_a_shared_vertex_has_one_contact(oracle):
    # A sphere resting on a vertex that six triangles share: each triangle's
    # tester reports the same vertex contact; the boundary sets keep them all
    # (the vertex blocks only the others), and IsContactUnique drops the
    # repeats, leaving one contact straight below.
    # The same on the middle of an edge two triangles share (two candidates,
    # fewer than the reducer would see).
    for x, z in ((0.0, 0.0), (0.0, 1.0)):
        w = _run("mesh", 1, [0.5, 0, 0], (x, 0.52, z), 20)
        ms = [m for m in w.manifolds() if m.kind == 5]
        assert len(ms) == 1 and ms[0].count == 1
        np.testing.assert_allclose(ms[0].contacts[0].normal, [0, -1, 0], atol=1e-6)
        np.testing.assert_allclose(ms[0].contacts[0].position[0::2], [x, z], atol=1e-6)


def test_transformed_mesh(oracle):
    # A mesh turned by 0.3 rad about z and lifted by 2: a box resting on it
    # settles at the plane's height below its centre.
    from oracle import refcpu
    R = rot([0, 0, 1], 0.3).astype(np.float32)
    V, I, _ = scenes.heightfield_mesh(cells=8, size=20.0)
    w = refcpu.World()
    w.add_static_mesh(w.create_mesh_shape(V, I, (R, np.array([0, 2, 0], np.float32))))
    n = np.array([0, 1, 0]) @ R  # the plane normal (row-vector transform)
    p0 = np.array([0.5, 2.0, 0.3]) + n * 0.6
    w.add_body(1, [0.5, 0, 0], 1.0, list(p0) + [0, 0, 0, 1, 0, 0, 0, 0, 0, 0], [1, 1, 0])
    w.step(30)
    p = w.motion_states()[1][:3]
    assert abs(np.dot(p - np.array([0, 2, 0]), n) - 0.49) < 5e-3


def test_mesh_pile_manifolds(oracle):
    # A pile of every shape on a bumpy mesh (amplitude 1): every mesh manifold
    # has at most four contacts and normals of unit length, and no body falls
    # through the mesh (spheres may still roll down its slopes).
    from tests.helpers import oracle_hulls
    sc, meshes = scenes.mesh_pile(24, seed=2)
    w = oracle.World()
    w.add_scene(sc, oracle_hulls(), meshes)
    w.step(180)
    ms = w.manifolds()
    mesh_manifolds = [m for m in ms if m.kind in (4, 5)]
    assert mesh_manifolds
    for m in mesh_manifolds:
        assert 1 <= m.count <= 4 and m.body_b == 0
        for i in range(m.count):
            assert abs(np.linalg.norm(m.contacts[i].normal) - 1) < 1e-3
            assert m.contacts[i].penetration_depth < 0.1
    st = w.motion_states()
    assert st[1:, 1].min() > -1.5


@pytest.mark.parametrize("sidedness,stops", [(0, True), (1, True), (2, False)])
def test_continuous_body_stops_at_the_mesh(oracle, sidedness, stops):
    # A sphere (radius 0.25) falling at 180 m/s (3 m per step) from 1 m above
    # a mesh: a discrete body tunnels through; a continuous one is clamped to
    # the time of impact of its core sphere (0.8 of its radius) against the
    # triangles, so it stops with its centre 0.2 above the mesh -- unless the
    # mesh is one-sided and keeps only the other side.
    from oracle import refcpu
    V, I, _ = scenes.heightfield_mesh(cells=4, size=20.0)
    for continuous in (False, True):
        w = refcpu.World()
        w.add_static_mesh(w.create_mesh_shape(V, I), sidedness)
        w.add_body(1, [0.25, 0, 0], 1.0, [0.3, 1.0, 0.2, 0, 0, 0, 1, 0, -180, 0, 0, 0, 0], [1, 0.8, 0],
                   flags=1 | (8 if continuous else 0))
        w.step(1)
        y = w.motion_states()[1][1]
        if continuous and stops:
            assert abs(y - 0.8 * 0.25) < 1e-3
        else:
            assert y < -1.9
