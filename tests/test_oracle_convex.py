"""Pins the oracle's general convex path (support mappings, GJK, MPR, the
pair tester's state machine and the incremental manifold) against what the
mathematics fixes: closed forms for spheres, capsules and boxes, the
distance and penetration depth of polytopes by brute force (a small QP and a
dense direction search), and invariants of the contacts.  The reference ships
no golden vectors for this path.  CPU only.
"""
import math

import numpy as np
import pytest
from scipy.optimize import minimize

I = [0.0, 0.0, 0.0, 1.0]
MARGIN = 0.04


def qaxis(axis, angle):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    s = math.sin(angle / 2)
    return [a[0] * s, a[1] * s, a[2] * s, math.cos(angle / 2)]


def rotmat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def rand_quat(rng):
    q = rng.normal(size=4)
    return list(q / np.linalg.norm(q))


def sphere(r):
    return {"kind": 1, "margin": r}


def box(hx, hy, hz, margin=MARGIN):
    return {"kind": 0, "half": [hx, hy, hz], "margin": margin}


def capsule(length, r):
    return {"kind": 3, "half_length": length / 2, "margin": r}


def hull(verts, margin=MARGIN):
    return {"kind": 2, "verts": np.asarray(verts, np.float32), "margin": margin}


def sphere_points(rng, n, radius=1.0, squash=(1.0, 1.0, 1.0)):
    """n points on a (squashed) sphere, recentred on their mean so that the
    shape's origin lies inside it, as ConvexHullShape's recentring ensures."""
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p = p * radius * np.asarray(squash)
    return (p - p.mean(axis=0)).astype(np.float32)


def core_points(sh, pose):
    """World-space vertices of a shape's core (the shape without its margin)."""
    k = sh["kind"]
    if k == 1:
        loc = np.zeros((1, 3))
    elif k == 3:
        loc = np.array([[0, sh["half_length"], 0], [0, -sh["half_length"], 0]])
    elif k == 0:
        c = np.asarray(sh["half"], float) - sh["margin"]
        loc = np.array([[sx * c[0], sy * c[1], sz * c[2]] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
    else:
        loc = np.asarray(sh["verts"], float)
    return loc @ rotmat(pose[3:]).T + np.asarray(pose[:3], float)


def core_distance(pa_pts, pb_pts):
    """Distance between the convex hulls of two point sets: the convex QP
    min |P'l - Q'm| over l, m >= 0 each summing to one (SLSQP; a convex
    problem, so its local minimum is the global one)."""
    P = np.asarray(pa_pts, float)
    Q = np.asarray(pb_pts, float)
    na, nb = len(P), len(Q)

    def f(x):
        d = x[:na] @ P - x[na:] @ Q
        return d @ d

    def g(x):
        d = x[:na] @ P - x[na:] @ Q
        return np.concatenate([2 * P @ d, -2 * Q @ d])

    x0 = np.concatenate([np.full(na, 1 / na), np.full(nb, 1 / nb)])
    res = minimize(f, x0, jac=g, method="SLSQP", bounds=[(0, 1)] * (na + nb),
                   constraints=[{"type": "eq", "fun": lambda x: x[:na].sum() - 1, "jac": lambda x: np.r_[np.ones(na), np.zeros(nb)]},
                                {"type": "eq", "fun": lambda x: x[na:].sum() - 1, "jac": lambda x: np.r_[np.zeros(na), np.ones(nb)]}],
                   options={"ftol": 1e-15, "maxiter": 1000})
    return math.sqrt(max(res.fun, 0.0))


def in_hull(pts, p, tol=2e-6):
    """Whether p lies in the convex hull of pts, up to FP32 rounding."""
    return core_distance(pts, np.asarray(p, float)[None, :]) < tol * max(1.0, float(np.abs(p).max()))


def fib_dirs(n):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = math.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], axis=1)


def support_gap(pa_pts, ma, pb_pts, mb, n):
    """h_A(n) + h_B(-n) for unit n: the depth of the Minkowski difference A-B along n."""
    return (pa_pts @ n).max(axis=0) - (pb_pts @ n).min(axis=0) + ma + mb


def brute_depth(pa_pts, ma, pb_pts, mb):
    d = fib_dirs(40000)
    gaps = support_gap(pa_pts, ma, pb_pts, mb, d.T)
    best = d[np.argmin(gaps)]

    def f(v):
        v = v / np.linalg.norm(v)
        return float(support_gap(pa_pts, ma, pb_pts, mb, v[:, None])[0])

    res = minimize(f, best, method="Nelder-Mead", options={"xatol": 1e-10, "fatol": 1e-12, "maxiter": 4000})
    return min(res.fun, gaps.min())


# ---------------------------------------------------------------- closed forms

@pytest.mark.parametrize("d", [0.35, 0.55, 0.699])
def test_sphere_sphere_general_path_closed_form(oracle, d):
    """Two spheres through the general path: GJK closest points of the cores are
    the centres, so the contact is the sphere-sphere closed form."""
    ra, rb = 0.3, 0.4
    pa = [1.0, 2.0, 3.0] + I
    pb = [1.0 + d * 0.6, 2.0 + d * 0.8, 3.0] + I
    t = oracle.ConvexTester()
    hit, c = t.candidate(sphere(ra), pa, sphere(rb), pb)
    assert hit and t.state.value == 1
    n = np.array([0.6, 0.8, 0.0])
    assert np.allclose(c[3:6], n, atol=1e-6)
    assert abs(c[6] - (ra + rb - d)) < 1e-6
    assert np.allclose(c[:3], np.array(pa[:3]) + n * d * ra / (ra + rb), atol=1e-6)
    assert c[7] == 0
    # Compare with the dedicated sphere-sphere test (SphereTester).
    ok, ref7 = oracle.sphere_sphere(ra, rb, pa[:3], pb[:3])
    assert ok
    assert np.allclose(np.abs(c[3:6]), np.abs(ref7[3:6]), atol=1e-4)
    assert abs(c[6] - ref7[6]) < 1e-4


def test_sphere_sphere_separated_stays_separated(oracle):
    t = oracle.ConvexTester()
    hit, _ = t.candidate(sphere(0.3), [0, 0, 0] + I, sphere(0.4), [0.71, 0, 0] + I)
    assert not hit and t.state.value == 0
    # The boolean GJK leaves the last search direction -(closest point of A-B),
    # which points from A towards B.
    assert t.sep_axis[0] > 0


@pytest.mark.parametrize("face", range(6))
def test_box_sphere_general_path_face_closed_form(oracle, face):
    """A sphere over a box face: the rounded box of the general path equals the
    box there, so depth and normal match the box-sphere closed form."""
    axis, sign = face // 3, 1 - 2 * (face % 2)
    half = np.array([0.5, 0.7, 0.9])
    sp = np.array([0.1, -0.05, 0.2])
    sp[axis] = sign * (half[axis] + 0.25 - 0.05)
    t = oracle.ConvexTester()
    hit, c = t.candidate(box(*half), [0, 0, 0] + I, sphere(0.25), list(sp) + I)
    assert hit
    n = np.zeros(3)
    n[axis] = sign
    assert np.allclose(c[3:6], n, atol=1e-5)
    assert abs(c[6] - 0.05) < 1e-5
    ok, ref7 = oracle.box_sphere(half, [0, 0, 0], I, 0.25, sp)
    assert ok
    assert abs(abs(ref7[6]) - c[6]) < 1e-4


def test_parallel_capsules_closed_form(oracle):
    """Two parallel capsules along y, 0.5 apart with radius 0.3: depth 0.1, normal +x."""
    t = oracle.ConvexTester()
    hit, c = t.candidate(capsule(2.0, 0.3), [0, 0, 0] + I, capsule(1.0, 0.3), [0.5, 0.2, 0] + I)
    assert hit
    assert np.allclose(c[3:6], [1, 0, 0], atol=1e-6)
    assert abs(c[6] - 0.1) < 1e-6
    assert abs(c[0] - 0.25) < 1e-6


def test_crossed_capsules_closed_form(oracle):
    """Capsule A along y, capsule B along x (rotated 90 degrees about z), offset
    along z by 0.5: the segments' closest points are the two centres."""
    t = oracle.ConvexTester()
    hit, c = t.candidate(capsule(2.0, 0.3), [0, 0, 0] + I, capsule(2.0, 0.3), [0, 0, 0.5] + qaxis([0, 0, 1], math.pi / 2))
    assert hit
    assert np.allclose(c[3:6], [0, 0, 1], atol=1e-6)
    assert abs(c[6] - 0.1) < 1e-6


def test_mpr_box_box_axis_aligned_depth(oracle):
    """Deep box-box overlap: MPR's depth is the smallest axis overlap."""
    a, b = box(0.5, 0.5, 0.5), box(0.5, 0.5, 0.5)
    hit, c = oracle.mpr_contact(a, [0, 0, 0] + I, b, [0.1, 0.7, 0.05] + I, [0, 1, 0])
    assert hit and c[7] == -1
    assert np.allclose(c[3:6], [0, 1, 0], atol=1e-4)
    assert abs(c[6] - 0.3) < 1e-4


# ---------------------------------------------------------------- brute force

@pytest.mark.parametrize("seed", range(8))
def test_gjk_distance_matches_qp(oracle, seed):
    """GJK closest points between hull/box/capsule cores: their distance equals
    the QP minimum distance between the cores' convex hulls."""
    rng = np.random.default_rng(seed)
    kinds = [hull(sphere_points(rng, 16, 0.8, (1.0, 0.7, 1.2))), box(0.4, 0.6, 0.5), capsule(1.2, 0.3),
             hull(sphere_points(rng, 12, 0.6))]
    sa, sb = kinds[seed % 4], kinds[(seed + 1) % 4]
    pa = list(rng.normal(size=3)) + rand_quat(rng)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    pb = list(np.array(pa[:3]) + direction * 2.2) + rand_quat(rng)
    inter, ca, cb = oracle.gjk_closest(sa, pa, sb, pb)
    assert not inter
    A, B = core_points(sa, pa), core_points(sb, pb)
    # Certificate of optimality: both points lie on their shapes, and the
    # shapes' separation along n = (cb - ca)/|cb - ca| equals |cb - ca|
    # (any separation is a lower bound of the distance, any pair of points an
    # upper bound).
    assert in_hull(A, ca) and in_hull(B, cb)
    got = float(np.linalg.norm(cb - ca))
    n = (cb - ca).astype(float) / got
    sep = float((B @ n).min() - (A @ n).max())
    assert abs(got - sep) < 2e-5 * max(1.0, got), (got, sep)
    qp = core_distance(A, B)
    assert abs(got - qp) < 2e-4 * max(1.0, qp), (got, qp)


@pytest.mark.parametrize("seed", range(8))
def test_mpr_depth_against_direction_search(oracle, seed):
    """Deeply overlapping hulls.  MPR's depth is the distance from the origin to
    the boundary of the Minkowski difference M = A - B along the returned unit
    normal n (the last ray cast's direction): the point depth*n lies on the
    boundary of M, and the depth is bracketed by the distance from the origin
    to the boundary (the minimum support gap over all directions, found by a
    dense search) below and by the support gap along n above.  MPR's
    refinement stops at a local minimum after at most three casts, so the
    depth need not reach the lower bound."""
    rng = np.random.default_rng(100 + seed)
    sa = hull(sphere_points(rng, 16, 1.0, (1.0, 0.6, 0.8)))
    sb = hull(sphere_points(rng, 16, 0.8, (0.7, 1.0, 1.0)))
    pa = [0, 0, 0] + rand_quat(rng)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    pb = list(d * 0.6) + rand_quat(rng)
    hit, c = oracle.mpr_contact(sa, pa, sb, pb, list(d))
    assert hit
    n = c[3:6].astype(float)
    assert abs(np.linalg.norm(n) - 1) < 1e-5
    A, B = core_points(sa, pa), core_points(sb, pb)
    own = float(support_gap(A, MARGIN, B, MARGIN, n[:, None] / np.linalg.norm(n))[0])
    bf = brute_depth(A, MARGIN, B, MARGIN)
    assert bf - 2e-4 <= c[6] <= own + 2e-5, (bf, c[6], own)
    # The point t*n lies on the boundary of M, which is the core difference
    # C = conv{a - b} grown by the summed margins: its distance to C equals
    # the summed margins.
    # dist(p, A - B) is the distance between A and B + p.
    p = c[6] * n / np.linalg.norm(n)
    dist = core_distance(A, B + p)
    assert abs(dist - 2 * MARGIN) < 2e-4, dist


@pytest.mark.parametrize("seed", range(6))
def test_boolean_gjk_agrees_with_qp(oracle, seed):
    """The boolean GJK (on the margin-expanded shapes) reports intersection
    exactly when the core distance is below the summed margins."""
    rng = np.random.default_rng(200 + seed)
    sa = hull(sphere_points(rng, 14, 0.7))
    sb = box(0.3, 0.5, 0.4)
    for k in range(12):
        pa = [0, 0, 0] + rand_quat(rng)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pb = list(d * rng.uniform(0.5, 1.8)) + rand_quat(rng)
        qp = core_distance(core_points(sa, pa), core_points(sb, pb))
        if abs(qp - 2 * MARGIN) < 1e-3:
            continue
        hit, _ = oracle.gjk_intersecting(sa, pa, sb, pb)
        assert hit == (qp < 2 * MARGIN), (qp, hit)


# ---------------------------------------------------------------- state machine

def test_state_machine_shallow_to_deep_and_back(oracle):
    """A sphere sinking into a hull: Separated -> Shallow (GJK on the cores)
    -> Deep (cores overlap, MPR) -> back to Shallow once the depth is below
    the summed margins."""
    rng = np.random.default_rng(7)
    h = hull(sphere_points(rng, 16, 1.0))
    t = oracle.ConvexTester()
    states = []
    for y in (2.0, 1.3, 0.9, 0.3, 0.9, 1.2):
        hit, c = t.candidate(h, [0, 0, 0] + I, sphere(0.3), [0, y, 0] + I, va=(0, 0, 0), vb=(0, -1, 0))
        states.append(t.state.value)
        if hit:
            assert abs(np.linalg.norm(c[3:6]) - 1) < 1e-5
    assert states[0] == 0
    assert 2 in states
    assert states[-1] in (0, 1)


def test_deep_from_separated_uses_relative_velocity(oracle):
    """Entering deep contact straight from separation, the MPR search starts
    along the relative linear velocity (A minus B)."""
    # The sphere's centre lies inside the box's core (half extent 0.46), so the
    # closest-point GJK reports intersection and MPR takes over.
    t = oracle.ConvexTester()
    hit, c = t.candidate(box(0.5, 0.5, 0.5), [0, 0, 0] + I, sphere(0.3), [0.0, 0.3, 0.0] + I,
                         va=(0, 0, 0), vb=(0, -2, 0))
    assert hit and t.state.value == 2
    assert np.allclose(t.local_dir, [0, 2, 0])
    assert np.allclose(c[3:6], [0, 1, 0], atol=1e-4)
    assert abs(c[6] - 0.5) < 1e-4


# ---------------------------------------------------------------- whole worlds

def _world(oracle, rows, hulls=None):
    from bepuphysics1_amd import scenes
    w = oracle.World()
    w.add_scene(scenes.make(rows), hulls)
    return w


def test_capsule_rests_on_ground(oracle):
    from bepuphysics1_amd import scenes
    rows = [scenes._row(scenes.SHAPE_BOX, (20, 1, 20), 0, (0, -0.5, 0)),
            scenes._row(scenes.SHAPE_CAPSULE, (1.0, 0.25, 0), 1.0, (0, 0.26, 0), quat=qaxis([0, 0, 1], math.pi / 2))]
    w = _world(oracle, rows)
    w.step(240)
    ms = w.motion_states()
    y = ms[1, 1]
    assert abs(y - 0.25) < 0.02, y
    assert np.linalg.norm(ms[1, 7:10]) < 0.05
    m = [x for x in w.manifolds() if x.count]
    assert len(m) == 1 and m[0].kind == 3
    assert 1 <= m[0].count <= 4
    for i in range(m[0].count):
        assert abs(abs(m[0].contacts[i].normal[1]) - 1) < 1e-3


def test_hull_rests_on_ground_with_several_contacts(oracle):
    """A box-shaped hull (8 corners) settles face-down on the ground: the
    incremental manifold accumulates several contacts, at most four, all with
    the ground normal, and the hull comes to rest at its half height."""
    from bepuphysics1_amd import scenes
    corners = np.array([[sx * 0.5, sy * 0.3, sz * 0.4] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)],
                       np.float32)
    vd = np.diag([(0.6 ** 2 + 0.8 ** 2) / 12, (1.0 ** 2 + 0.8 ** 2) / 12, (1.0 ** 2 + 0.6 ** 2) / 12]).astype(np.float32)
    rows = [scenes._row(scenes.SHAPE_BOX, (20, 1, 20), 0, (0, -0.5, 0)),
            scenes._row(scenes.SHAPE_CONVEX_HULL, (0, 0, 0), 1.0, (0, 0.5, 0), quat=qaxis([1, 0, 0], 0.05))]
    w = _world(oracle, rows, hulls=[(corners, vd)])
    w.step(300)
    ms = w.motion_states()
    assert abs(ms[1, 1] - (0.3 + MARGIN)) < 0.03, ms[1, 1]
    m = [x for x in w.manifolds() if x.count]
    assert len(m) == 1
    assert 2 <= m[0].count <= 4
