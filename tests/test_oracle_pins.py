"""Pins the oracle (the CPU restatement) against what the mathematics fixes:
closed forms, brute force and invariants.  The reference has no golden
vectors or test fixtures for this path (its "tests" are interactive demos),
so every oracle function is pinned here by an independent computation that a
plausible transcription mistake (dropped term, wrong sign, swapped operand)
would fail.  CPU only.
"""
import math

import numpy as np
import pytest

from bepuphysics1_amd import scenes

I = [0.0, 0.0, 0.0, 1.0]


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


# ---------------------------------------------------------------- box-box

def test_box_face_on_face_closed_form(oracle):
    """Axis-aligned unit box sunk 0.02 into a wide slab: the manifold is the
    4 corners of the box's bottom face, normal +-y, depth 0.02."""
    hit, axis, d, c, ids = oracle.box_box([25, 0.5, 25], [0, -0.5, 0], I, [0.5, 0.5, 0.5], [0.3, 0.48, 0.1], I)
    assert hit
    assert np.allclose(np.abs(axis), [0, 1, 0])
    assert len(c) == 4 and len(set(ids.tolist())) == 4
    corners = {(round(float(x), 4), round(float(z), 4)) for x, _, z, _ in c}
    assert corners == {(0.8, 0.6), (-0.2, 0.6), (-0.2, -0.4), (0.8, -0.4)}
    assert np.allclose(c[:, 1], -0.02, atol=1e-6)
    assert np.allclose(c[:, 3], -0.02, atol=1e-6)


def test_box_overlap_rectangle_closed_form(oracle):
    """Two unit boxes stacked with a lateral offset: contacts are the corners
    of the overlap rectangle of the two faces."""
    hit, axis, d, c, ids = oracle.box_box([0.5] * 3, [0, 0, 0], I, [0.5] * 3, [0.2, 0.99, 0.0], I)
    assert hit
    xs = sorted({round(float(v), 4) for v in c[:, 0]})
    zs = sorted({round(float(v), 4) for v in c[:, 2]})
    assert xs == [-0.3, 0.5] and zs == [-0.5, 0.5]
    assert np.allclose(c[:, 3], -0.01, atol=1e-6)


def test_box_separated(oracle):
    hit, *_ = oracle.box_box([0.5] * 3, [0, 0, 0], I, [0.5] * 3, [0, 1.001, 0], I)
    assert not hit
    hit, *_ = oracle.box_box([0.5] * 3, [0, 0, 0], qaxis([1, 1, 0], 0.7), [0.5] * 3, [3, 0, 0], I)
    assert not hit


def test_box_edge_edge_closed_form(oracle):
    """Box A rotated 45 deg about x, box B rotated 45 deg about z, stacked:
    their edges cross at x=z=0; depth = 2*(sqrt(2)/2) - separation."""
    qa, qb = qaxis([1, 0, 0], math.pi / 4), qaxis([0, 0, 1], math.pi / 4)
    sep = 1.38
    hit, axis, d, c, ids = oracle.box_box([0.5] * 3, [0, 0, 0], qa, [0.5] * 3, [0, sep, 0], qb)
    assert hit
    assert len(c) == 1
    assert np.allclose(np.abs(axis), [0, 1, 0], atol=1e-6)
    assert np.allclose(c[0, :3], [0, math.sqrt(0.5), 0], atol=1e-5)
    assert c[0, 3] == pytest.approx(-(2 * math.sqrt(0.5) - sep), abs=1e-5)


def _sat_depth(ha, pa, qa, hb, pb, qb):
    """Brute-force SAT over the 15 axes: minimum overlap (penetration depth)."""
    Ra, Rb = rotmat(qa), rotmat(qb)
    axes = [Ra[:, i] for i in range(3)] + [Rb[:, i] for i in range(3)]
    for i in range(3):
        for j in range(3):
            c = np.cross(Ra[:, i], Rb[:, j])
            if np.linalg.norm(c) > 1e-6:
                axes.append(c / np.linalg.norm(c))
    t = np.asarray(pb) - np.asarray(pa)
    best = np.inf
    for ax in axes:
        ra = sum(ha[k] * abs(ax @ Ra[:, k]) for k in range(3))
        rb = sum(hb[k] * abs(ax @ Rb[:, k]) for k in range(3))
        best = min(best, ra + rb - abs(ax @ t))
    return best


def test_box_box_depth_matches_brute_force_sat(oracle):
    """For random overlapping pairs whose deepest contact is a face contact,
    the manifold's deepest point is no deeper than the SAT depth (+biases),
    and for face-face-dominated cases it equals the SAT depth."""
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(400):
        ha = rng.uniform(0.25, 1.0, 3)
        hb = rng.uniform(0.25, 1.0, 3)
        qa = scenes.random_quaternions(rng, 1)[0]
        qb = scenes.random_quaternions(rng, 1)[0]
        pb = rng.uniform(-1.2, 1.2, 3)
        sat = _sat_depth(ha, [0, 0, 0], qa, hb, pb, qb)
        if sat <= 0.01:
            continue
        hit, axis, d, c, ids = oracle.box_box(ha, [0, 0, 0], qa, hb, pb, qb)
        assert hit
        assert np.linalg.norm(axis) == pytest.approx(1.0, abs=1e-5)
        # -distance returned is the chosen axis' penetration; SAT picks the
        # minimum over axes, and the reference biases (<= 0.02) the choice.
        assert -d <= sat + 0.02 + 1e-4
        assert -d >= sat - 1e-4
        if len(c):
            assert (c[:, 3] <= 1e-6).all()
        checked += 1
    assert checked > 100


def test_box_ids_stable_under_translation(oracle):
    """Contact IDs come from features, so they do not change when both boxes
    move together (persistent manifold matching relies on this)."""
    q = qaxis([0.3, 1, 0.2], 0.4)
    _, _, _, c1, i1 = oracle.box_box([0.5] * 3, [0, 0, 0], I, [0.5] * 3, [0.1, 0.95, 0.05], q)
    _, _, _, c2, i2 = oracle.box_box([0.5] * 3, [3, 1, -2], I, [0.5] * 3, [3.1, 1.95, -1.95], q)
    assert sorted(i1.tolist()) == sorted(i2.tolist())


# ---------------------------------------------------------------- spheres

def test_sphere_sphere_closed_form(oracle):
    hit, out = oracle.sphere_sphere(0.5, 0.25, [0, 0, 0], [0, 0.7, 0])
    assert hit
    assert np.allclose(out[3:6], [0, 1, 0])
    assert out[6] == pytest.approx(0.05, abs=1e-6)
    assert np.allclose(out[0:3], [0, 0.7 * 0.5 / 0.75, 0], atol=1e-6)
    hit, _ = oracle.sphere_sphere(0.5, 0.25, [0, 0, 0], [0, 0.851, 0])  # beyond 0.1 speculative margin
    assert not hit


def test_box_sphere_closed_form(oracle):
    hit, out = oracle.box_sphere([0.5] * 3, [0, 0, 0], I, 0.5, [0.2, 0.95, 0.1])
    assert hit
    assert np.allclose(out[0:3], [0.2, 0.5, 0.1], atol=1e-6)
    assert np.allclose(out[3:6], [0, 1, 0], atol=1e-6)
    assert out[6] == pytest.approx(0.05, abs=1e-6)
    # centre inside the box: pushed out through the nearest face
    hit, out = oracle.box_sphere([0.5, 0.5, 0.5], [0, 0, 0], I, 0.2, [0.4, 0.0, 0.0])
    assert hit and np.allclose(out[3:6], [1, 0, 0]) and out[6] == pytest.approx(0.3, abs=1e-6)


# ---------------------------------------------------------------- broadphase

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grid_and_sweep_pairs_equal_brute_force(oracle, seed):
    """The oracle's grid broadphase (default) and its sort-and-sweep give the
    brute-force pair set, including bodies far larger than the grid cell."""
    sc = scenes.stress(1500, seed=seed)
    if seed == 2:
        sc = np.concatenate([scenes.mixed_sizes(400, seed=5, edge=60.0), sc])
    a = oracle.World()
    a.add_scene(sc)
    a.broadphase_only()
    b = oracle.World()
    b.add_scene(sc)
    b.set_bruteforce_broadphase(True)
    b.broadphase_only()
    c = oracle.World()
    c.add_scene(sc)
    c.set_broadphase_mode(2)
    c.broadphase_only()
    assert np.array_equal(a.pairs(), b.pairs())
    assert np.array_equal(c.pairs(), b.pairs())
    # and all equal an independent numpy brute force on the oracle's AABBs
    # (kinematic-kinematic pairs excluded)
    ab = a.aabbs()
    lo, hi = ab[:, :3], ab[:, 3:]
    ov = np.all((lo[:, None, :] <= hi[None, :, :]) & (lo[None, :, :] <= hi[:, None, :]), axis=2)
    dyn = sc["mass"] > 0
    ov &= dyn[:, None] | dyn[None, :]
    i, j = np.nonzero(np.triu(ov, 1))
    keys = np.sort((i.astype(np.uint64) << np.uint64(32)) | j.astype(np.uint64))
    assert np.array_equal(a.pairs(), keys)


def test_box_aabb_closed_form(oracle):
    """AABB half extent of a rotated box is |R| h, expanded by v*dt on the
    velocity's side (BoxShape.GetBoundingBox + ExpandBoundingBox)."""
    rng = np.random.default_rng(5)
    q = scenes.random_quaternions(rng, 1)[0]
    st = np.zeros(13, np.float32)
    st[0:3] = [1, 2, 3]
    st[3:7] = q
    st[7:10] = [0.5, -1.0, 0.0]
    w = oracle.World()
    s = oracle.default_settings()
    s.gravity[:] = [0, 0, 0]
    w = oracle.World(s)
    w.add_body(0, (1.0, 2.0, 3.0), 0.0, st, (1, 1, 0))  # kinematic: no damping, no gravity
    w.broadphase_only()
    ab = w.aabbs()[0]
    R = rotmat(q)
    ext = np.abs(R) @ np.array([0.5, 1.0, 1.5])
    dt = 1 / 60
    assert np.allclose(ab[0:3], [1 - ext[0], 2 - ext[1] - 1.0 * dt, 3 - ext[2]], atol=1e-5)
    assert np.allclose(ab[3:6], [1 + ext[0] + 0.5 * dt, 2 + ext[1], 3 + ext[2]], atol=1e-5)


# ---------------------------------------------------------------- integration / solver

def test_free_fall_closed_form(oracle):
    """No contacts: v_{n+1} = (v_n + g dt) * (1-0.03)^dt; x += v dt.
    (Stabilization is switched off: it legitimately boosts damping while the
    body is slow, EntityBase.cs:927-936.)"""
    s = oracle.default_settings()
    s.use_stabilization = 0
    w = oracle.World(s)
    st = np.zeros(13, np.float32)
    st[1] = 100.0
    st[6] = 1.0
    w.add_body(1, (0.5, 0, 0), 1.0, st, (1, 0.8, 0))
    v, y = 0.0, 100.0
    damp = (1 - 0.03) ** (1 / 60)
    for _ in range(30):
        v = (v - 9.81 / 60) * damp
        y += v / 60
    w.step(30)
    ms = w.motion_states()[0]
    assert ms[8] == pytest.approx(v, rel=1e-5)
    assert ms[1] == pytest.approx(y, rel=1e-6)


def test_spin_quaternion_stays_unit(oracle):
    s = oracle.default_settings()
    s.gravity[:] = [0, 0, 0]
    w = oracle.World(s)
    st = np.zeros(13, np.float32)
    st[6] = 1.0
    st[10:13] = [0.0, 3.0, 0.0]
    w.add_body(0, (1, 1, 1), 1.0, st, (1, 0.8, 0))
    w.step(60)
    ms = w.motion_states()[0]
    q = ms[3:7]
    assert np.linalg.norm(q) == pytest.approx(1.0, abs=1e-6)
    # rotation about +y only; the damped angle is sum of w_k dt
    damp = 0.85 ** (1 / 60)
    angle = sum(3.0 * damp ** (k + 1) / 60 for k in range(60))
    assert abs(q[0]) < 1e-6 and abs(q[2]) < 1e-6
    assert 2 * math.atan2(q[1], q[3]) == pytest.approx(angle, rel=2e-3)


def test_resting_sphere_impulse_equals_weight(oracle):
    """A sphere resting on a static box: after settling, the normal impulse
    per step equals m*g*dt (the contact carries the weight), friction ~0."""
    sc = scenes.make([scenes.ground(), scenes._row(scenes.SHAPE_SPHERE, (0.5, 0, 0), 2.0, (0, 0.49, 0))])
    w = oracle.World()
    w.add_scene(sc)
    w.step(240)
    imp = w.impulses()
    assert len(imp) == 1
    assert imp[0].normal[0] == pytest.approx(2.0 * 9.81 / 60, rel=2e-2)
    assert abs(imp[0].friction[0]) < 1e-4


def test_third_law_and_cone(oracle):
    """Every contact row applies equal and opposite linear impulses, so total
    momentum of an isolated two-box system (no ground, no gravity) is
    conserved; friction impulses stay inside the Coulomb cone."""
    s = oracle.default_settings()
    s.gravity[:] = [0, 0, 0]
    s.linear_damping = 0.0
    s.angular_damping = 0.0
    s.use_stabilization = 0
    w = oracle.World(s)
    st = np.zeros(13, np.float32)
    st[6] = 1.0
    st[7] = 1.0
    st[9] = 0.2
    w.add_body(0, (1, 1, 1), 1.0, st, (0.6, 0.5, 0))
    st2 = np.zeros(13, np.float32)
    st2[0] = 1.02
    st2[1] = 0.3
    st2[6] = 1.0
    w.add_body(0, (1, 1, 1), 3.0, st2, (0.6, 0.5, 0))
    p0 = 1.0 * st[7:10] + 3.0 * st2[7:10]
    for _ in range(20):
        w.step(1)
        ms = w.motion_states()
        p = 1.0 * ms[0, 7:10] + 3.0 * ms[1, 7:10]
        assert np.allclose(p, p0, atol=1e-5)
        for r in w.impulses():
            ntot = sum(r.normal[:r.count])
            assert all(x >= 0 for x in r.normal[:r.count])
            assert math.hypot(*r.sliding) <= 0.6 * ntot + 1e-6


def test_pyramid_stable(oracle):
    """C1 pyramid, 300 steps: nothing falls, penetration stays bounded."""
    sc = scenes.pyramid()
    w = oracle.World()
    w.add_scene(sc)
    w.step(300)
    ms = w.motion_states()
    assert np.abs(ms[1:, 0:3] - sc["state"][1:, 0:3]).max() < 0.2
    assert w.stats().max_penetration < 0.05
    assert np.abs(ms[1:, 7:13]).max() < 0.05


def test_colouring_is_proper_and_greedy(oracle):
    """Same colour => no shared dynamic body; colour = mex of the colours of
    higher-priority neighbours (checked by an independent recomputation)."""
    sc = scenes.random_pile(200, seed=7)
    w = oracle.World()
    w.add_scene(sc)
    w.step(90)
    imps = w.impulses()
    dyn = sc["mass"] > 0
    col = {int(r.key): int(r.color) for r in imps}
    bodies = {k: (k >> 32, k & 0xffffffff) for k in col}
    by_body = {}
    for k, (a, b) in bodies.items():
        for x in (a, b):
            if dyn[x]:
                by_body.setdefault(x, []).append(k)
    prio = {k: oracle.mix64(k) for k in col}
    for k, (a, b) in bodies.items():
        nb = set()
        for x in (a, b):
            if dyn[x]:
                nb.update(by_body[x])
        nb.discard(k)
        assert all(col[n] != col[k] for n in nb)
        higher = {col[n] for n in nb if prio[n] > prio[k]}
        mex = 0
        while mex in higher:
            mex += 1
        assert col[k] == mex
    # solve order is (colour, key)
    order = [(int(r.color), int(r.key)) for r in imps]
    assert order == sorted(order)


def test_mix64_is_a_bijection_sample(oracle):
    keys = np.arange(20000, dtype=np.uint64) * np.uint64(0x100000001)
    vals = {oracle.mix64(int(k)) for k in keys}
    assert len(vals) == len(keys)


@pytest.mark.parametrize("threads", [2, 5])
def test_multithreaded_oracle_is_bit_identical(oracle, threads):
    """The multithreaded step (body stages and pair updates in parallel, the
    solver colour by colour) computes the same bits as the single-threaded one."""
    from helpers import oracle_hulls
    sc = scenes.random_pile(150, seed=4, shapes=(0, 1, 2, 3))
    a, b = oracle.World(), oracle.World()
    a.add_scene(sc, oracle_hulls())
    b.add_scene(sc, oracle_hulls())
    b.set_threads(threads)
    a.step(40)
    b.step(40)
    assert np.array_equal(a.motion_states().view(np.uint32), b.motion_states().view(np.uint32))
    assert np.array_equal(a.pairs(), b.pairs())
    ia = {int(r.key): bytes(r) for r in a.impulses()}
    ib = {int(r.key): bytes(r) for r in b.impulses()}
    assert ia == ib


# ---------------------------------------------------------------- host API between steps

def test_removed_body_leaves_every_stage(oracle):
    """Space.Remove: the body has no pairs from the next step on (the pair set
    equals brute force over the other bodies) and no longer moves; a second
    removal is a foreign remove (not found)."""
    sc = scenes.random_pile(60, seed=9)
    w = oracle.World()
    w.add_scene(sc)
    w.step(20)
    gone = [4, 11, 30]
    for b in gone:
        assert w.remove(b) == 0
    assert w.remove(11) == 3
    assert w.remove(len(sc)) == 3
    frozen = w.motion_states()[gone].copy()
    w.step(5)
    keys = w.pairs()
    ends = set((keys >> np.uint64(32)).tolist()) | set((keys & np.uint64(0xffffffff)).tolist())
    assert not ends & set(gone)
    assert np.array_equal(w.motion_states()[gone], frozen)
    # the rest: brute force over the AABBs of the bodies still present
    ab = w.aabbs()
    keep = np.ones(len(sc), bool)
    keep[gone] = False
    lo, hi = ab[:, :3], ab[:, 3:]
    ov = np.all((lo[:, None, :] <= hi[None, :, :]) & (lo[None, :, :] <= hi[:, None, :]), axis=2)
    dyn = sc["mass"] > 0
    ov &= (dyn[:, None] | dyn[None, :]) & keep[:, None] & keep[None, :]
    i, j = np.nonzero(np.triu(ov, 1))
    assert np.array_equal(np.sort((i.astype(np.uint64) << np.uint64(32)) | j.astype(np.uint64)), keys)


def test_single_field_setters_touch_only_their_field(oracle):
    """Entity.Position (1), Orientation (2), LinearVelocity (4) and
    AngularVelocity (8) setters write their own part only; the orientation is
    normalised as the reference's setter does (a unit quaternion stays put)."""
    s = oracle.default_settings()
    s.gravity[:] = [0, 0, 0]
    w = oracle.World(s)
    st0 = np.zeros(13, np.float32)
    st0[6] = 1.0
    st0[0:3] = [1, 2, 3]
    w.add_body(0, (1, 1, 1), 1.0, st0, (1, 0.8, 0))
    new = np.arange(13, dtype=np.float32)[None, :] + 1
    for mask, sl in ((1, slice(0, 3)), (2, slice(3, 7)), (4, slice(7, 10)), (8, slice(10, 13))):
        before = w.motion_states()[0].copy()
        assert w.set_motion_states([0], new, mask) == 0
        after = w.motion_states()[0]
        others = np.ones(13, bool)
        others[sl] = False
        assert np.array_equal(after[others], before[others])
        if mask == 2:
            q = new[0, 3:7] / np.linalg.norm(new[0, 3:7])
            assert np.allclose(after[sl], q, atol=1e-7)
        else:
            assert np.array_equal(after[sl], new[0, sl])
    assert w.set_motion_states([0], new, 0) == 1
    bad = new.copy()
    bad[0, 8] = np.nan
    assert w.set_motion_states([0], bad, 4) == 1
    assert w.set_motion_states([0], bad, 1) == 0  # the NaN is not in the written part


def test_settings_change_takes_effect_next_step(oracle):
    """ForceUpdater.Gravity and the time step changed between steps: a free
    body's velocity follows v += g dt with the new values from the next step."""
    s = oracle.default_settings()
    s.linear_damping = s.angular_damping = 0.0
    s.use_stabilization = 0
    w = oracle.World(s)
    st0 = np.zeros(13, np.float32)
    st0[6] = 1.0
    w.add_body(0, (1, 1, 1), 1.0, st0, (1, 0.8, 0))
    w.step(3)
    v3 = w.motion_states()[0][8]
    assert v3 == pytest.approx(3 * -9.81 / 60, rel=1e-6)
    s2 = oracle.default_settings()
    s2.linear_damping = s2.angular_damping = 0.0
    s2.use_stabilization = 0
    s2.gravity[:] = [2.0, 0.0, 0.0]
    s2.dt = 0.01
    assert w.set_settings(s2) == 0
    w.step(4)
    ms = w.motion_states()[0]
    assert ms[7] == pytest.approx(4 * 2.0 * 0.01, rel=1e-6)
    assert ms[8] == pytest.approx(v3, rel=1e-6)
    bad = oracle.default_settings()
    bad.dt = 0.0
    assert w.set_settings(bad) == 1
