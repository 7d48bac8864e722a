"""The reference host API calls beyond add and step, on the GPU against the
oracle doing the same (run on an MI355X):

  Space.Remove(entity)                       Space.cs:364-482          bepu_remove
  Entity.Position / Orientation / LinearVelocity / AngularVelocity /
  MotionState setters                        EntityBase.cs:51-235      bepu_set_motion_states
  Solver.IterationLimit, ForceUpdater.Gravity, TimeStepDuration
                                             Solver.cs:19, ForceUpdater.cs:19-29   bepu_set_settings

and the error contract of the step: a capacity overflow in any step is
sticky and every later readback reports it (BEPU_E_CAPACITY) until cleared.
Results are compared bit for bit (motion states) and exactly (pair sets,
manifold and impulse tables).
"""
import ctypes

import numpy as np
import pytest

from bepuphysics1_amd import Box, Space, Sphere, _lib, scenes
from helpers import f32bits, impulse_table, manifold_table, oracle_hulls, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

E_NOT_FOUND, E_CAPACITY = 3, 7


def _same(g, o):
    assert np.array_equal(g.pairs(), o.pairs())
    assert manifold_table(g.manifolds()) == manifold_table(o.manifolds())
    assert impulse_table(g.impulses()) == impulse_table(o.impulses())
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))


def _pile():
    sc = scenes.random_pile(90, seed=11, shapes=(0, 1, 2, 3))
    return sc


def test_remove_and_state_setters_mid_run_exact():
    """Bodies removed, teleported, re-aimed and added between steps: the GPU
    world and the oracle stay bit-identical."""
    sc = _pile()
    g, o = run_gpu(sc, 40, capacity=len(sc) + 8), run_oracle(sc, 40)
    _same(g, o)
    L, h = g._lib, g._handle()
    # Space.Remove on a few bodies, one of them twice (ArgumentException -> not found)
    for bid in (5, 17, 40):
        g.Remove(bid)
        assert o.remove(bid) == 0
    assert L.bepu_remove(h, 17) == E_NOT_FOUND
    assert o.remove(17) == E_NOT_FOUND
    assert L.bepu_remove(h, 10_000) == E_NOT_FOUND
    # MotionState setter on two bodies, single-field setters on three others
    ms = o.motion_states()
    st = ms[[3, 8]].copy()
    st[0, 0:3] += (0.5, 3.0, -0.25)
    st[0, 3:7] = (0.3, 0.1, -0.2, 0.9)  # not unit: the setter normalises
    st[1, 7:13] = (1.0, 2.0, 0.0, 0.0, 3.0, -1.0)
    g.set_motion_states([3, 8], st, 15)
    assert o.set_motion_states([3, 8], st, 15) == 0
    one = np.zeros((1, 13), np.float32)
    one[0, 7:10] = (0.0, 4.0, 0.0)
    g.set_motion_states([10], one, 4)  # LinearVelocity only
    o.set_motion_states([10], one, 4)
    one[0, 0:3] = (1.0, 12.0, -2.0)
    g.set_motion_states([12], one, 1)  # Position only
    o.set_motion_states([12], one, 1)
    one[0, 10:13] = (0.0, 0.0, 6.0)
    g.set_motion_states([21], one, 8)  # AngularVelocity only
    o.set_motion_states([21], one, 8)
    # a removed body cannot be set; nothing is written then
    bad = np.zeros((2, 13), np.float32)
    bad[:, 6] = 1
    with pytest.raises(Exception):
        g.set_motion_states([2, 5], bad, 15)
    assert o.set_motion_states([2, 5], bad, 15) == E_NOT_FOUND
    # a body added mid-run joins at the next step on both
    row = scenes._row(scenes.SHAPE_SPHERE, (0.6, 0, 0), 2.0, (0.0, 14.0, 0.0), vel=(0.0, -3.0, 0.0))
    g.add_scene(scenes.make([row]))
    o.add_scene(scenes.make([row]))
    for _ in range(4):
        g.Update(15)
        o.step(15)
        _same(g, o)
    # removed bodies kept their last state and have no pairs
    keys = g.pairs()
    ends = set((keys >> np.uint64(32)).tolist()) | set((keys & np.uint64(0xffffffff)).tolist())
    assert not ends & {5, 17, 40}


def test_settings_changed_between_steps_exact():
    """Solver.IterationLimit, ForceUpdater.Gravity and the time step changed
    mid-run take effect at the next step, identically on both sides."""
    sc = scenes.make(list(scenes.pyramid(rows=6)) + list(scenes.random_pile(30, seed=3, with_ground=False)))
    g, o = run_gpu(sc, 25), run_oracle(sc, 25)
    s = g._settings
    s.iteration_limit = 4
    s.gravity[:] = [1.5, -6.0, 0.5]
    s.dt = 1.0 / 120.0
    g.set_settings(s)
    assert o.set_settings(s) == 0
    g.Solver.IterationLimit = 4  # the same through the reference-named property
    for _ in range(3):
        g.Update(20)
        o.step(20)
        _same(g, o)
    # invalid settings are refused and leave the world as it was
    bad = _lib.Settings.from_buffer_copy(g._settings)
    bad.dt = -1.0
    assert g._lib.bepu_set_settings(g._handle(), ctypes.byref(bad)) == 1
    got = type(bad)()
    assert g._lib.bepu_get_settings(g._handle(), ctypes.byref(got)) == 0
    assert got.dt == pytest.approx(1.0 / 120.0) and got.iteration_limit == 4


def test_entity_setters_through_the_space_mirror():
    """Entity.LinearVelocity / Position setters on the Python mirror after the
    body is on the device, against the oracle's single-field setters."""
    sp = Space(capacity=8)
    ground = scenes.ground()
    gsc = scenes.make([ground])
    sp.add_scene(gsc)
    box = sp.Add(Box((0.0, 0.5, 0.0), 1, 1, 1, 1.0))
    ball = sp.Add(Sphere((3.0, 0.5, 0.0), 0.5, 2.0))
    sp.Update(10)
    box.LinearVelocity = (2.0, 0.0, 0.5)
    ball.Position = (3.0, 4.0, 1.0)
    sp.Update(30)
    from oracle import refcpu
    o = refcpu.World()
    o.add_scene(gsc)
    o.add_body(0, (1, 1, 1), 1.0, np.r_[0.0, 0.5, 0.0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0], (1.0, 0.8, 0.0), 1)
    o.add_body(1, (0.5, 0, 0), 2.0, np.r_[3.0, 0.5, 0.0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0], (1.0, 0.8, 0.0), 1)
    o.step(10)
    one = np.zeros((1, 13), np.float32)
    one[0, 7:10] = (2.0, 0.0, 0.5)
    o.set_motion_states([1], one, 4)
    one[0, 0:3] = (3.0, 4.0, 1.0)
    o.set_motion_states([2], one, 1)
    o.step(30)
    assert np.array_equal(f32bits(sp.motion_states()), f32bits(o.motion_states()))
    assert box.Position == tuple(float(x) for x in o.motion_states()[1, 0:3])
    sp.Remove(ball)
    assert ball.Space is None
    with pytest.raises(ValueError):
        sp.Remove(ball)


def _lattice(n):
    """n x n x n unit boxes at spacing 1.0 on a ground: every box's AABB
    touches its 26 neighbours' (closed-interval overlap), about 11-13 pairs
    per body."""
    rows = [scenes.ground(4 * n, 1.0, 4 * n)]
    for i in range(n):
        for j in range(n):
            for k in range(n):
                rows.append(scenes._row(scenes.SHAPE_BOX, (1, 1, 1), 1.0,
                                        (i - n / 2 + 0.5, 0.5 + j, k - n / 2 + 0.5)))
    return scenes.make(rows)


def test_touching_lattice_parity_with_enough_pairs():
    """A touching 10 x 10 x 10 box lattice holds more pairs than the default
    capacity (4 per body); with the capacity given it matches the oracle."""
    sc = _lattice(10)
    o = run_oracle(sc, 6)
    npairs = len(o.pairs())
    assert npairs > 4 * len(sc)
    g = Space(capacity=len(sc), pair_capacity=2 * npairs)
    g.add_scene(sc)
    g.Update(6)
    _same(g, o)
    g.synchronize()


def test_overflow_is_sticky_and_loud():
    """With the default pair capacity the lattice overflows in its first
    step; the error stays through a multi-step call, every readback reports
    it, bepu_step refuses to run until bepu_clear_error, and it comes back as
    soon as a step overflows again."""
    sc = _lattice(10)
    g = Space(capacity=len(sc))
    g.add_scene(sc)
    L, h = g._lib, g._handle()
    assert L.bepu_step(h, 5) == 0  # queued; the overflow is seen at the next wait
    assert L.bepu_synchronize(h) == E_CAPACITY
    buf = (_lib.MotionState * len(sc))()
    n = ctypes.c_uint32()
    assert L.bepu_get_motion_states(h, buf, len(sc), ctypes.byref(n)) == E_CAPACITY
    assert L.bepu_get_pairs(h, None, 0, ctypes.byref(n)) == E_CAPACITY
    assert L.bepu_step(h, 1) == E_CAPACITY  # refused
    assert L.bepu_clear_error(h) == 0
    assert L.bepu_synchronize(h) == 0
    assert L.bepu_step(h, 1) == 0
    assert L.bepu_synchronize(h) == E_CAPACITY


def test_world_ids_stay_within_bounds():
    """Insert, insert, remove the last, insert (capacity 2): the world ids
    stay within the per-world arrays, and the worlds step as if alone."""
    desc = scenes.make([scenes._row(scenes.SHAPE_SPHERE, (0.5, 0, 0), 1.0, (0.0, 3.0, 0.0))])
    sp = Space(capacity=2)
    w1, f1 = sp.insert_world(desc)
    w2, f2 = sp.insert_world(desc)
    sp.remove_world(w2)
    w3, f3 = sp.insert_world(desc)
    assert max(w1, w2, w3) <= 3
    sp.Update(30)
    sp.synchronize()
    a = sp.motion_states_range(f1, 1)
    b = sp.motion_states_range(f3, 1)
    assert np.array_equal(f32bits(a), f32bits(b))
    with pytest.raises(Exception):
        sp.insert_world(desc)  # both slots taken


def test_async_readback_matches_the_synchronous_one():
    """bepu_request_motion_states_ranges / bepu_collect_motion_states give the
    states bepu_get_motion_states_ranges gives at the same point of the
    stream, while later steps are already queued; the ring holds 8
    outstanding readbacks and a collected ticket is gone."""
    worlds = scenes.batched_worlds(40, seed=2, shapes=(0, 1, 2))
    sp = Space(capacity=sum(len(w) for w in worlds))
    sp.register_hulls(scenes.hull_library())
    spans = [(int(sp.add_world(w)[1][0]), len(w)) for w in worlds]
    sp.Update(20)
    f = [s[0] for s in spans[::3]]
    c = [s[1] for s in spans[::3]]
    req = sp.request_motion_states_ranges(f, c)
    sp.Update(5)  # queued behind the readback
    got = np.concatenate(sp.collect_motion_states(req))
    sp2 = Space(capacity=sum(len(w) for w in worlds))
    sp2.register_hulls(scenes.hull_library())
    for w in worlds:
        sp2.add_world(w)
    sp2.Update(20)
    want = np.concatenate(sp2.motion_states_ranges(f, c))
    assert np.array_equal(f32bits(got), f32bits(want))
    with pytest.raises(Exception):
        sp.collect_motion_states(req)  # already collected
    reqs = [sp.request_motion_states_ranges(f[:1], c[:1]) for _ in range(8)]
    L, h = sp._lib, sp._handle()
    t = ctypes.c_uint32()
    one = np.array(f[:1], np.uint32)
    cnt = np.array(c[:1], np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    assert L.bepu_request_motion_states_ranges(h, one.ctypes.data_as(u32p), cnt.ctypes.data_as(u32p), 1,
                                               ctypes.byref(t)) == E_CAPACITY
    for r in reqs:
        sp.collect_motion_states(r)
    sp.synchronize()


def test_non_finite_world_is_evicted_and_the_others_go_on():
    """One sub-world of a batch is driven to a non-finite state (position x
    3.4e38 with velocity x 3e38 overflows at the next integration): it is
    listed by bepu_get_faulted_worlds, synchronous readbacks touching it
    return BEPU_E_NONFINITE, and every other world steps on bit-identical to
    the same batch without the fault.  The fault survives bepu_clear_error;
    removing the world and inserting a new one into its id starts clean."""
    E_NONFINITE = 6
    worlds = scenes.batched_worlds(24, seed=3, shapes=(0, 1, 2))
    cap = sum(len(w) for w in worlds) + 64

    def make():
        sp = Space(capacity=cap)
        sp.register_hulls(scenes.hull_library())
        spans = [sp.add_world(w) for w in worlds]
        return sp, [(int(wid), int(ids[0]), len(ids)) for wid, ids in spans]

    a, spans = make()
    b, _ = make()
    a.Update(10)
    b.Update(10)
    bad = 5
    wid, first, n = spans[bad]
    dyn = np.flatnonzero(worlds[bad]["mass"] > 0)
    body = first + int(dyn[0])
    st = a.motion_states_range(body, 1).copy()
    st[0, 0] = 3.4e38
    st[0, 7] = 3e38
    a.set_motion_states([body], st)
    assert a.faulted_worlds() == []
    a.Update(30)
    b.Update(30)
    assert a.faulted_worlds() == [wid]
    assert b.faulted_worlds() == []
    others = [s for k, s in enumerate(spans) if k != bad]
    ga = a.motion_states_ranges([s[1] for s in others], [s[2] for s in others])
    gb = b.motion_states_ranges([s[1] for s in others], [s[2] for s in others])
    for x, y in zip(ga, gb):
        assert np.array_equal(f32bits(x), f32bits(y))
    L, h = a._lib, a._handle()
    out = (_lib.MotionState * n)()
    assert L.bepu_get_motion_states_range(h, first, n, out) == E_NONFINITE
    got = np.frombuffer(out, dtype=np.float32).reshape(n, 13)
    assert not np.isfinite(got[int(dyn[0])]).all()
    cnt = ctypes.c_uint32()
    whole = (_lib.MotionState * cap)()
    assert L.bepu_get_motion_states(h, whole, cap, ctypes.byref(cnt)) == E_NONFINITE
    assert L.bepu_clear_error(h) == 0
    assert a.faulted_worlds() == [wid]
    a.remove_world(wid)
    assert a.faulted_worlds() == []
    wid2, first2 = a.insert_world(worlds[bad])
    a.Update(5)
    assert a.faulted_worlds() == []
    assert np.isfinite(a.motion_states_range(first2, n)).all()
    a.close()
    b.close()
