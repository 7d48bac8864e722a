"""Deactivation (sleeping) on the GPU against the oracle, bit for bit (run on
an MI355X).  Scenes use bodies that are not always active, so islands fall
asleep after LowVelocityTimeMinimum, are looked at within the per-step
budget in turn, and wake on new constraints, removals, setters and
Activate():

  * a mixed pile (boxes, spheres) settling to sleep on a quiet ground;
  * small box towers that sleep, then a sphere dropped on one tower wakes it
    (its frozen pairs go stale and come back the next step);
  * 150 singleton islands in zero gravity: the budget and the rotation;
  * setters, Activate() and removal mid-run on a sleeping pile;
  * a hull pile (frozen general convex pairs);
  * batched sub-worlds, each with its own rotation, against each world alone.

Compared at checkpoints: motion states (bits), activity, pair set, manifold
and impulse tables.
"""
import numpy as np
import pytest

from bepuphysics1_amd import Space, scenes
from helpers import f32bits, impulse_table, manifold_table, oracle_hulls, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

SPHERE, BOX = scenes.SHAPE_SPHERE, scenes.SHAPE_BOX


def _sleepy(sc):
    sc = sc.copy()
    sc["flags"] = 0
    return sc


def _check(g, o, what):
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states())), what + ": motion states"
    assert np.array_equal(g.activity(), o.activity()), what + ": activity"
    assert np.array_equal(g.pairs(), o.pairs()), what + ": pairs"
    assert manifold_table(g.manifolds()) == manifold_table(o.manifolds()), what + ": manifolds"
    assert impulse_table(g.impulses()) == impulse_table(o.impulses()), what + ": impulses"


def _run_both(sc, checkpoints, settings_g=None, settings_o=None, capacity=None):
    g = run_gpu(sc, 0, settings=settings_g, capacity=capacity)
    o = run_oracle(sc, 0, settings=settings_o)
    done = 0
    for c in checkpoints:
        g.Update(c - done)
        o.step(c - done)
        done = c
        _check(g, o, "step %d" % c)
    return g, o


def test_pile_settles_to_sleep_exact():
    sc = _sleepy(scenes.random_pile(60, seed=5, shapes=(BOX, SPHERE)))
    g, o = _run_both(sc, [1, 30, 61, 62, 100, 160, 240, 320, 400])
    act = o.activity()
    assert not act[sc["mass"] > 0].all()  # some islands did fall asleep
    g.close()
    o.close()


def test_towers_sleep_and_wake_on_impact_exact():
    sc = _sleepy(scenes.towers(nx=4, nz=4, height=10, jitter=0.0))
    g, o = _run_both(sc, [1, 61, 90, 200, 300], capacity=len(sc) + 1)
    assert not o.activity()[sc["mass"] > 0].any()
    # a sphere falls onto the first tower
    top = sc["state"][sc["mass"] > 0][:, 1].max()
    x0, z0 = sc["state"][1, 0], sc["state"][1, 2]
    ball = scenes.make([scenes._row(SPHERE, (0.3, 0, 0), 2.0, (x0, top + 3.0, z0), flags=0)])
    g.add_scene(ball)
    o.add_scene(ball, None)
    done = 300
    # the ball reaches the tower about 45 steps after it is added: compare
    # every step around the impact
    for c in [320, 335] + list(range(336, 366)) + [400, 500]:
        g.Update(c - done)
        o.step(c - done)
        done = c
        _check(g, o, "after impact, step %d" % c)
    g.close()
    o.close()


def test_budget_and_rotation_exact():
    from oracle import refcpu
    rows = [scenes._row(SPHERE, (0.4, 0, 0), 1.0, (3.0 * (i % 15), 0, 3.0 * (i // 15)), flags=0) for i in range(150)]
    sc = scenes.make(rows)
    from bepuphysics1_amd import _lib
    s = _lib.default_settings()
    s.gravity[1] = 0.0
    so = refcpu.default_settings()
    so.gravity[1] = 0.0
    g = run_gpu(sc, 0, settings=s)
    o = run_oracle(sc, 0, settings=so)
    for k in range(1, 70):
        g.Update(1)
        o.step(1)
        assert np.array_equal(g.activity(), o.activity()), "step %d" % k
    assert not o.activity().any()
    g.close()
    o.close()


def test_setters_activate_and_remove_on_sleeping_pile_exact():
    sc = _sleepy(scenes.random_pile(40, seed=9, shapes=(BOX, SPHERE)))
    g, o = _run_both(sc, [250])
    dyn = np.flatnonzero(sc["mass"] > 0)
    # Activate() one body, set another's velocity, remove a third
    assert o.activate([int(dyn[0])]) == 0
    g.activate([int(dyn[0])])
    st = o.motion_states()[dyn[1]:dyn[1] + 1].copy()
    st[0, 7] = 2.0
    assert o.set_motion_states([int(dyn[1])], st, fields=4) == 0
    g.set_motion_states([int(dyn[1])], st, fields=4)
    assert o.remove(int(dyn[2])) == 0
    g.Remove(int(dyn[2]))
    done = 250
    for c in (251, 252, 260, 320, 420):
        g.Update(c - done)
        o.step(c - done)
        done = c
        _check(g, o, "after edits, step %d" % c)
    g.close()
    o.close()


def test_hull_pile_sleeps_exact():
    sc = _sleepy(scenes.random_pile(30, seed=3, shapes=(BOX, SPHERE, scenes.SHAPE_CONVEX_HULL)))
    g, o = _run_both(sc, [1, 80, 200, 320])
    g.close()
    o.close()


def test_batched_sleeping_worlds_match_each_alone_and_the_oracle():
    worlds = [_sleepy(w) for w in scenes.batched_worlds(12, seed=4, nmin=8, nmax=140, shapes=(BOX, SPHERE))]
    sp = Space(capacity=sum(len(w) for w in worlds))
    spans = []
    for w in worlds:
        _, ids = sp.add_world(w)
        spans.append((int(ids[0]), len(w)))
    sp.Update(260)
    got = sp.motion_states()
    act = sp.activity()
    for k, w in enumerate(worlds):
        f, n = spans[k]
        o = run_oracle(w, 260)
        assert np.array_equal(f32bits(got[f:f + n]), f32bits(o.motion_states())), "world %d" % k
        assert np.array_equal(act[f:f + n], o.activity()), "world %d activity" % k
        o.close()
    sp.close()
