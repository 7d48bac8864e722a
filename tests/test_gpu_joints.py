"""Ball-socket joints on the GPU against the oracle, bit for bit (run on an
MI355X): BallSocketJoint (Constraints/TwoEntity/Joints/BallSocketJoint.cs)
as a constraint of the islands, the colouring and the solve beside the
contact manifold constraints.

  * chains of boxes, spheres, hulls and capsules swinging from kinematic
    anchors onto a ground and loose boxes (small and medium islands);
  * a 300-link chain lying on the ground (an island of the global path);
  * chains of bodies that may sleep: they settle and sleep as one island,
    and a moved anchor wakes them; a link removed mid-run; spring settings
    changed mid-run;
  * chain worlds batched with ordinary piles against each world alone.

Compared at checkpoints: motion states (bits), pair sets, joint impulses
(bits), and at the end the manifold and contact impulse tables.
"""
import numpy as np
import pytest

from bepuphysics1_amd import Space, scenes
from helpers import f32bits, impulse_table, manifold_table, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _both(sc, joints, settings=None):
    g, o = run_gpu(sc, 0, settings=settings), run_oracle(sc, 0, settings=settings)
    for a, b, p in joints:
        ja = g.add_ball_socket(a, b, p)
        jb = o.add_ball_socket(a, b, p)
        assert ja == jb
    return g, o


def _same(g, o, what):
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states())), what + ": motion states"
    assert np.array_equal(g.pairs(), o.pairs()), what + ": pairs"
    assert np.array_equal(f32bits(g.joint_impulses()), f32bits(o.joint_impulses())), what + ": joint impulses"


def _end(g, o):
    assert manifold_table(g.manifolds()) == manifold_table(o.manifolds())
    assert impulse_table(g.impulses()) == impulse_table(o.impulses())


def test_swinging_chains_exact():
    sc, joints = scenes.chains(6, 8, seed=3)
    g, o = _both(sc, joints)
    done = 0
    for c in (1, 30, 60, 120, 180):
        g.Update(c - done)
        o.step(c - done)
        done = c
        _same(g, o, "step %d" % c)
    _end(g, o)
    assert np.linalg.norm(o.joint_impulses(), axis=1).min() > 0
    g.close()


def test_long_chain_on_the_global_path_exact():
    """300 links in one island: more than a workgroup's 256 constraints, so
    it is coloured and solved on the global path."""
    rows = [scenes.ground(400.0, 1.0, 20.0)]
    joints = []
    for k in range(300):
        rows.append(scenes._row(scenes.SHAPE_BOX, (0.5, 0.3, 0.3), 1.0, (-150.0 + 0.6 * k, 0.15, 0.0)))
        if k:
            joints.append((k, k + 1, (-150.0 + 0.6 * k - 0.3, 0.15, 0.0)))
    # lift one end so the chain is pulled up and over
    rows.append(scenes._row(scenes.SHAPE_BOX, (0.4, 0.4, 0.4), 0.0, (-150.0, 3.0, 0.0), flags=0))
    joints.append((301, 1, (-150.3, 0.6, 0.0)))
    sc = scenes.make(rows)
    g, o = _both(sc, joints)
    done = 0
    for c in (1, 20, 60):
        g.Update(c - done)
        o.step(c - done)
        done = c
        _same(g, o, "step %d" % c)
        assert g.stats().global_manifolds > 256  # the chain's constraints took the global path
    _end(g, o)
    g.close()


def test_sleeping_chains_removal_and_settings_exact():
    sc, joints = scenes.chains(4, 5, seed=7, flags=0)
    g, o = _both(sc, joints)
    done = 0
    for c in (40, 120, 240):
        g.Update(c - done)
        o.step(c - done)
        done = c
        _same(g, o, "step %d" % c)
        assert np.array_equal(g.activity(), o.activity())
    # anchors are kinematic bodies 1, 7, 13, 19: move one, which wakes its chain
    ms = o.motion_states()
    st = ms[7:8].copy()
    st[0, 7:10] = (0.3, 0.0, 0.0)
    g.set_motion_states([7], st, 4)
    assert o.set_motion_states([7], st, 4) == 0
    g.Update(30)
    o.step(30)
    _same(g, o, "after the anchor moved")
    assert np.array_equal(g.activity(), o.activity())
    # a link removed mid-chain ends its two joints
    g.Remove(10)
    assert o.remove(10) == 0
    # softer joints on the first chain
    for j in range(5):
        g.set_joint_settings(j, 20000.0, 5000.0, 2.0)
        assert o.set_joint_settings(j, 20000.0, 5000.0, 2.0) == 0
    g.Update(60)
    o.step(60)
    _same(g, o, "end")
    assert np.array_equal(g.activity(), o.activity())
    _end(g, o)
    g.close()


def test_batched_chain_worlds_match_alone():
    worlds = [scenes.chains(3, int(n), seed=int(s)) for n, s in ((4, 1), (6, 2), (3, 3))]
    piles = [scenes.random_pile(20, seed=4, shapes=(0, 1, 2, 3))]
    sp = Space(capacity=sum(len(w[0]) for w in worlds) + sum(len(p) for p in piles))
    sp.register_hulls(scenes.hull_library())
    wids = []
    for sc, joints in worlds:
        wid, ids = sp.add_world(sc)
        wids.append(wid)
        first = int(ids[0])
        for a, b, p in joints:
            sp.add_ball_socket(first + a, first + b, p)
    for p in piles:
        sp.add_world(p)
    sp.Update(60)
    ms = sp.motion_states()
    ji = sp.joint_impulses()
    jbase = 0
    for k, (sc, joints) in enumerate(worlds):
        first, count = sp.world_range(wids[k])
        alone = run_gpu(sc, 0)
        for a, b, p in joints:
            alone.add_ball_socket(a, b, p)
        alone.Update(60)
        assert np.array_equal(f32bits(ms[first:first + count]), f32bits(alone.motion_states())), "world %d" % k
        assert np.array_equal(f32bits(ji[jbase:jbase + len(joints)]), f32bits(alone.joint_impulses()))
        jbase += len(joints)
        alone.close()
    sp.close()
