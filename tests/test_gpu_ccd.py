"""Continuous collision detection on the GPU against the oracle, bit for bit
(run on an MI355X): ContinuousPositionUpdater (ContinuousPositionUpdater.cs:
76-160), ConvexPairHandler.UpdateTimeOfImpact (ConvexPairHandler.cs:37-111)
and GJKToolbox.CCDSphereCast (GJKToolbox.cs:470-495).

  * a shooting gallery: fast boxes, spheres, hulls and capsules, continuous,
    passive and discrete, fired at two thin kinematic walls;
  * position update modes changed between steps (Entity.PositionUpdateMode)
    on bodies that may fall asleep;
  * gallery worlds batched with ordinary piles, against each world alone;
  * the same gallery replayed from a captured step graph.

Compared: motion states (bits) and pair sets at checkpoints, manifold and
impulse tables at the end.
"""
import numpy as np
import pytest

from bepuphysics1_amd import Space, scenes
from helpers import f32bits, impulse_table, manifold_table, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _same(g, o, what):
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states())), what + ": motion states"
    assert np.array_equal(g.pairs(), o.pairs()), what + ": pairs"


def _end(g, o):
    assert manifold_table(g.manifolds()) == manifold_table(o.manifolds())
    assert impulse_table(g.impulses()) == impulse_table(o.impulses())


def test_shooting_gallery_exact():
    sc = scenes.shooting_gallery(96, seed=1)
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    g.Update(1)
    o.step(1)
    _same(g, o, "step 1")
    # after the first step every continuous projectile is still between the walls
    cont = (sc["flags"] & scenes.FLAG_CCD_CONTINUOUS) != 0
    x = g.motion_states()[:, 0]
    assert ((x[cont] > 0.0) & (x[cont] < 12.0)).all()
    disc = (sc["flags"] & (scenes.FLAG_CCD_CONTINUOUS | scenes.FLAG_CCD_PASSIVE)) == 0
    disc[:3] = False
    done = 1
    for c in (10, 30, 60, 90):
        g.Update(c - done)
        o.step(c - done)
        done = c
        _same(g, o, "step %d" % c)
    _end(g, o)
    x = g.motion_states()[:, 0]
    assert ((x[disc] < 0.0) | (x[disc] > 12.0)).mean() > 0.5  # most discrete ones went through
    g.close()


def test_mode_changes_between_steps_exact():
    """Modes switched between steps on both sides (continuous -> passive ->
    discrete and back), in a gallery whose bodies may sleep: pairs keep their
    time of impact while no continuous body asks for it."""
    sc = scenes.shooting_gallery(48, seed=4)
    sc["flags"] &= ~np.uint32(scenes.FLAG_ALWAYS_ACTIVE)
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    rng = np.random.default_rng(2)
    n = len(sc)
    for call in range(8):
        steps = int(rng.integers(1, 12))
        g.Update(steps)
        o.step(steps)
        _same(g, o, "call %d" % call)
        ids = [int(i) for i in rng.choice(np.arange(3, n), 8, replace=False)]
        mode = int(call % 3)
        g.set_position_update_mode(ids, mode)
        assert o.set_position_update_mode(ids, mode) == 0
    g.Update(120)
    o.step(120)
    _same(g, o, "end")
    assert np.array_equal(g.activity(), o.activity())
    _end(g, o)
    g.close()


def test_batched_gallery_worlds_match_alone():
    """Gallery worlds and ordinary piles in one batch: each equals itself
    stepped alone (the piles have no passive or continuous body; the batch
    runs the continuous pass for all)."""
    worlds = [scenes.shooting_gallery(int(n), seed=int(s)) for n, s in ((24, 7), (40, 8), (16, 9))]
    worlds += [scenes.random_pile(30, seed=3, shapes=(0, 1, 2, 3)), scenes.random_pile(12, seed=5)]
    sp = Space(capacity=sum(len(w) for w in worlds))
    sp.register_hulls(scenes.hull_library())
    wids = [sp.add_world(w)[0] for w in worlds]
    sp.Update(50)
    ms = sp.motion_states()
    imp = impulse_table(sp.impulses())
    for k, w in enumerate(worlds):
        first, count = sp.world_range(wids[k])
        alone = run_gpu(w, 50)
        assert np.array_equal(f32bits(ms[first:first + count]), f32bits(alone.motion_states())), "world %d" % k
        shift = (first << 32) | first
        for key, v in impulse_table(alone.impulses()).items():
            assert imp[key + shift] == v, "world %d impulses" % k
        alone.close()
    o = run_oracle(worlds[1], 50)
    first, count = sp.world_range(wids[1])
    assert np.array_equal(f32bits(ms[first:first + count]), f32bits(o.motion_states()))
    sp.close()


def test_gallery_replayed_from_step_graph_matches():
    sc = scenes.shooting_gallery(64, seed=11)
    g = run_gpu(sc, 0)
    g.set_tuning(step_graph=1)
    e = run_gpu(sc, 0)
    for n in (1, 2, 5, 9, 23):
        g.Update(n)
        e.Update(n)
        _same(g, e, "after %d" % n)
    g.close()
    e.close()
