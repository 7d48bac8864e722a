"""Steps replayed from captured HIP graphs against steps launched one by one
(run on an MI355X).

A world captures the launches of a step once per parity of its step count
and replays them; it captures again when its device description or launch
configuration changes.  Between the calls that step the worlds here sit the
edits a caller makes (bodies added and removed, state setters, Activate,
settings, sub-worlds inserted and removed), readbacks, stage-timed steps
(launched one by one) and configuration changes, so that replays follow
eager work on the same stream and recaptures happen mid-run.  Every call's
results are compared bit for bit with a world that launches every step one
by one, and at the end with the oracle.
"""
import numpy as np
import pytest

from bepuphysics1_amd import Space, scenes
from helpers import f32bits, impulse_table, manifold_table, oracle_hulls, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert np.array_equal(f32bits(a.motion_states()), f32bits(b.motion_states())), what + ": motion states"
    assert np.array_equal(a.pairs(), b.pairs()), what + ": pairs"
    assert impulse_table(a.impulses()) == impulse_table(b.impulses()), what + ": impulses"


def _edits(rng, spaces, oracle, call):
    """One edit per call, the same on every space and the oracle."""
    kind = call % 5
    if kind == 0:
        ids = [int(i) for i in rng.choice(np.arange(1, 40), 2, replace=False)]
        ms = oracle.motion_states()[ids].copy()
        ms[:, 7:10] += rng.normal(0, 1.5, (2, 3)).astype(np.float32)
        for s in spaces:
            s.set_motion_states(ids, ms, 4)
        assert oracle.set_motion_states(ids, ms, 4) == 0
    elif kind == 1:
        row = scenes._row(scenes.SHAPE_SPHERE, (0.4, 0, 0), 1.0, (float(rng.uniform(-3, 3)), 9.0, 0.0),
                          vel=(0.0, -2.0, 0.0), flags=0)
        for s in spaces:
            s.add_scene(scenes.make([row]))
        oracle.add_scene(scenes.make([row]))
    elif kind == 2:
        bid = 41 + call
        for s in spaces:
            s.Remove(bid)
        assert oracle.remove(bid) == 0
    elif kind == 3:
        ids = list(range(2, 40, 7))  # none of them removed
        for s in spaces:
            s.activate(ids)
        assert oracle.activate(ids) == 0
    else:
        for s in spaces:
            st = s._settings
            st.iteration_limit = 6 + call % 4
            s.set_settings(st)
        assert oracle.set_settings(spaces[0]._settings) == 0


def test_graph_steps_match_eager_steps_through_edits_and_recaptures():
    """A sleeping pile of boxes, spheres, hulls and capsules: calls of 1 to 9
    steps with an edit, readbacks and now and then stage timing or another
    launch configuration between them.  The graph world equals the eager
    world after every call and the oracle at the end."""
    sc = scenes.random_pile(70, seed=23, shapes=(0, 1, 2, 3))
    sc["flags"] = 0  # bodies may fall asleep: the deactivation stages run too
    cap = len(sc) + 16
    g = run_gpu(sc, 0, capacity=cap)
    e = run_gpu(sc, 0, capacity=cap)
    g.set_tuning(step_graph=1)
    assert g.tuning()["step_graph"] == 1 and e.tuning()["step_graph"] == 0
    o = run_oracle(sc, 0)
    rng = np.random.default_rng(4)
    total = 0
    for call in range(24):
        n = int(rng.integers(1, 10))
        if call % 6 == 5:
            g.set_timing(True)  # these steps are launched one by one with stage events
        g.Update(n)
        e.Update(n)
        o.step(n)
        total += n
        g.set_timing(False)
        _same(g, e, "call %d (%d steps)" % (call, total))
        if call % 8 == 3:
            # another configuration: both parities are captured again
            g.set_tuning(narrowphase_block=64 if call % 16 == 3 else 128, convex_side_stream=call % 16 == 3)
        _edits(rng, (g, e), o, call)
    g.Update(30)
    e.Update(30)
    o.step(30)
    _same(g, e, "end")
    assert np.array_equal(g.activity(), o.activity())
    assert manifold_table(g.manifolds()) == manifold_table(o.manifolds())
    assert impulse_table(g.impulses()) == impulse_table(o.impulses())
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))
    for s in (g, e):
        s.close()


def test_graph_steps_of_batched_worlds_inserted_and_removed():
    """Sub-worlds inserted into and removed from a batch between calls (each
    change captures the step again): the graph batch equals the eager batch
    after every call, and each remaining world equals itself stepped alone."""
    worlds = scenes.batched_worlds(40, seed=8, nmin=8, nmax=60, shapes=(0, 1, 2, 3))
    cap = sum(len(w) for w in worlds) + 400
    spaces = []
    for graph in (1, 0):
        sp = Space(capacity=cap)
        sp.register_hulls(scenes.hull_library())
        sp.set_tuning(step_graph=graph)
        spaces.append(sp)
    wids = [[s.insert_world(w)[0] for w in worlds[:30]] for s in spaces]
    assert wids[0] == wids[1]
    live = list(wids[0])
    steps_of = {w: 0 for w in live}
    scene_of = dict(zip(live, worlds[:30]))
    nxt = 30
    rng = np.random.default_rng(9)
    for call in range(12):
        n = int(rng.integers(1, 6))
        for s in spaces:
            s.Update(n)
        for w in live:
            steps_of[w] += n
        _same(spaces[0], spaces[1], "call %d" % call)
        if call % 3 == 0 and nxt < len(worlds):
            new = [s.insert_world(worlds[nxt])[0] for s in spaces]
            assert new[0] == new[1]
            live.append(new[0])
            steps_of[new[0]] = 0
            scene_of[new[0]] = worlds[nxt]
            nxt += 1
        if call % 3 == 1:
            gone = live.pop(int(rng.integers(0, len(live))))
            for s in spaces:
                s.remove_world(gone)
    for w in live[:6]:
        first, count = spaces[0].world_range(w)
        alone = run_gpu(scene_of[w], steps_of[w])
        assert np.array_equal(f32bits(spaces[0].motion_states()[first:first + count]),
                              f32bits(alone.motion_states())), "world %d" % w
        alone.close()
    for s in spaces:
        s.close()


def test_graph_replay_time_for_a_small_world():
    """A 70-body pile (the plain single-world case): reports ms per step with
    and without graph replay.  The device is busy for most of each step either
    way (15 kernels whose own durations add up to about 0.23 ms), so the
    replay is within 10% of launching one by one."""
    import time
    sc = scenes.random_pile(70, seed=3, shapes=(0, 1, 2, 3))
    out = {}
    for graph in (0, 1, 0, 1):
        sp = run_gpu(sc, 0)
        sp.set_tuning(step_graph=graph)
        sp.Update(20)
        sp.synchronize()
        t0 = time.perf_counter()
        sp.Update(400)
        sp.synchronize()
        out.setdefault(graph, []).append((time.perf_counter() - t0) * 1e3 / 400)
        sp.close()
    eager, replay = min(out[0]), min(out[1])
    print("70-body pile: %.4f ms/step launched one by one, %.4f ms/step replayed" % (eager, replay))
    assert replay <= 1.10 * eager
