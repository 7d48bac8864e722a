"""General convex pairs (hulls and capsules) on the GPU vs the oracle, and
in the batched and served forms vs the same world alone (run on an MI355X).

The device runs the same GJK / MPR / manifold arithmetic as the oracle in
the same operation order, so states, pairs, manifolds and impulses are
compared bit for bit; the 1e-4 relative tolerance of the manifold and
impulse comparisons is the stated fallback.
"""
import numpy as np
import pytest

from bepuphysics1_amd import scenes
from helpers import f32bits, impulse_table, run_gpu, run_oracle
from test_gpu_parity import _compare_impulses, _compare_manifolds

pytestmark = pytest.mark.gpu

ALL_SHAPES = (scenes.SHAPE_BOX, scenes.SHAPE_SPHERE, scenes.SHAPE_CONVEX_HULL, scenes.SHAPE_CAPSULE)


def test_convex_pile_exact():
    """A pile of boxes, spheres, hulls and capsules dropped on a ground: every
    30 steps the pair sets match, and at the end manifolds, impulses and
    motion states are bit-identical."""
    sc = scenes.random_pile(100, seed=5, shapes=ALL_SHAPES)
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    for _ in range(4):
        g.Update(30)
        o.step(30)
        assert np.array_equal(g.pairs(), o.pairs())
        assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))
    _compare_manifolds(g, o)
    _compare_impulses(g, o)
    kinds = {m.kind for m in o.manifolds() if m.count}
    assert 3 in kinds  # general convex manifolds took part


def test_stress_hull_mix_first_step():
    """The stress mix (50% boxes, 25% spheres, 25% hulls) for one step: pairs,
    manifolds (hull pairs through GJK / MPR) and states bit-identical."""
    sc = scenes.stress(4096, seed=12, shapes=scenes.STRESS_MIX)
    g, o = run_gpu(sc, 1), run_oracle(sc, 1)
    assert np.array_equal(g.pairs(), o.pairs())
    _compare_manifolds(g, o)
    _compare_impulses(g, o)
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))
    assert sum(1 for m in o.manifolds() if m.count and m.kind == 3) > 50


def test_hull_and_capsule_stack_settles():
    """Hulls and capsules piled on each other for 300 steps: bit-identical to
    the oracle throughout, nothing sinks through the ground and contacts stay
    shallow."""
    sc = scenes.random_pile(40, seed=9, shapes=(scenes.SHAPE_CONVEX_HULL, scenes.SHAPE_CAPSULE), drop=(1.0, 4.0))
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    for _ in range(3):
        g.Update(100)
        o.step(100)
        assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))
    _compare_impulses(g, o)
    ms = g.motion_states()
    assert np.isfinite(ms).all()
    assert ms[1:, 1].min() > -0.1  # nothing fell through the ground
    assert g.stats().max_penetration < 0.2


def test_batched_hull_worlds_match_alone():
    """Worlds with hulls stepped as sub-worlds of one batch are bit-identical
    to each world alone; the hull library is registered once per handle."""
    from bepuphysics1_amd import Space
    worlds = scenes.batched_worlds(12, seed=21, nmin=8, nmax=48, shapes=ALL_SHAPES)
    steps = 30
    sp = Space(capacity=sum(len(w) for w in worlds))
    sp.register_hulls(scenes.hull_library())
    for sc in worlds:
        sp.add_world(sc)
    sp.Update(steps)
    ms_b = sp.motion_states()
    imp_b = impulse_table(sp.impulses())
    for k, sc in enumerate(worlds):
        first, count = sp.world_range(k)
        g = run_gpu(sc, steps)
        assert np.array_equal(f32bits(ms_b[first:first + count]), f32bits(g.motion_states())), "world %d" % k
        shift = (first << 32) | first
        for key, v in impulse_table(g.impulses()).items():
            assert imp_b[key + shift] == v, "world %d impulses" % k
        g.close()
        if k < 2:
            o = run_oracle(sc, steps)
            assert np.array_equal(f32bits(ms_b[first:first + count]), f32bits(o.motion_states()))


def test_served_hull_worlds_match_alone():
    """Continuous batching with hull and capsule worlds: results equal each
    world stepped alone."""
    from bepuphysics1_amd.serve import Server
    rng = np.random.default_rng(17)
    reqs = []
    for _ in range(8):
        n = int(rng.integers(10, 50))
        reqs.append((int(rng.integers(0, 30)), scenes.random_pile_fast(n, rng, sphere_fraction=0.25,
                                                                       hull_fraction=0.4), int(rng.integers(5, 25))))
    reqs.sort(key=lambda r: r[0])
    srv = Server(capacity=160, hull_shapes=scenes.hull_library())
    tick = 0
    while reqs or srv.queue or srv.resident:
        while reqs and reqs[0][0] <= tick:
            _, sc, st = reqs.pop(0)
            srv.submit(sc, st)
        srv.admit()
        srv.tick()
        tick += 1
        assert tick < 5000
    assert len(srv.completed) == 8
    for r in srv.completed:
        g = run_gpu(r.scene, r.steps)
        assert np.array_equal(f32bits(r.result), f32bits(g.motion_states())), "request %d" % r.rid
        g.close()
    srv.close()


def test_launch_configurations_give_identical_results():
    """Every launch configuration the autotuner chooses from steps a world to
    the same bits, and the autotuner leaves a valid one set."""
    from bepuphysics1_amd import Space
    worlds = scenes.batched_worlds(24, seed=31, nmin=8, nmax=120, shapes=ALL_SHAPES)
    configs = [dict(narrowphase_block=64, convex_lds=0, convex_side_stream=0, small_island_max=32, step_graph=0),
               dict(narrowphase_block=256, convex_lds=1, convex_side_stream=1, small_island_max=64, step_graph=1),
               dict(narrowphase_block=128, convex_lds=1, convex_side_stream=0, small_island_max=32, step_graph=1),
               dict(narrowphase_block=128, convex_lds=2, convex_side_stream=1, small_island_max=64, step_graph=0)]
    results = []
    for cfg in configs:
        sp = Space(capacity=sum(len(w) for w in worlds))
        sp.register_hulls(scenes.hull_library())
        for w in worlds:
            sp.add_world(w)
        sp.set_tuning(**cfg)
        sp.Update(60)
        assert sp.tuning() == cfg
        results.append((f32bits(sp.motion_states()), impulse_table(sp.impulses())))
        sp.close()
    for ms, imp in results[1:]:
        assert np.array_equal(ms, results[0][0])
        assert imp == results[0][1]
    # The tuner steps the world with the candidates: same bits as plain steps.
    sp = Space(capacity=sum(len(w) for w in worlds))
    sp.register_hulls(scenes.hull_library())
    for w in worlds:
        sp.add_world(w)
    sp.Update(10)
    chosen = sp.autotune(trial_steps=1)
    # 2 rounds x (1 untimed + 1 timed step) per candidate: convex_lds 3,
    # convex_side_stream 2, narrowphase_block 3, small_island_max 2, and
    # step_graph 2 when the steps are short enough for replay to be tried
    assert chosen["steps"] in (2 * 2 * (3 + 2 + 3 + 2), 2 * 2 * (3 + 2 + 3 + 2 + 2)) and chosen["ms_per_step"] > 0
    assert chosen["narrowphase_block"] in (64, 128, 256) and chosen["small_island_max"] in (32, 64)
    sp.Update(60 - 10 - chosen["steps"])
    assert np.array_equal(f32bits(sp.motion_states()), results[0][0])
    sp.close()
