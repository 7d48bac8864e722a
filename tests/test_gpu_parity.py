"""GPU library vs oracle on identical seeded scenes (run on an MI355X).

Bar (from the task's parity contract):
  * broadphase pair sets: bit-exact;
  * contact counts: exact; positions, normals, depths: within 1e-4 relative;
  * accumulated impulses per step: within 1e-4 relative, constraints
    compared in the colour order both sides use;
  * stacks stay stable.
The library is expected to be bit-identical to the oracle for boxes and
spheres (same operation order, no FMA contraction, IEEE division/sqrt), so
most checks below are exact; the tolerance is the stated fallback.
"""
import numpy as np
import pytest

from bepuphysics1_amd import scenes
from helpers import f32bits, impulse_table, manifold_table, rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _compare_manifolds(g, o):
    mg, mo = manifold_table(g.manifolds()), manifold_table(o.manifolds())
    assert set(mg) == set(mo), "manifold key sets differ: %d vs %d" % (len(mg), len(mo))
    worst = 0.0
    for k in mo:
        cg, ag, bg, lg = mg[k]
        co, ao, bo, lo = mo[k]
        assert cg == co, "contact count differs for pair %x" % k
        assert (ag, bg) == (ao, bo)
        ids_g = sorted(c[0] for c in lg)
        ids_o = sorted(c[0] for c in lo)
        assert ids_g == ids_o
        for cgc, coc in zip(sorted(lg), sorted(lo)):
            worst = max(worst, rel_err(cgc[1], coc[1]).max(), rel_err(cgc[2], coc[2]).max(),
                        rel_err(cgc[3], coc[3]).max())
    assert worst <= TOL, worst
    return worst


def _compare_impulses(g, o):
    ig, io = impulse_table(g.impulses()), impulse_table(o.impulses())
    assert set(ig) == set(io)
    worst = 0.0
    for k in io:
        col_g, n_g, pen_g, fr_g, sl_g, tw_g = ig[k]
        col_o, n_o, pen_o, fr_o, sl_o, tw_o = io[k]
        assert col_g == col_o, "colour differs for %x" % k
        assert n_g == n_o
        worst = max(worst, rel_err(pen_g, pen_o).max(initial=0), rel_err(fr_g, fr_o).max(initial=0),
                    rel_err(sl_g, sl_o).max(), rel_err([tw_g], [tw_o]).max())
    assert worst <= TOL, worst
    return worst


def test_pyramid_first_step_exact():
    sc = scenes.pyramid()
    g, o = run_gpu(sc, 1), run_oracle(sc, 1)
    assert np.array_equal(g.pairs(), o.pairs())
    _compare_manifolds(g, o)
    _compare_impulses(g, o)
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))


def test_pyramid_300_steps():
    """C1: 55-box pyramid, 300 steps, 10 iterations, warm starting."""
    sc = scenes.pyramid()
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    for chunk in range(6):
        g.Update(50)
        o.step(50)
        assert np.array_equal(g.pairs(), o.pairs()), "pairs differ after %d steps" % (50 * (chunk + 1))
        _compare_impulses(g, o)
    msg, mso = g.motion_states(), o.motion_states()
    assert np.abs(msg - mso).max() < 1e-4
    st = g.stats()
    assert st.max_penetration < 0.05
    # every box still stands where it started (within 5 cm)
    assert np.abs(msg[1:, 0:3] - sc["state"][1:, 0:3]).max() < 0.2


def test_box_sphere_pile_steps():
    sc = scenes.random_pile(120, seed=3)
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    for _ in range(4):
        g.Update(30)
        o.step(30)
        assert np.array_equal(g.pairs(), o.pairs())
    _compare_manifolds(g, o)
    _compare_impulses(g, o)


@pytest.mark.parametrize("broadphase", ["grid", "tree"])
def test_stress_broadphase_exact(broadphase):
    sc = scenes.stress(4096, seed=11)
    g, o = run_gpu(sc, 1, broadphase=broadphase), run_oracle(sc, 1)
    pg, po = g.pairs(), o.pairs()
    assert len(po) > 1000
    assert np.array_equal(pg, po)
    _compare_manifolds(g, o)


@pytest.mark.parametrize("broadphase", ["grid", "tree"])
@pytest.mark.parametrize("offset", [(0.0, 0.0, 0.0), (-3.0e4, 1.0e3, 5.0e4)])
def test_mixed_sizes_broadphase_exact(broadphase, offset):
    """Sizes over three orders of magnitude (several grid levels), static
    bodies, and a scene far from the origin: pair set equals the oracle's
    brute force exactly."""
    sc = scenes.mixed_sizes(3000, seed=5, offset=offset)
    g = run_gpu(sc, 1, broadphase=broadphase)
    o = run_oracle(sc, 0)
    o.set_bruteforce_broadphase(True)
    o.step(1)
    pg, po = g.pairs(), o.pairs()
    assert len(po) > 500
    assert np.array_equal(pg, po)


def test_pyramid_tree_broadphase():
    sc = scenes.pyramid()
    g, o = run_gpu(sc, 30, broadphase="tree"), run_oracle(sc, 30)
    assert np.array_equal(g.pairs(), o.pairs())
    _compare_impulses(g, o)
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))


def test_towers_small():
    sc = scenes.towers(nx=8, nz=8, height=20, seed=1)
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    g.Update(60)
    o.step(60)
    assert np.array_equal(g.pairs(), o.pairs())
    _compare_manifolds(g, o)
    _compare_impulses(g, o)
    ms = g.motion_states()
    assert np.abs(ms[1:, 0] - sc["state"][1:, 0]).max() < 0.1


def test_large_island_global_path():
    """A 28-row pyramid (406 boxes) is one island of ~800 constraints: it is
    solved by the cooperative global path and must match the oracle exactly."""
    sc = scenes.pyramid(rows=28)
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    for _ in range(3):
        g.Update(10)
        o.step(10)
        assert np.array_equal(g.pairs(), o.pairs())
        _compare_impulses(g, o)
    st = g.stats()
    assert st.global_manifolds == st.manifolds and st.manifolds > 256
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))


def test_mixed_tiers_one_world():
    """Small (towers), medium (a 10-row pyramid) and large (a 28-row pyramid)
    islands in one world, side by side."""
    t = scenes.towers(nx=4, nz=4, height=6, seed=2)
    p1 = scenes.pyramid(rows=10)[1:]
    p2 = scenes.pyramid(rows=28)[1:]
    p1["state"][:, 2] += 20.0
    p2["state"][:, 2] -= 20.0
    sc = np.concatenate([scenes.make([scenes.ground(width=200.0, length=200.0)]), t[1:], p1, p2])
    g, o = run_gpu(sc, 0), run_oracle(sc, 0)
    g.Update(15)
    o.step(15)
    assert np.array_equal(g.pairs(), o.pairs())
    _compare_impulses(g, o)
    st = g.stats()
    assert 0 < st.global_manifolds < st.manifolds
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))


def test_batched_worlds_match_alone():
    """24 independent piles stepped as sub-worlds of one batch: each world's
    state and impulses are bit-identical to the same world stepped alone on
    the GPU, its pairs never cross worlds, and a sample matches the oracle."""
    from bepuphysics1_amd import Space
    worlds = scenes.batched_worlds(24, seed=7, nmin=8, nmax=64)
    steps = 40
    sp = Space(capacity=sum(len(w) for w in worlds))
    for sc in worlds:
        sp.add_world(sc)
    sp.Update(steps)
    ms_b, pairs_b = sp.motion_states(), sp.pairs()
    imp_b = impulse_table(sp.impulses())
    lo, hi = (pairs_b >> 32).astype(np.int64), (pairs_b & 0xffffffff).astype(np.int64)
    for k, sc in enumerate(worlds):
        first, count = sp.world_range(k)
        assert count == len(sc)
        inside_lo = (lo >= first) & (lo < first + count)
        inside_hi = (hi >= first) & (hi < first + count)
        assert np.array_equal(inside_lo, inside_lo & inside_hi) or not inside_lo.any()
        mine = pairs_b[inside_lo] - np.uint64((first << 32) | first)
        g = run_gpu(sc, steps)
        assert np.array_equal(np.sort(mine), g.pairs()), "world %d pairs" % k
        assert np.array_equal(f32bits(ms_b[first:first + count]), f32bits(g.motion_states())), "world %d" % k
        ia = impulse_table(g.impulses())
        shift = (first << 32) | first
        for key, v in ia.items():
            assert imp_b[key + shift] == v, "world %d impulses" % k
        g.close()
        if k < 4:
            o = run_oracle(sc, steps)
            assert np.array_equal(f32bits(ms_b[first:first + count]), f32bits(o.motion_states())), "oracle %d" % k
    # no pair joins two worlds
    wid = np.searchsorted(np.array([sp.world_range(k)[0] for k in range(len(worlds))]), lo, side="right")
    wid2 = np.searchsorted(np.array([sp.world_range(k)[0] for k in range(len(worlds))]), hi, side="right")
    assert np.array_equal(wid, wid2)


def test_served_continuous_batching_matches_alone():
    """Ten worlds arrive at different ticks into a server whose capacity holds
    only a few at a time, so later worlds reuse the id ranges of retired ones
    while others are mid-run.  Every result is bit-identical to stepping the
    same world alone, and two of them to the oracle."""
    from bepuphysics1_amd.serve import Server
    rng = np.random.default_rng(3)
    reqs = []
    for _ in range(10):
        n = int(rng.integers(10, 60))
        reqs.append((int(rng.integers(0, 40)), scenes.random_pile_fast(n, rng), int(rng.integers(5, 30))))
    reqs.sort(key=lambda r: r[0])
    srv = Server(capacity=200)
    tick = 0
    while reqs or srv.queue or srv.resident:
        while reqs and reqs[0][0] <= tick:
            _, sc, st = reqs.pop(0)
            srv.submit(sc, st)
        srv.admit()
        srv.tick()
        tick += 1
        assert tick < 5000
    assert len(srv.completed) == 10
    firsts = [r.first for r in srv.completed]
    assert len(set(firsts)) < len(firsts), "no id range was reused"
    for k, r in enumerate(srv.completed):
        g = run_gpu(r.scene, r.steps)
        assert np.array_equal(f32bits(r.result), f32bits(g.motion_states())), "request %d" % r.rid
        g.close()
        if k < 2:
            o = run_oracle(r.scene, r.steps)
            assert np.array_equal(f32bits(r.result), f32bits(o.motion_states()))
    srv.close()
