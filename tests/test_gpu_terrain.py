"""Terrains on the GPU vs the oracle (run on an MI355X).

The device scans each terrain's grid under the convex's query box, runs the
triangle pair testers and the mesh manifold's boundary handling and reduction
in the oracle's operation order, so pair sets, motion states, manifolds and
impulses are compared bit for bit.
"""
import numpy as np
import pytest

from bepuphysics1_amd import scenes
from helpers import f32bits, run_gpu, run_oracle
from test_gpu_parity import _compare_impulses, _compare_manifolds

pytestmark = pytest.mark.gpu

ALL_SHAPES = (scenes.SHAPE_BOX, scenes.SHAPE_SPHERE, scenes.SHAPE_CONVEX_HULL, scenes.SHAPE_CAPSULE)


def _run_both(sc, meshes, rounds, steps):
    g, o = run_gpu(sc, 0, meshes=meshes), run_oracle(sc, 0, meshes=meshes)
    for r in range(rounds):
        g.Update(steps)
        o.step(steps)
        print("round", r, "pairs", len(o.pairs()), flush=True)
        assert np.array_equal(g.pairs(), o.pairs()), r
        assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states())), r
    return g, o


@pytest.mark.parametrize("organization", [0, 1])
def test_terrain_pile_exact(organization):
    """Every shape dropped on a bumpy terrain: pair sets and states
    bit-identical every 30 steps, manifolds and impulses at the end."""
    sc, meshes = scenes.terrain_pile(48, seed=3, organization=organization)
    g, o = _run_both(sc, meshes, 4, 30)
    _compare_manifolds(g, o)
    _compare_impulses(g, o)
    kinds = {m.kind for m in o.manifolds() if m.count}
    assert {4, 5} <= kinds


@pytest.mark.parametrize("improve", [True, False])
def test_flat_terrain_sliding(improve):
    """Bodies sliding across the quads of a flat terrain, with and without the
    improved boundary handling."""
    sc, meshes = scenes.terrain_pile(32, seed=11, amplitude=0.0, spacing=1.5, improve_boundary=improve)
    rng = np.random.default_rng(4)
    sc["state"][1:, 7] = rng.uniform(-3, 3, len(sc) - 1)
    g, o = _run_both(sc, meshes, 3, 40)
    _compare_manifolds(g, o)
    _compare_impulses(g, o)


def test_transformed_terrains_and_a_mesh():
    """A mirrored terrain, a sheared and tilted one, and a static mesh in one
    scene; bodies fall across all three."""
    t0 = scenes.heightfield_terrain(9, 9, 2.0, amplitude=0.8, seed=2, center=(-9, 0),
                                    linear=np.diag([-2.0, 1.0, 2.0]))
    t0 = scenes.TerrainDesc(t0.heights, 0, (t0.transform[0], t0.transform[1] + np.array([16, 0, 0], np.float32)))
    c, s = np.cos(0.2), np.sin(0.2)
    lin = np.array([[1.5 * c, 1.5 * s, 0], [-s, c, 0.05], [0.1, 0, 1.5]], np.float32)
    t1 = scenes.heightfield_terrain(8, 8, 1.5, amplitude=0.5, seed=3, center=(8, 0), organization=1, linear=lin)
    V, I, _ = scenes.heightfield_mesh(cells=6, size=10.0, amplitude=0.3, seed=4, center=(0, 10))
    meshes = [t0, t1, (V, I, None)]
    rows = [scenes.terrain(0), scenes.terrain(1, improve_boundary=False), scenes.static_mesh(2)]
    rng = np.random.default_rng(9)
    for k in range(40):
        shp = int(rng.choice(ALL_SHAPES))
        pos = (rng.uniform(-14, 12), rng.uniform(3, 8), rng.uniform(-6, 12))
        rows.append(scenes._row(shp, scenes._shape_dims(rng, shp), 1.0, pos,
                                quat=tuple(scenes.random_quaternions(rng, 1)[0])))
    sc = scenes.make(rows)
    g, o = _run_both(sc, meshes, 3, 40)
    _compare_manifolds(g, o)
    _compare_impulses(g, o)


def test_bodies_sleep_on_a_terrain():
    sc, meshes = scenes.terrain_pile(40, seed=6, amplitude=0.5, drop=(0.5, 2.0))
    sc["flags"][1:] = 0
    g, o = _run_both(sc, meshes, 4, 90)
    assert np.array_equal(g.activity(), o.activity())
    assert (~o.activity()[1:]).sum() > 10
    _compare_manifolds(g, o)
    _compare_impulses(g, o)


def test_terrain_pairs_removal_and_setters():
    from bepuphysics1_amd import BepuError
    sc, meshes = scenes.terrain_pile(30, seed=8)
    g, o = _run_both(sc, meshes, 1, 40)
    for bid in (3, 7, 12):
        g.Remove(bid)
        assert o.remove(bid) == 0
    ids = np.array([5, 9], np.uint32)
    st = o.motion_states()[ids].copy()
    st[:, 7:10] = [[2, 4, 0], [-1, 3, 1]]
    g.set_motion_states(ids, st)
    o.set_motion_states(ids, st)
    with pytest.raises(BepuError):
        g.set_motion_states(np.array([0], np.uint32), o.motion_states()[:1])
    with pytest.raises(BepuError):
        g.activate([0])
    for r in range(3):
        g.Update(30)
        o.step(30)
        assert np.array_equal(g.pairs(), o.pairs()), r
        assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states())), r
    _compare_impulses(g, o)


def test_batched_terrain_worlds_match_alone():
    from bepuphysics1_amd import Space
    worlds = [scenes.terrain_pile(int(n), seed=20 + k, organization=k % 2) for k, n in enumerate((12, 30, 7, 22))]
    sp = Space(capacity=sum(len(w[0]) for w in worlds) + 4)
    sp.register_hulls(scenes.hull_library())
    firsts = []
    for k, (sc, meshes) in enumerate(worlds):
        t = meshes[0]
        sid = sp.create_terrain_shape(t.heights, t.organization, t.transform)
        rows = sc.copy()
        rows["dims"][0, 0] = sid
        sp._mesh_ids = list(range(sid + 1))
        wid, ids = sp.add_world(rows)
        firsts.append(int(ids[0]))
    sp.Update(60)
    allst = sp.motion_states()
    for k, (sc, meshes) in enumerate(worlds):
        o = run_oracle(sc, 60, meshes=meshes)
        assert np.array_equal(f32bits(allst[firsts[k]:firsts[k] + len(sc)]), f32bits(o.motion_states())), k


@pytest.mark.parametrize("mirrored", [False, True])
def test_fast_bodies_against_a_terrain(mirrored):
    """Fast bodies of every shape fired at a bumpy terrain, discrete, passive
    and continuous: times of impact against the terrain's triangles (from
    the query box of the manifold's update) bit-identical to the oracle."""
    sc, meshes = scenes.terrain_pile(40, seed=13, amplitude=0.6, drop=(1.0, 3.0))
    if mirrored:
        t = meshes[0]
        lin = t.transform[0] * np.array([[-1], [1], [1]], np.float32)
        tr = t.transform[1] + np.array([-lin[0, 0] * (t.heights.shape[0] - 1), 0, 0], np.float32)
        meshes = [scenes.TerrainDesc(t.heights, t.organization, (lin, tr))]
    rng = np.random.default_rng(14)
    n = len(sc) - 1
    sc["state"][1:, 7:10] = np.stack([rng.uniform(-20, 20, n), rng.uniform(-240, -90, n), rng.uniform(-20, 20, n)], 1)
    sc["flags"][1:] = scenes.FLAG_ALWAYS_ACTIVE | rng.choice([0, scenes.FLAG_CCD_PASSIVE, scenes.FLAG_CCD_CONTINUOUS],
                                                              n, p=[0.2, 0.2, 0.6]).astype(np.uint32)
    g, o = _run_both(sc, meshes, 4, 5)
    st = o.motion_states()
    cont = (sc["flags"][1:] & scenes.FLAG_CCD_CONTINUOUS) != 0
    assert (st[1:][cont, 1] > -1.5).mean() > 0.9
    _compare_impulses(g, o)


def test_terrain_arguments_refused():
    from bepuphysics1_amd import BepuError, Space
    sp = Space(capacity=8)
    lin, tr = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    with pytest.raises(BepuError):
        sp.create_terrain_shape(np.zeros((1, 5), np.float32), 0, (lin, tr))
    with pytest.raises(BepuError):
        sp.create_terrain_shape(np.zeros((3, 3), np.float32), 2, (lin, tr))
    with pytest.raises(BepuError):
        sp.create_terrain_shape(np.full((3, 3), np.inf, np.float32), 0, (lin, tr))
    sid = sp.create_terrain_shape(np.zeros((3, 3), np.float32), 0, (lin, tr))
    sp._mesh_ids = [sid + 1]
    with pytest.raises(BepuError):
        sp.add_scene(scenes.make([scenes.terrain(0)]))
