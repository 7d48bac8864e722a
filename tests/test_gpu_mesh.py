"""Static meshes on the GPU vs the oracle (run on an MI355X).

The device walks each mesh's bounding box tree, runs the triangle pair
testers and the mesh manifold's boundary handling and reduction in the
oracle's operation order, so pair sets, motion states, manifolds and
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


def test_mesh_pile_exact():
    """Every shape dropped on a bumpy heightfield mesh: pair sets and states
    bit-identical every 30 steps, manifolds and impulses at the end."""
    sc, meshes = scenes.mesh_pile(48, seed=3)
    g, o = _run_both(sc, meshes, 4, 30)
    _compare_manifolds(g, o)
    _compare_impulses(g, o)
    kinds = {m.kind for m in o.manifolds() if m.count}
    assert {4, 5} <= kinds  # mesh-convex and mesh-sphere manifolds took part


@pytest.mark.parametrize("sidedness,improve", [(1, False), (0, True), (2, True)])
def test_flat_mesh_sidedness_and_boundary(sidedness, improve):
    """A flat tessellated mesh, one-sided or double-sided, with and without
    the improved boundary handling: bodies sliding across the triangle edges."""
    sc, meshes = scenes.mesh_pile(32, seed=11, amplitude=0.0, cells=24, sidedness=sidedness,
                                  improve_boundary=improve)
    rng = np.random.default_rng(4)
    sc["state"][1:, 7] = rng.uniform(-3, 3, len(sc) - 1)  # sideways velocities
    g, o = _run_both(sc, meshes, 3, 40)
    _compare_manifolds(g, o)
    _compare_impulses(g, o)


def test_transformed_meshes_and_two_grounds():
    """Two mesh shapes, one turned, lifted and skewed by its affine transform,
    the other shared by two static meshes; bodies fall across all three."""
    V0, I0, _ = scenes.heightfield_mesh(cells=10, size=16.0, amplitude=0.8, seed=2, center=(-6, 0))
    c, s = np.cos(0.35), np.sin(0.35)
    lin = np.array([[c, s, 0], [-s, c, 0], [0.1, 0, 1.0]], np.float32)
    V1, I1, _ = scenes.heightfield_mesh(cells=6, size=12.0, amplitude=0.0, seed=3, center=(7, 0))
    meshes = [(V0, I0, None), (V1, I1, (lin, np.array([0, 1.5, 0], np.float32)))]
    rows = [scenes.static_mesh(0), scenes.static_mesh(1, sidedness=1), scenes.static_mesh(1, improve_boundary=False)]
    rng = np.random.default_rng(9)
    for k in range(40):
        shp = int(rng.choice(ALL_SHAPES))
        pos = (rng.uniform(-12, 12), rng.uniform(3, 8), rng.uniform(-5, 5))
        rows.append(scenes._row(shp, scenes._shape_dims(rng, shp), 1.0, pos,
                                quat=tuple(scenes.random_quaternions(rng, 1)[0])))
    sc = scenes.make(rows)
    g, o = _run_both(sc, meshes, 3, 40)
    _compare_manifolds(g, o)
    _compare_impulses(g, o)


def test_bodies_sleep_on_a_mesh():
    """Bodies that may sleep settle on a mesh and go to sleep; activity,
    states and pairs stay bit-identical through 360 steps."""
    sc, meshes = scenes.mesh_pile(40, seed=6, amplitude=0.5, drop=(0.5, 2.0))
    sc["flags"][1:] = 0  # not always active: deactivation on
    g, o = _run_both(sc, meshes, 4, 90)
    assert np.array_equal(g.activity(), o.activity())
    assert (~o.activity()[1:]).sum() > 10  # most came to rest
    _compare_manifolds(g, o)
    _compare_impulses(g, o)


def test_mesh_pairs_removal_and_setters():
    """Bodies removed and set moving over a mesh mid-run; the mesh itself
    refuses state setters and activation."""
    from bepuphysics1_amd import BepuError
    sc, meshes = scenes.mesh_pile(30, seed=8)
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


def test_batched_mesh_worlds_match_alone():
    """Sub-worlds, each with its own mesh ground, stepped in one batch: every
    world bit-identical to the oracle stepping it alone."""
    from bepuphysics1_amd import Space
    from helpers import oracle_hulls
    worlds = [scenes.mesh_pile(int(n), seed=20 + k) for k, n in enumerate((12, 30, 7, 22))]
    sp = Space(capacity=sum(len(w[0]) for w in worlds) + 4)
    sp.register_hulls(scenes.hull_library())
    firsts = []
    for k, (sc, meshes) in enumerate(worlds):
        sid = sp.create_mesh_shape(*meshes[0])
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


def test_fast_bodies_against_a_mesh():
    """Fast bodies of every shape fired at a bumpy mesh, discrete, passive and
    continuous: times of impact against the triangles and the clamped motion
    bit-identical to the oracle."""
    sc, meshes = scenes.mesh_pile(40, seed=13, amplitude=0.6, drop=(1.0, 3.0))
    rng = np.random.default_rng(14)
    n = len(sc) - 1
    sc["state"][1:, 7:10] = np.stack([rng.uniform(-20, 20, n), rng.uniform(-240, -90, n), rng.uniform(-20, 20, n)], 1)
    sc["flags"][1:] = scenes.FLAG_ALWAYS_ACTIVE | rng.choice([0, scenes.FLAG_CCD_PASSIVE, scenes.FLAG_CCD_CONTINUOUS],
                                                              n, p=[0.2, 0.2, 0.6]).astype(np.uint32)
    g, o = _run_both(sc, meshes, 4, 5)
    st = o.motion_states()
    cont = (sc["flags"][1:] & scenes.FLAG_CCD_CONTINUOUS) != 0
    assert (st[1:][cont, 1] > -1.5).mean() > 0.9  # the continuous ones stayed above the mesh
    _compare_impulses(g, o)


def test_repeated_triangles_and_large_spheres():
    """A mesh listing every triangle twice (the second copy meets the tester
    the first updated, in the same step) and large spheres over a fine mesh
    and a fine terrain (more overlapped triangles than one pass tests side by
    side): bit-identical to the oracle."""
    V, I, _ = scenes.heightfield_mesh(cells=12, size=12.0, amplitude=0.4, seed=5, center=(-8, 0))
    I2 = np.concatenate([I, I])
    Vf, If, _ = scenes.heightfield_mesh(cells=40, size=14.0, amplitude=0.3, seed=6, center=(8, 0))
    t = scenes.heightfield_terrain(29, 29, 0.5, amplitude=0.3, seed=7, center=(0, 14))
    meshes = [(V, I2, None), (Vf, If, None), t]
    rows = [scenes.static_mesh(0), scenes.static_mesh(1), scenes.terrain(2)]
    rng = np.random.default_rng(21)
    for k in range(30):
        shp = int(rng.choice(ALL_SHAPES))
        rows.append(scenes._row(shp, scenes._shape_dims(rng, shp), 1.0,
                                (rng.uniform(-12, -4), rng.uniform(1.5, 4), rng.uniform(-4, 4)),
                                quat=tuple(scenes.random_quaternions(rng, 1)[0])))
    for cx, cz in ((8, 0), (0, 14)):
        for k in range(6):
            rows.append(scenes._row(scenes.SHAPE_SPHERE, (rng.uniform(1.0, 1.5), 0, 0), 1.0,
                                    (cx + rng.uniform(-4, 4), rng.uniform(3, 6), cz + rng.uniform(-4, 4))))
    sc = scenes.make(rows)
    g, o = _run_both(sc, meshes, 3, 30)
    _compare_manifolds(g, o)
    _compare_impulses(g, o)
    # within the device's per-pair lists (32 candidates, 32 edge and 32
    # vertex contacts, 64 blocked vertices and edges, 32 testers), with more
    # than one pass of triangles for the spheres
    peaks = o.mesh_peaks()
    assert np.all(peaks <= [32, 32, 32, 64, 64, 32]) and peaks[1] > 11, peaks
