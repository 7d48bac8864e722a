"""Whole-world snapshots on the GPU (run on an MI355X): stepping on from a
restored snapshot gives the same bits as stepping on from the moment it was
taken, in the same world and in a second world built the same way; a
snapshot of another structure is refused.
"""
import numpy as np
import pytest

from bepuphysics1_amd import scenes
from helpers import f32bits, impulse_table, manifold_table, run_gpu

pytestmark = pytest.mark.gpu


def _world():
    """A pile of every shape on a terrain and a static mesh, some bodies free
    to sleep, a continuous one, and a ball socket joint."""
    sc, meshes = scenes.terrain_pile(40, seed=31, amplitude=0.8)
    V, I, _ = scenes.heightfield_mesh(cells=8, size=10.0, amplitude=0.3, seed=32, center=(30, 0))
    rows = [scenes.static_mesh(1)]
    rng = np.random.default_rng(33)
    for k in range(12):
        shp = int(rng.choice([0, 1, 2, 3]))
        rows.append(scenes._row(shp, scenes._shape_dims(rng, shp), 1.0,
                                (30 + rng.uniform(-3, 3), rng.uniform(1.5, 4), rng.uniform(-3, 3))))
    scene = np.concatenate([sc, scenes.make(rows)])
    scene["flags"][1:21] = 0  # may sleep
    scene["flags"][21] = scenes.FLAG_ALWAYS_ACTIVE | scenes.FLAG_CCD_CONTINUOUS
    g = run_gpu(scene, 0, meshes=[meshes[0], (V, I, None)])
    g.add_ball_socket(22, 23, g.motion_states()[22][:3])
    return g


def _state(g):
    return (f32bits(g.motion_states()), g.pairs(), manifold_table(g.manifolds()), impulse_table(g.impulses()),
            g.activity())


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]
    assert a[3] == b[3]
    assert np.array_equal(a[4], b[4])


def test_restore_continues_bit_identically():
    g = _world()
    g.Update(40)
    snap = g.save_snapshot()
    steps = g.stats().steps
    g.Update(45)  # an odd count: the double-buffered arrays change roles
    first = _state(g)
    g.restore_snapshot(snap)
    assert g.stats().steps == steps
    g.Update(45)
    _same(first, _state(g))
    # a second world built the same way takes it after another number of steps
    h = _world()
    h.Update(7)
    h.restore_snapshot(snap)
    h.Update(45)
    _same(first, _state(h))


def test_restore_refuses_another_structure():
    from bepuphysics1_amd import BepuError
    g = _world()
    g.Update(10)
    snap = g.save_snapshot()
    g.Remove(5)
    with pytest.raises(BepuError):
        g.restore_snapshot(snap)
    bad = snap.copy()
    bad[0] ^= 0xFF
    with pytest.raises(BepuError):
        g.restore_snapshot(bad)
    with pytest.raises(BepuError):
        g.restore_snapshot(snap[:-8])
    g.Update(10)  # the world goes on
    assert np.isfinite(g.motion_states()).all()
