"""The benchmark workloads at their own sizes and parameters, on the GPU
against the oracle (run on an MI355X):

  C4  4,096 worlds, N ~ logU[8, 256] of boxes, spheres and 16-point hulls,
      8 iterations, 300 steps: every world's motion states bit for bit and its
      impulse table exactly, against that world run alone on the oracle;
  C2  the 32,768-body stress mix with hulls, one step: pair set, manifolds,
      impulses and motion states exact;
  C3  64 x 64 towers of 20 boxes (81,920 bodies), 10 iterations: the first
      steps exact at full size, then all 600 steps: the deepest contact
      bounded and the towers standing counted; and under the reference's
      defaults (sleeping) all 600 steps exact, standing and sleeping reported.
"""
import concurrent.futures
import ctypes
import os

import numpy as np
import pytest

from bepuphysics1_amd import Space, _lib, scenes
from helpers import f32bits, impulse_table, manifold_table, oracle_hulls, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

MIX = (scenes.SHAPE_BOX, scenes.SHAPE_SPHERE, scenes.SHAPE_CONVEX_HULL)
IMP_DTYPE = np.dtype([("key", np.uint64), ("color", np.int32), ("count", np.int32), ("normal", np.float32, 4),
                      ("friction", np.float32, 4), ("sliding", np.float32, 2), ("twist", np.float32),
                      ("pad", np.float32)])
assert IMP_DTYPE.itemsize == ctypes.sizeof(_lib.Impulses)


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _gpu_impulses(sp):
    L, h = sp._lib, sp._handle()
    n = ctypes.c_uint32()
    L.bepu_get_impulses(h, None, 0, ctypes.byref(n))
    out = np.zeros(max(n.value, 1), IMP_DTYPE)
    assert L.bepu_get_impulses(h, out.ctypes.data_as(ctypes.POINTER(_lib.Impulses)), n.value,
                               ctypes.byref(ctypes.c_uint32())) == 0
    return out[:n.value]


def _oracle_impulses(w):
    from oracle import refcpu
    L = refcpu.lib()
    n = ctypes.c_uint32()
    L.refcpu_get_impulses(w.h, None, 0, ctypes.byref(n))
    out = np.zeros(max(n.value, 1), IMP_DTYPE)
    L.refcpu_get_impulses(w.h, out.ctypes.data_as(ctypes.POINTER(refcpu.Impulses)), n.value,
                          ctypes.byref(ctypes.c_uint32()))
    return out[:n.value]


def _by_key(a):
    return a[np.argsort(a["key"], kind="stable")]


def test_batched_c4_every_world_matches_the_oracle():
    """C4 at its own parameters: all 4,096 worlds batched in one handle for
    300 steps; each world alone on the oracle (one task per world)."""
    from oracle import refcpu
    steps = 300
    worlds = scenes.batched_worlds(4096, seed=0, shapes=MIX)
    s = _lib.default_settings()
    s.iteration_limit = 8
    sp = Space(capacity=sum(len(w) for w in worlds), settings=s)
    sp.register_hulls(scenes.hull_library())
    firsts = []
    for w in worlds:
        _, ids = sp.add_world(w)
        firsts.append(int(ids[0]))
    sp.Update(steps)
    gstates = sp.motion_states()
    gimp = _gpu_impulses(sp)
    sp.close()
    # the world of every constraint, from its lower body id
    lo = (gimp["key"] >> np.uint64(32)).astype(np.int64)
    wk = np.searchsorted(np.asarray(firsts), lo, side="right") - 1
    hulls = oracle_hulls()
    so = refcpu.default_settings()
    so.iteration_limit = 8

    def run(k):
        w = refcpu.World(so)
        w.add_scene(worlds[k], hulls)
        w.step(steps)
        out = w.motion_states(), _oracle_impulses(w)
        w.close()
        return out

    with concurrent.futures.ThreadPoolExecutor(_threads()) as ex:
        results = list(ex.map(run, range(len(worlds))))
    bodies = 0
    for k, (ostates, oimp) in enumerate(results):
        f, n = firsts[k], len(worlds[k])
        assert np.array_equal(f32bits(gstates[f:f + n]), f32bits(ostates)), "world %d motion states" % k
        mine = gimp[wk == k].copy()
        base = np.uint64(f)
        mine["key"] -= (base << np.uint64(32)) | base
        assert np.array_equal(_by_key(mine).view(np.uint8), _by_key(oimp).view(np.uint8)), "world %d impulses" % k
        bodies += n
    assert bodies == len(gstates)


def test_stress_c2_full_size_exact():
    """C2 at 32,768 bodies (half boxes, a quarter spheres, a quarter hulls),
    one step from rest: pair set, manifolds, impulses and motion states."""
    sc = scenes.stress(32768, shapes=scenes.STRESS_MIX)
    g = run_gpu(sc, 1)
    o = run_oracle(sc, 0)
    o.set_threads(_threads())
    o.step(1)
    assert np.array_equal(g.pairs(), o.pairs())
    assert manifold_table(g.manifolds()) == manifold_table(o.manifolds())
    assert impulse_table(g.impulses()) == impulse_table(o.impulses())
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))
    assert g.stats().pairs > 60000


def test_towers_c3_full_size_600_steps():
    """C3: the first 3 steps of all 81,920 boxes exact against the oracle,
    then 600 steps with every body always active (no sleeping): no contact
    goes deeper than 0.1, and the towers standing (top box within 0.5
    horizontally and 1.0 vertically of where it started: twenty soft contacts
    compress the column by about 0.3) are counted.  With
    every body kept awake the towers lean: four corner contacts solved in
    sequence carry unequal impulses, the 0.01 allowed penetration leaves a
    dead band without restoring bias, and a 20-box column amplifies the tilt
    (the oracle shows the same lean on a single jitter-free tower), so
    standing is reported here, not required."""
    sc = scenes.towers()
    g = run_gpu(sc, 3)
    o = run_oracle(sc, 0)
    o.set_threads(_threads())
    o.step(3)
    assert np.array_equal(g.pairs(), o.pairs())
    assert impulse_table(g.impulses()) == impulse_table(o.impulses())
    assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states()))
    o.close()
    deepest = 0.0
    for _ in range(597 // 199):
        g.Update(199)
        deepest = max(deepest, g.stats().max_penetration)
    assert g.stats().steps == 600
    ms = g.motion_states()
    start = sc["state"]
    dyn = sc["mass"] > 0
    top = np.zeros(len(sc), bool)
    ys = start[:, 1]
    top[dyn] = ys[dyn] == ys[dyn].max()
    assert top.sum() == 64 * 64
    drift = np.hypot(ms[top, 0] - start[top, 0], ms[top, 2] - start[top, 2])
    standing = int(((drift < 0.5) & (np.abs(ms[top, 1] - start[top, 1]) < 1.0)).sum())
    print("C3 always active, 600 steps: %d of 4096 towers standing, top drift max %.3f median %.3f, deepest %.4f"
          % (standing, drift.max(), np.median(drift), deepest))
    assert deepest < 0.1
    assert np.isfinite(ms).all()


@pytest.mark.timeout(600)
def test_towers_c3_full_size_with_sleeping_exact():
    """C3 under the reference's defaults (no body always active, so towers
    go to sleep once settled): all 81,920 boxes for 600 steps, motion states
    and activity bit for bit against the oracle at steps 100 and 600; the
    deepest contact, the towers standing and the bodies asleep reported.
    The deactivation budget looks at 100 island members per step, 5 towers,
    so a full turn over the 4,096 towers takes 820 steps: a tower not yet put
    to sleep leans as an always-active one does (and one that fell sleeps
    once its boxes rest), so standing is reported, not required."""
    sc = scenes.towers()
    sc["flags"] = 0
    g = run_gpu(sc, 0)
    o = run_oracle(sc, 0)
    o.set_threads(_threads())
    deepest = 0.0
    done = 0
    for c in (100, 300, 600):
        while done < c:
            k = min(100, c - done)
            g.Update(k)
            deepest = max(deepest, g.stats().max_penetration)
            done += k
        o.step(c - o.stats().steps)
        if c in (100, 600):
            assert np.array_equal(f32bits(g.motion_states()), f32bits(o.motion_states())), "step %d" % c
            assert np.array_equal(g.activity(), o.activity()), "step %d activity" % c
    o.close()
    ms = g.motion_states()
    act = g.activity()
    dyn = sc["mass"] > 0
    start = sc["state"]
    ys = start[:, 1]
    top = np.zeros(len(sc), bool)
    top[dyn] = ys[dyn] == ys[dyn].max()
    drift = np.hypot(ms[top, 0] - start[top, 0], ms[top, 2] - start[top, 2])
    standing = int(((drift < 0.5) & (np.abs(ms[top, 1] - start[top, 1]) < 1.0)).sum())
    asleep = int((~act[dyn]).sum())
    up = (drift < 0.5) & (np.abs(ms[top, 1] - start[top, 1]) < 1.0)
    top_asleep = ~act[top]
    print("C3 with sleeping, 600 steps: %d of 4096 towers standing, %d of them asleep, %d of %d boxes asleep, "
          "deepest %.4f" % (standing, int(top_asleep.sum()), asleep, int(dyn.sum()), deepest))
    assert deepest < 0.1
    assert top_asleep.sum() > 1000
