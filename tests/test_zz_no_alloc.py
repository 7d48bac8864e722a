"""bepu_get_motion_states_ranges_noalloc against bepu_get_motion_states_ranges
(run on an MI355X): the same bits, on two calls in a row (the second one
smaller, over the contents the first one left in the shared device array),
and its limits reported as errors."""
import ctypes

import numpy as np
import pytest

from bepuphysics1_amd import Space, _lib, scenes
from helpers import f32bits

pytestmark = pytest.mark.gpu

E_INVALID_ARG, E_CAPACITY = 1, 7


def _ranges(sp, firsts, counts, fn):
    f = np.ascontiguousarray(firsts, np.uint32)
    c = np.ascontiguousarray(counts, np.uint32)
    total = int(c.sum())
    out = np.zeros((max(total, 1), 13), np.float32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    st = getattr(sp._lib, fn)(sp._handle(), f.ctypes.data_as(u32p), c.ctypes.data_as(u32p), len(f),
                              out.ctypes.data_as(ctypes.POINTER(_lib.MotionState)))
    return st, out[:total]


def test_noalloc_readback_matches_the_allocating_one():
    mix = (scenes.SHAPE_BOX, scenes.SHAPE_SPHERE, scenes.SHAPE_CONVEX_HULL)
    worlds = scenes.batched_worlds(300, seed=5, shapes=mix)
    sp = Space(capacity=sum(len(w) for w in worlds) + 1)
    sp.register_hulls(scenes.hull_library())
    spans = []
    for w in worlds:
        _, ids = sp.add_world(w)
        spans.append((int(ids[0]), len(w)))
    sp.Update(40)
    spans = np.array(spans, np.int64)
    # first call: every world
    st, a = _ranges(sp, spans[:, 0], spans[:, 1], "bepu_get_motion_states_ranges_noalloc")
    assert st == 0
    st, b = _ranges(sp, spans[:, 0], spans[:, 1], "bepu_get_motion_states_ranges")
    assert st == 0
    assert np.array_equal(f32bits(a), f32bits(b))
    # second, smaller call after more steps: some worlds, partial ranges, in another order
    sp.Update(7)
    pick = np.array([17, 3, 250, 120, 4, 4])
    firsts = spans[pick, 0] + np.array([0, 1, 2, 0, 0, 3])
    counts = np.minimum(spans[pick, 1] - (firsts - spans[pick, 0]), np.array([5, 9, 1, 30, 2, 4]))
    st, a = _ranges(sp, firsts, counts, "bepu_get_motion_states_ranges_noalloc")
    assert st == 0
    st, b = _ranges(sp, firsts, counts, "bepu_get_motion_states_ranges")
    assert st == 0
    assert len(a) == counts.sum() < 100
    assert np.array_equal(f32bits(a), f32bits(b))
    # no ranges: nothing to do
    st, _ = _ranges(sp, [], [], "bepu_get_motion_states_ranges_noalloc")
    assert st == 0
    # limits: too many ranges, too many bodies, a range past the end, a pending upload
    st, _ = _ranges(sp, np.zeros(16385, np.uint32), np.ones(16385, np.uint32),
                    "bepu_get_motion_states_ranges_noalloc")
    assert st == E_CAPACITY
    st, _ = _ranges(sp, [0] * 2, [len(sp.motion_states()), 262144], "bepu_get_motion_states_ranges_noalloc")
    assert st in (E_CAPACITY, E_INVALID_ARG)
    st, _ = _ranges(sp, [int(spans[-1, 0])], [int(spans[-1, 1]) + 1], "bepu_get_motion_states_ranges_noalloc")
    assert st == E_INVALID_ARG
    sp.add_scene(scenes.make([scenes._row(scenes.SHAPE_SPHERE, (0.5, 0, 0), 1.0, (0.0, 50.0, 0.0))]))
    st, _ = _ranges(sp, [0], [1], "bepu_get_motion_states_ranges_noalloc")
    assert st == E_INVALID_ARG
    sp.close()
