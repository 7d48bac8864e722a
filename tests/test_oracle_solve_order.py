"""The reference's solve order as an oracle mode, and how far the colour order
(the GPU's, and the oracle's default) moves results from it.  CPU only.

Solver.UpdateSingleThreaded (Solver.cs:237-263) presteps and warm-starts the
solver list in order, then visits iterationLimit x n entries in the order
(i * prime + offset) mod n of PermutationMapper (PermutationMapper.cs:27-127),
advancing the permutation every step.  The oracle restates that scheme
(`set_solve_order(1)`); the list order (the order constraints were added,
which follows the reference's broad phase) is taken as pair-key order, and
mode 2 takes the reverse, another order of addition.

Pinned here: the visit order is one permutation of the list repeated per
iteration and changes from step to step; constraints that share no dynamic
body give bit-identical results in every order.  Reported: the impulse and
position drift between the colour order and the permuted order beside the
drift of the permuted order under a reversed list, on the pyramid and on
8 x 8 towers of 20 boxes.
"""
import numpy as np

from bepuphysics1_amd import scenes


def _run(oracle, sc, steps, mode):
    w = oracle.World(oracle.default_settings())
    w.add_scene(sc, None)
    assert w.set_solve_order(mode) == 0
    w.step(steps)
    return w


def _impulses(w):
    return {i.key: np.array(list(i.normal)[:i.count] + [i.sliding[0], i.sliding[1], i.twist]) for i in w.impulses()}


def _drift(a, b):
    ia, ib = _impulses(a), _impulses(b)
    keys = sorted(k for k in set(ia) & set(ib) if len(ia[k]) == len(ib[k]))
    A = np.concatenate([ia[k] for k in keys])
    B = np.concatenate([ib[k] for k in keys])
    return float(np.linalg.norm(A - B) / np.linalg.norm(A))


def test_permuted_order_is_one_permutation_repeated_and_moves_per_step(oracle):
    for n in (1, 2, 3, 7, 100, 1000):
        for k in (1, 2, 57):
            order = oracle.permuted_order(n, k, 3 * n)
            first = order[:n]
            assert sorted(first.tolist()) == list(range(n))
            assert np.array_equal(order[n:2 * n], first) and np.array_equal(order[2 * n:], first)
    a, b = oracle.permuted_order(1000, 5, 1000), oracle.permuted_order(1000, 6, 1000)
    assert not np.array_equal(a, b)


def test_independent_constraints_give_the_same_bits_in_every_order(oracle):
    """Spheres resting apart on a kinematic ground: no two constraints share a
    dynamic body, so the order of visits cannot matter."""
    rows = [scenes.ground()] + [scenes._row(scenes.SHAPE_SPHERE, (0.5, 0, 0), 1.0, (3.0 * i, 0.49, 0.0), flags=0)
                                for i in range(12)]
    sc = scenes.make(rows)
    ref = _run(oracle, sc, 120, 0).motion_states()
    for mode in (1, 2):
        assert np.array_equal(_run(oracle, sc, 120, mode).motion_states().view(np.uint32), ref.view(np.uint32))


def test_colour_order_drift_is_that_of_the_reference_order_itself(oracle):
    report = []
    for name, sc in (("pyramid", scenes.pyramid()), ("towers 8x8x20", scenes.towers(nx=8, nz=8))):
        c10, p10, r10 = (_run(oracle, sc, 10, m) for m in (0, 1, 2))
        colour, own = _drift(p10, c10), _drift(p10, r10)
        c, p, r = (_run(oracle, sc, 300, m) for m in (0, 1, 2))
        dyn = sc["mass"] > 0
        pos = lambda x, y: float(np.abs(x.motion_states()[dyn, :3] - y.motion_states()[dyn, :3]).max())
        pens = [w.stats().max_penetration for w in (c, p, r)]
        report.append("%s: step-10 impulse L2 drift colour %.3f vs reversed list %.3f; 300-step max position "
                      "drift colour %.4f vs reversed list %.4f; deepest contact %.4f / %.4f / %.4f"
                      % (name, colour, own, pos(c, p), pos(r, p), *pens))
        # the colour order is no further from the reference's order than a
        # different order of addition takes the reference from itself (2x)
        assert colour < 2.0 * own
        assert pos(c, p) < 2.0 * pos(r, p) + 0.01
        assert max(pens) < 1.05 * min(pens)
    print("\n".join(report))
