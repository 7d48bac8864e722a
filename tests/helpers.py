"""Shared helpers for the parity tests: run one scene on the GPU library and
on the oracle, and compare their outputs."""
import numpy as np


def f32bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def oracle_hulls():
    """The scene hull library as the oracle takes it: (surface vertices, volume
    distribution) per shape, computed by the library's host hull builder."""
    global _ORACLE_HULLS
    if _ORACLE_HULLS is None:
        from bepuphysics1_amd import compute_hull, compute_hull_minimum_radius, scenes
        _ORACLE_HULLS = [compute_hull(p)[:2] + (compute_hull_minimum_radius(p),) for p in scenes.hull_library()]
    return _ORACLE_HULLS


_ORACLE_HULLS = None


def run_gpu(scene, steps, settings=None, capacity=None, broadphase=None, meshes=None):
    from bepuphysics1_amd import Space, scenes
    sp = Space(capacity=capacity or max(len(scene), 2), settings=settings)
    if broadphase:
        sp.set_broadphase(broadphase)
    if scenes.has_hulls(scene):
        sp.register_hulls(scenes.hull_library())
    if meshes:
        sp.register_meshes(meshes)
    sp.add_scene(scene)
    if steps:
        sp.Update(steps)
    return sp


def run_oracle(scene, steps, settings=None, meshes=None):
    from oracle import refcpu
    from bepuphysics1_amd import scenes
    w = refcpu.World(settings)
    w.add_scene(scene, oracle_hulls() if scenes.has_hulls(scene) else None, meshes)
    if steps:
        w.step(steps)
    return w


def manifold_table(manifolds):
    """{key: (count, [(id, pos3, normal3, depth) sorted by id])}"""
    out = {}
    for m in manifolds:
        cs = []
        for i in range(m.count):
            c = m.contacts[i]
            cs.append((int(c.id), tuple(c.position), tuple(c.normal), float(c.penetration_depth)))
        out[int(m.key)] = (m.count, int(m.body_a), int(m.body_b), cs)
    return out


def impulse_table(imps):
    out = {}
    for r in imps:
        out[int(r.key)] = (int(r.color), int(r.count), list(r.normal)[:r.count], list(r.friction)[:r.count],
                           list(r.sliding), float(r.twist))
    return out


def rel_err(a, b, floor=1e-3):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
