"""ctypes binding of the oracle library (liboracle_refcpu.so).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import it; the
product package (bepuphysics1_amd) never does.

The oracle restates BEPUphysics v1's per-step path on the CPU, single
threaded (BEPUphysics/Space.cs:496-521 and the files cited in the C++
sources).  Its API mirrors include/bepu_gpu.h with a ``refcpu_`` prefix.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_refcpu.so")


# The structs are include/bepu_gpu.h's: one Python definition of that layout
# serves the product binding and the oracle (refcpu_world.cpp fills the same
# header structs), and tests/test_abi.py checks it against both compiled
# libraries' sizeof.
from bepuphysics1_amd._lib import (Contact, Impulses, Manifold, Material,  # noqa: E402,F401
                                   MotionState, Settings, StepStats)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.refcpu_create.restype = vp
        L.refcpu_create.argtypes = [ctypes.POINTER(Settings)]
        L.refcpu_destroy.argtypes = [vp]
        L.refcpu_set_broadphase.argtypes = [vp, ctypes.c_int]
        L.refcpu_set_threads.argtypes = [vp, ctypes.c_int]
        L.refcpu_settings_default.argtypes = [ctypes.POINTER(Settings)]
        L.refcpu_add_box.argtypes = [vp, ctypes.POINTER(MotionState), ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.POINTER(Material),
                                     ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_add_sphere.argtypes = [vp, ctypes.POINTER(MotionState), ctypes.c_float, ctypes.c_float,
                                        ctypes.POINTER(Material), ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_add_capsule.argtypes = [vp, ctypes.POINTER(MotionState), ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.POINTER(Material), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_uint32)]
        fpp = ctypes.POINTER(ctypes.c_float)
        L.refcpu_add_convex_hull.argtypes = [vp, ctypes.POINTER(MotionState), fpp, ctypes.c_uint32, fpp,
                                             ctypes.c_float, ctypes.c_float, ctypes.POINTER(Material),
                                             ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_set_position_update_mode.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                                                      ctypes.c_int]
        L.refcpu_add_ball_socket.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, fpp,
                                             ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_set_joint_settings.argtypes = [vp, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.refcpu_get_joint_impulses.argtypes = [vp, fpp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_cast.argtypes = [ctypes.c_int, fpp, fpp, ctypes.c_int, fpp, fpp, fpp, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, fpp]
        shape_args = [fpp, fpp, ctypes.c_int, fpp, fpp, fpp, ctypes.c_int, fpp]
        L.refcpu_gjk_intersecting.argtypes = shape_args + [fpp]
        L.refcpu_gjk_closest.argtypes = shape_args + [fpp]
        L.refcpu_mpr_contact.argtypes = shape_args + [fpp, fpp]
        L.refcpu_convex_candidate.argtypes = shape_args + [fpp, fpp, ctypes.POINTER(ctypes.c_int), fpp, fpp, fpp]
        L.refcpu_step.argtypes = [vp, ctypes.c_int]
        L.refcpu_broadphase_only.argtypes = [vp]
        for name, typ in (("refcpu_get_motion_states", MotionState), ("refcpu_get_manifolds", Manifold),
                          ("refcpu_get_impulses", Impulses)):
            getattr(L, name).argtypes = [vp, ctypes.POINTER(typ), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_get_pairs.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_get_aabbs.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_get_stats.argtypes = [vp, ctypes.POINTER(StepStats)]
        fp = ctypes.POINTER(ctypes.c_float)
        L.refcpu_box_box.argtypes = [fp, fp, fp, fp, fp, fp, fp, fp, fp, ctypes.POINTER(ctypes.c_int)]
        L.refcpu_sphere_sphere.argtypes = [ctypes.c_float, ctypes.c_float, fp, fp, fp]
        L.refcpu_box_sphere.argtypes = [fp, fp, fp, ctypes.c_float, fp, fp]
        L.refcpu_mix64.restype = ctypes.c_uint64
        L.refcpu_mix64.argtypes = [ctypes.c_uint64]
        L.refcpu_remove.argtypes = [vp, ctypes.c_uint32]
        L.refcpu_activate.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]
        L.refcpu_set_solve_order.argtypes = [vp, ctypes.c_int]
        L.refcpu_permuted_order.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64)]
        L.refcpu_get_activity.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_uint32)]
        L.refcpu_set_motion_states.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(MotionState),
                                               ctypes.c_uint32, ctypes.c_uint32]
        L.refcpu_set_settings.argtypes = [vp, ctypes.POINTER(Settings)]
        L.refcpu_reduce_contacts.argtypes = [fp, fp, fp, fp, ctypes.c_float, ctypes.POINTER(ctypes.c_int)]
        ip = ctypes.POINTER(ctypes.c_int32)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.refcpu_create_mesh_shape.argtypes = [vp, fpp, ctypes.c_uint32, ip, ctypes.c_uint32, fpp, u32p]
        L.refcpu_add_static_mesh.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(Material), u32p]
        L.refcpu_get_mesh_tree.argtypes = [vp, ctypes.c_uint32, fpp, ip, ctypes.c_uint32, u32p]
        L.refcpu_mesh_overlaps.argtypes = [vp, ctypes.c_uint32, fpp, fpp, ip, ctypes.c_uint32, u32p]
        L.refcpu_create_terrain_shape.argtypes = [vp, fpp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, u32p]
        L.refcpu_add_terrain.argtypes = [vp, ctypes.c_uint32, fpp, ctypes.c_int, ctypes.POINTER(Material), u32p]
        L.refcpu_terrain_overlaps.argtypes = [vp, ctypes.c_uint32, fpp, fpp, ip, ctypes.c_uint32, u32p]
        L.refcpu_terrain_triangle.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, fpp, ip]
        L.refcpu_terrain_query_box.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, fpp]
        L.refcpu_mesh_peaks.argtypes = [vp, ip]
        L.refcpu_triangle_convex.argtypes = [fpp, fpp, ctypes.c_int, ctypes.c_float, ctypes.c_float, fpp,
                                             ctypes.c_int, ctypes.c_float, ip, fpp, fpp]
        L.refcpu_triangle_region.argtypes = [fpp, fpp]
        L.refcpu_triangle_sphere.argtypes = [ctypes.c_float, fpp, ctypes.c_int, ctypes.c_float, fpp,
                                             ctypes.POINTER(ctypes.c_int)]
        L.refcpu_closest_point_triangle.argtypes = [fpp, fpp, fpp]
        L.refcpu_sizeof.restype = ctypes.c_uint32
        L.refcpu_sizeof.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def permuted_order(n, k, count):
    """The permuted visit order of step k over n constraints (first `count` visits)."""
    out = np.zeros(max(count, 1), np.int64)
    lib().refcpu_permuted_order(int(n), int(k), int(count), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out[:count]


def default_settings():
    s = Settings()
    lib().refcpu_settings_default(ctypes.byref(s))
    return s


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class World:
    """One oracle world; mirrors bepu_world of include/bepu_gpu.h."""

    def __init__(self, settings=None):
        self._s = settings if settings is not None else default_settings()
        self.h = lib().refcpu_create(ctypes.byref(self._s))

    def close(self):
        if self.h:
            lib().refcpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n):
        """Threads for step(); the results are identical for any count."""
        lib().refcpu_set_threads(self.h, int(n))

    def set_broadphase_mode(self, mode):
        """0: uniform grid (default), 1: brute force, 2: sort and sweep -- one pair set."""
        lib().refcpu_set_broadphase(self.h, int(mode))

    def set_bruteforce_broadphase(self, on=True):
        lib().refcpu_set_broadphase(self.h, 1 if on else 0)

    def add_body(self, shape, dims, mass, state, material, flags=0, hull=None):
        """`hull` = (surface vertices [n,3], volume distribution [3,3][, minimum
        radius without the margin]) for a BEPU_SHAPE_CONVEX_HULL body (as
        bepu_compute_hull and bepu_compute_hull_minimum_radius return them;
        only continuous collision detection reads the radius, 0 if absent)."""
        ms = MotionState()
        ms.position[:] = [float(x) for x in state[0:3]]
        ms.orientation[:] = [float(x) for x in state[3:7]]
        ms.linear_velocity[:] = [float(x) for x in state[7:10]]
        ms.angular_velocity[:] = [float(x) for x in state[10:13]]
        mat = Material(*[float(x) for x in material])
        bid = ctypes.c_uint32()
        if shape == 0:
            rc = lib().refcpu_add_box(self.h, ctypes.byref(ms), float(dims[0]), float(dims[1]), float(dims[2]),
                                      float(mass), ctypes.byref(mat), int(flags), ctypes.byref(bid))
        elif shape == 1:
            rc = lib().refcpu_add_sphere(self.h, ctypes.byref(ms), float(dims[0]), float(mass),
                                         ctypes.byref(mat), int(flags), ctypes.byref(bid))
        elif shape == 3:
            rc = lib().refcpu_add_capsule(self.h, ctypes.byref(ms), float(dims[0]), float(dims[1]), float(mass),
                                          ctypes.byref(mat), int(flags), ctypes.byref(bid))
        elif shape == 2:
            verts = np.ascontiguousarray(hull[0], np.float32)
            vd = np.ascontiguousarray(hull[1], np.float32)
            rc = lib().refcpu_add_convex_hull(self.h, ctypes.byref(ms), _fptr(verts), len(verts), _fptr(vd),
                                              float(hull[2] if len(hull) > 2 else 0.0), float(mass),
                                              ctypes.byref(mat), int(flags),
                                              ctypes.byref(bid))
        else:
            raise NotImplementedError("shape %d" % shape)
        if rc != 0:
            raise ValueError("refcpu add failed with status %d" % rc)
        return bid.value

    def add_scene(self, scene, hulls=None, meshes=None):
        """`hulls[k]` = (vertices, volume distribution) of hull shape id k, which
        a hull body names in dims[0]; `meshes[k]` = (vertices, indices,
        transform) of the mesh a static mesh row names in dims[0]."""
        mesh_ids = {}
        for i in range(len(scene)):
            r = scene[i]
            shape = int(r["shape"])
            if shape == 4:
                k = int(r["dims"][0])
                if k not in mesh_ids:
                    mesh_ids[k] = self.create_mesh_shape(*meshes[k])
                self.add_static_mesh(mesh_ids[k], int(r["dims"][1]), bool(r["dims"][2]), r["material"])
                continue
            if shape == 5:
                t = meshes[int(r["dims"][0])]
                sid = self.create_terrain_shape(t.heights, t.organization)
                self.add_terrain(sid, t.transform, bool(r["dims"][1]), r["material"])
                continue
            hull = hulls[int(r["dims"][0])] if shape == 2 else None
            self.add_body(shape, r["dims"], r["mass"], r["state"], r["material"], int(r["flags"]), hull=hull)

    def create_mesh_shape(self, vertices, indices, transform=None):
        """new StaticMeshShape(vertices, indices[, worldTransform]); `transform`
        is (3x3 linear part as Matrix3x3 rows, translation) or None.  Returns
        the mesh shape id."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        ix = np.ascontiguousarray(indices, np.int32).reshape(-1)
        t = None
        if transform is not None:
            t = np.concatenate([np.asarray(transform[0], np.float32).reshape(9),
                                np.asarray(transform[1], np.float32).reshape(3)]).astype(np.float32)
        sid = ctypes.c_uint32()
        rc = lib().refcpu_create_mesh_shape(self.h, _fptr(v), len(v), ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            len(ix), _fptr(t) if t is not None else None, ctypes.byref(sid))
        if rc != 0:
            raise ValueError("refcpu_create_mesh_shape failed with status %d" % rc)
        return sid.value

    def add_static_mesh(self, shape_id, sidedness=0, improve_boundary=True, material=(1.0, 0.8, 0.0)):
        """A StaticMesh of a mesh shape added to the space; returns its body id."""
        mat = Material(*[float(x) for x in material])
        bid = ctypes.c_uint32()
        rc = lib().refcpu_add_static_mesh(self.h, int(shape_id), int(sidedness), 1 if improve_boundary else 0,
                                          ctypes.byref(mat), ctypes.byref(bid))
        if rc != 0:
            raise ValueError("refcpu_add_static_mesh failed with status %d" % rc)
        return bid.value

    def create_terrain_shape(self, heights, organization=0):
        """new TerrainShape(heights[width, length], organization); returns its id."""
        h = np.ascontiguousarray(heights, np.float32)
        sid = ctypes.c_uint32()
        rc = lib().refcpu_create_terrain_shape(self.h, _fptr(h), h.shape[0], h.shape[1], int(organization),
                                               ctypes.byref(sid))
        if rc != 0:
            raise ValueError("refcpu_create_terrain_shape failed with status %d" % rc)
        return sid.value

    def add_terrain(self, shape_id, transform, improve_boundary=True, material=(1.0, 0.8, 0.0)):
        """A Terrain of a terrain shape under the affine `transform` (3x3 linear
        part as Matrix3x3 rows, translation); returns its body id."""
        t = np.concatenate([np.asarray(transform[0], np.float32).reshape(9),
                            np.asarray(transform[1], np.float32).reshape(3)]).astype(np.float32)
        mat = Material(*[float(x) for x in material])
        bid = ctypes.c_uint32()
        rc = lib().refcpu_add_terrain(self.h, int(shape_id), _fptr(t), 1 if improve_boundary else 0,
                                      ctypes.byref(mat), ctypes.byref(bid))
        if rc != 0:
            raise ValueError("refcpu_add_terrain failed with status %d" % rc)
        return bid.value

    def terrain_overlaps(self, body, lmin, lmax):
        """TerrainShape.GetOverlaps over a local box: triangle ids in scan order."""
        lo = np.ascontiguousarray(lmin, np.float32)
        hi = np.ascontiguousarray(lmax, np.float32)
        n = ctypes.c_uint32()
        lib().refcpu_terrain_overlaps(self.h, int(body), _fptr(lo), _fptr(hi), None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), np.int32)
        lib().refcpu_terrain_overlaps(self.h, int(body), _fptr(lo), _fptr(hi),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n.value, ctypes.byref(n))
        return out[:n.value]

    def terrain_triangle(self, body, tri, world):
        """(vertices [3, 3], sequential indices or None) of a terrain triangle."""
        v = np.zeros(9, np.float32)
        ix = np.zeros(3, np.int32)
        rc = lib().refcpu_terrain_triangle(self.h, int(body), int(tri), 1 if world else 0, _fptr(v),
                                           ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        if rc != 0:
            raise ValueError("refcpu_terrain_triangle failed with status %d" % rc)
        return v.reshape(3, 3), (None if world else ix)

    def mesh_peaks(self):
        """Largest per-pair sizes met by the mesh manifold updates so far:
        candidates, edge contacts, vertex contacts, blocked vertices, blocked
        edges, pair testers (before stale ones left)."""
        out = np.zeros(6, np.int32)
        lib().refcpu_mesh_peaks(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def terrain_query_box(self, convex, terrain):
        out = np.zeros(6, np.float32)
        rc = lib().refcpu_terrain_query_box(self.h, int(convex), int(terrain), _fptr(out))
        if rc != 0:
            raise ValueError("refcpu_terrain_query_box failed with status %d" % rc)
        return out[:3], out[3:]

    def mesh_tree(self, shape_id):
        """The flattened MeshBoundingBoxTree: (boxes [n, 6], links [n, 2])."""
        n = ctypes.c_uint32()
        lib().refcpu_get_mesh_tree(self.h, int(shape_id), None, None, 0, ctypes.byref(n))
        boxes = np.zeros((max(n.value, 1), 6), np.float32)
        links = np.zeros((max(n.value, 1), 2), np.int32)
        lib().refcpu_get_mesh_tree(self.h, int(shape_id), _fptr(boxes),
                                   links.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n.value, ctypes.byref(n))
        return boxes[:n.value], links[:n.value]

    def mesh_overlaps(self, shape_id, bmin, bmax):
        """MeshBoundingBoxTree.GetOverlaps(box): triangles (first-index positions) in traversal order."""
        lo = np.ascontiguousarray(bmin, np.float32)
        hi = np.ascontiguousarray(bmax, np.float32)
        n = ctypes.c_uint32()
        lib().refcpu_mesh_overlaps(self.h, int(shape_id), _fptr(lo), _fptr(hi), None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), np.int32)
        lib().refcpu_mesh_overlaps(self.h, int(shape_id), _fptr(lo), _fptr(hi),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n.value, ctypes.byref(n))
        return out[:n.value]

    def step(self, n=1):
        lib().refcpu_step(self.h, int(n))

    def remove(self, body_id):
        """Space.Remove(entity); returns the status code."""
        return lib().refcpu_remove(self.h, int(body_id))

    def set_motion_states(self, ids, states, fields=15):
        """Entity state setters for each listed body; `states` is [n, 13];
        `fields` is a BEPU_STATE_* mask (15: the MotionState setter)."""
        ids = np.ascontiguousarray(ids, np.uint32)
        st = np.ascontiguousarray(states, np.float32).reshape(-1, 13)
        return lib().refcpu_set_motion_states(self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              st.ctypes.data_as(ctypes.POINTER(MotionState)), len(ids), int(fields))

    def set_solve_order(self, mode):
        """0: colour by colour (the GPU's order); 1: the reference's permuted
        single-threaded order (Solver.cs:237-263, PermutationMapper.cs)."""
        return lib().refcpu_set_solve_order(self.h, int(mode))

    def add_ball_socket(self, a, b, anchor):
        """new BallSocketJoint(a, b, anchor) added to the space; returns the
        joint id (raises on a bad argument)."""
        p = np.ascontiguousarray(anchor, np.float32)
        jid = ctypes.c_uint32()
        rc = lib().refcpu_add_ball_socket(self.h, int(a), int(b), _fptr(p), ctypes.byref(jid))
        if rc != 0:
            raise ValueError("refcpu_add_ball_socket failed with status %d" % rc)
        return jid.value

    def set_joint_settings(self, joint, stiffness=600000.0, damping=90000.0, max_corrective=3.4028234663852886e38):
        """SpringSettings and MaxCorrectiveVelocity of a joint; returns the status code."""
        return lib().refcpu_set_joint_settings(self.h, int(joint), float(stiffness), float(damping),
                                               float(max_corrective))

    def joint_impulses(self):
        """BallSocketJoint.TotalImpulse of every joint, [n, 3]."""
        n = ctypes.c_uint32()
        lib().refcpu_get_joint_impulses(self.h, None, 0, ctypes.byref(n))
        out = np.zeros((max(n.value, 1), 3), np.float32)
        lib().refcpu_get_joint_impulses(self.h, _fptr(out), n.value, ctypes.byref(n))
        return out[:n.value]

    def set_position_update_mode(self, ids, mode):
        """Entity.PositionUpdateMode for each listed body: 0 Discrete, 1
        Passive, 2 Continuous.  Returns the status code."""
        ids = np.ascontiguousarray(ids, np.uint32)
        return lib().refcpu_set_position_update_mode(self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                     len(ids), int(mode))

    def activate(self, ids):
        """Entity.ActivityInformation.Activate() for each listed body."""
        ids = np.ascontiguousarray(ids, np.uint32)
        return lib().refcpu_activate(self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(ids))

    def activity(self):
        """Entity.ActivityInformation.IsActive of every body (bool array)."""
        n = ctypes.c_uint32()
        lib().refcpu_get_activity(self.h, None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), np.uint8)
        lib().refcpu_get_activity(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n.value,
                                  ctypes.byref(n))
        return out[:n.value].astype(bool)

    def set_settings(self, settings):
        """Settings changed between steps (iteration limit, gravity, dt, ...)."""
        self._s = settings
        return lib().refcpu_set_settings(self.h, ctypes.byref(settings))

    def broadphase_only(self):
        lib().refcpu_broadphase_only(self.h)

    def _count(self, fn):
        n = ctypes.c_uint32()
        fn(self.h, None, 0, ctypes.byref(n))
        return n.value

    def motion_states(self):
        n = self._count(lib().refcpu_get_motion_states)
        buf = (MotionState * max(n, 1))()
        lib().refcpu_get_motion_states(self.h, buf, n, ctypes.byref(ctypes.c_uint32()))
        arr = np.frombuffer(buf, dtype=np.float32).reshape(-1, 13)[:n]
        return arr.copy()

    def aabbs(self):
        n = self._count(lib().refcpu_get_aabbs)
        out = np.zeros((max(n, 1), 6), np.float32)
        lib().refcpu_get_aabbs(self.h, _fptr(out), n, ctypes.byref(ctypes.c_uint32()))
        return out[:n]

    def pairs(self):
        n = self._count(lib().refcpu_get_pairs)
        out = np.zeros(max(n, 1), np.uint64)
        lib().refcpu_get_pairs(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n,
                               ctypes.byref(ctypes.c_uint32()))
        return out[:n]

    def manifolds(self):
        n = self._count(lib().refcpu_get_manifolds)
        buf = (Manifold * max(n, 1))()
        lib().refcpu_get_manifolds(self.h, buf, n, ctypes.byref(ctypes.c_uint32()))
        return list(buf)[:n]

    def impulses(self):
        n = self._count(lib().refcpu_get_impulses)
        buf = (Impulses * max(n, 1))()
        lib().refcpu_get_impulses(self.h, buf, n, ctypes.byref(ctypes.c_uint32()))
        return list(buf)[:n]

    def stats(self):
        s = StepStats()
        lib().refcpu_get_stats(self.h, ctypes.byref(s))
        return s


def box_box(half_a, pos_a, q_a, half_b, pos_b, q_b):
    """BoxBoxCollider.AreBoxesColliding on one pair: returns (hit, axis, distance, contacts[n,4], ids[n])."""
    f = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    ha, pa, qa, hb, pb, qb = map(f, (half_a, pos_a, q_a, half_b, pos_b, q_b))
    axis = np.zeros(3, np.float32)
    dist = np.zeros(1, np.float32)
    cont = np.zeros((8, 4), np.float32)
    ids = np.zeros(8, np.int32)
    n = lib().refcpu_box_box(_fptr(ha), _fptr(pa), _fptr(qa), _fptr(hb), _fptr(pb), _fptr(qb), _fptr(axis),
                             _fptr(dist), _fptr(cont), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if n < 0:
        return False, axis, float(dist[0]), cont[:0], ids[:0]
    return True, axis, float(dist[0]), cont[:n].copy(), ids[:n].copy()


def sphere_sphere(ra, rb, pa, pb):
    out = np.zeros(7, np.float32)
    f = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    pa, pb = f(pa), f(pb)
    hit = lib().refcpu_sphere_sphere(float(ra), float(rb), _fptr(pa), _fptr(pb), _fptr(out))
    return bool(hit), out


def box_sphere(half, bp, bq, r, sp):
    out = np.zeros(7, np.float32)
    f = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    half, bp, bq, sp = map(f, (half, bp, bq, sp))
    hit = lib().refcpu_box_sphere(_fptr(half), _fptr(bp), _fptr(bq), float(r), _fptr(sp), _fptr(out))
    return bool(hit), out


def mix64(k):
    return int(lib().refcpu_mix64(int(k)))


def reduce_contacts(positions, depths, normal, cand_pos, cand_depth):
    """ContactReducer.ReduceContacts on four contacts (list order) and a
    candidate: returns (list index removed or -1, candidate added)."""
    f = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    p, d, n, c = f(positions), f(depths), f(normal), f(cand_pos)
    add = ctypes.c_int()
    r = lib().refcpu_reduce_contacts(_fptr(p), _fptr(d), _fptr(n), _fptr(c), float(cand_depth), ctypes.byref(add))
    return int(r), bool(add.value)


# ---------------------------------------------------------------------------
# General convex queries on two posed shapes.  A shape is a dict with
# "kind" (0 box, 1 sphere, 2 hull, 3 capsule), "half" (box half extents,
# margin included), "half_length" (capsule), "verts" (hull), "margin".

def _shape_arrays(sh):
    kind = int(sh["kind"])
    d = np.zeros(5, np.float32)
    d[0] = kind
    if kind == 0:
        d[1:4] = sh["half"]
    elif kind == 3:
        d[1] = sh["half_length"]
    d[4] = sh["margin"]
    v = np.ascontiguousarray(sh.get("verts", np.zeros((1, 3))), np.float32)
    return d, v, (len(v) if kind == 2 else 0)


def _pose(p):
    return np.ascontiguousarray(p, np.float32)


def gjk_intersecting(sa, pa, sb, pb, axis=(0, 0, 0)):
    da, va, na = _shape_arrays(sa)
    db, vb, nb = _shape_arrays(sb)
    pa, pb = _pose(pa), _pose(pb)
    ax = np.array(axis, np.float32)
    hit = lib().refcpu_gjk_intersecting(_fptr(da), _fptr(va), na, _fptr(pa), _fptr(db), _fptr(vb), nb, _fptr(pb),
                                        _fptr(ax))
    return bool(hit), ax


def gjk_closest(sa, pa, sb, pb):
    da, va, na = _shape_arrays(sa)
    db, vb, nb = _shape_arrays(sb)
    pa, pb = _pose(pa), _pose(pb)
    out = np.zeros(6, np.float32)
    hit = lib().refcpu_gjk_closest(_fptr(da), _fptr(va), na, _fptr(pa), _fptr(db), _fptr(vb), nb, _fptr(pb),
                                   _fptr(out))
    return bool(hit), out[:3].copy(), out[3:].copy()


def mpr_contact(sa, pa, sb, pb, axis):
    da, va, na = _shape_arrays(sa)
    db, vb, nb = _shape_arrays(sb)
    pa, pb = _pose(pa), _pose(pb)
    ax = np.array(axis, np.float32)
    out = np.zeros(8, np.float32)
    hit = lib().refcpu_mpr_contact(_fptr(da), _fptr(va), na, _fptr(pa), _fptr(db), _fptr(vb), nb, _fptr(pb),
                                   _fptr(ax), _fptr(out))
    return bool(hit), out


class ConvexTester:
    """GeneralConvexPairTester state carried across calls."""

    def __init__(self):
        self.state = ctypes.c_int(0)
        self.sep_axis = np.zeros(3, np.float32)
        self.local_dir = np.zeros(3, np.float32)

    def candidate(self, sa, pa, sb, pb, va=(0, 0, 0), vb=(0, 0, 0)):
        da, xa, na = _shape_arrays(sa)
        db, xb, nb = _shape_arrays(sb)
        pa, pb = _pose(pa), _pose(pb)
        va, vb = np.array(va, np.float32), np.array(vb, np.float32)
        out = np.zeros(8, np.float32)
        hit = lib().refcpu_convex_candidate(_fptr(da), _fptr(xa), na, _fptr(pa), _fptr(db), _fptr(xb), nb, _fptr(pb),
                                            _fptr(va), _fptr(vb), ctypes.byref(self.state), _fptr(self.sep_axis),
                                            _fptr(self.local_dir), _fptr(out))
        return bool(hit), out


def cast(kind, shape, pose, origin, direction, radius=0.0, max_length=1.0, scaling=0.8):
    """GJKToolbox casts on one posed shape: kind "sphere" (SphereCast), "ray"
    (RayCast) or "ccd" (CCDSphereCast); returns (hit, t)."""
    k = {"sphere": 0, "ray": 1, "ccd": 2}[kind]
    d, v, n = _shape_arrays(shape)
    p = _pose(pose)
    o = np.ascontiguousarray(origin, np.float32)
    dr = np.ascontiguousarray(direction, np.float32)
    t = np.zeros(1, np.float32)
    hit = lib().refcpu_cast(k, _fptr(d), _fptr(v), n, _fptr(p), _fptr(o), _fptr(dr), float(radius),
                            float(max_length), float(scaling), _fptr(t))
    return bool(hit), float(t[0])


# ---------------------------------------------------------------------------
# Triangle pair testers on one triangle in the convex's local space.

class TriangleTester:
    """TriangleConvexPairTester state (state, escape attempts, separating axis)
    carried across calls."""

    def __init__(self):
        self.st = np.zeros(2, np.int32)
        self.axis = np.zeros(3, np.float32)

    @property
    def state(self):
        return int(self.st[0])

    def candidates(self, shape, tri, sidedness=0, min_radius=0.0, core_scaling=0.8, tri_margin=0.0):
        d, v, n = _shape_arrays(shape)
        t = np.ascontiguousarray(tri, np.float32).reshape(9)
        out = np.zeros((2, 8), np.float32)
        k = lib().refcpu_triangle_convex(_fptr(d), _fptr(v), n, float(min_radius), float(core_scaling), _fptr(t),
                                         int(sidedness), float(tri_margin),
                                         self.st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _fptr(self.axis),
                                         _fptr(out))
        return out[:k].copy()


def triangle_region(tri, normal):
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    nn = np.ascontiguousarray(normal, np.float32)
    return int(lib().refcpu_triangle_region(_fptr(t), _fptr(nn)))


def triangle_sphere(radius, tri, sidedness=0, tri_margin=0.0):
    """(candidates [k, 8], Voronoi region)."""
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    out = np.zeros(8, np.float32)
    reg = ctypes.c_int(-1)
    k = lib().refcpu_triangle_sphere(float(radius), _fptr(t), int(sidedness), float(tri_margin), _fptr(out),
                                     ctypes.byref(reg))
    return out[None, :].copy()[:k], reg.value


def closest_point_triangle(tri, p):
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    pp = np.ascontiguousarray(p, np.float32)
    out = np.zeros(3, np.float32)
    r = lib().refcpu_closest_point_triangle(_fptr(t), _fptr(pp), _fptr(out))
    return r, out
