"""Host-side mirror of the reference's Space / Entity / SolverSettings API for
the step path, driving the gfx950 library through its C ABI.

Names and argument meaning follow BEPUphysics v1 (paths relative to the
reference repository root):
  Space()                      BEPUphysics/Space.cs:199-233
  Space.Add(entity)            Space.cs:241-358  (ArgumentException on a double add -> ValueError)
  Space.Remove(entity)         Space.cs:364-482  (ArgumentException on a foreign remove -> ValueError)
  Space.Update()               Space.cs:527-531
  Box(position, w, h, l, mass) Entities/Prefabs/Box.cs:33-37 (mass omitted: kinematic)
  Sphere(position, r, mass)    Entities/Prefabs/Sphere.cs:47-51
  Capsule(position, length, r, mass)    Entities/Prefabs/Capsule.cs:124-128
  ConvexHull(position, points, mass)    Entities/Prefabs/ConvexHull.cs:83-87 (body at `position`,
                                        the shape recentred on its volume centroid)
  Entity.Position/Orientation/LinearVelocity/AngularVelocity/MotionState (EntityBase.cs:51-235;
                               after the body is on the device each setter writes its field there)
  Material, IsAlwaysActive
  Space.Solver.IterationLimit  Constraints/Solver.cs:15
  Space.ForceUpdater.Gravity   OtherSpaceStages/ForceUpdater.cs
  Space.TimeStepSettings.TimeStepDuration  TimeStepSettings.cs:19
Entities added before the first Update are uploaded to the device in one
batch; state lives on the device afterwards and entity properties read it back.
Settings (Solver.IterationLimit, ForceUpdater.Gravity, TimeStepDuration) can
be changed at any time; later steps use them.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import BepuError, check

SHAPE_BOX, SHAPE_SPHERE, SHAPE_CONVEX_HULL, SHAPE_CAPSULE, SHAPE_STATIC_MESH, SHAPE_TERRAIN = 0, 1, 2, 3, 4, 5
STATE_POSITION, STATE_ORIENTATION, STATE_LINEAR_VELOCITY, STATE_ANGULAR_VELOCITY = 1, 2, 4, 8
STATE_ALL = 15
_STATE_FIELD = {"position": (STATE_POSITION, slice(0, 3)), "orientation": (STATE_ORIENTATION, slice(3, 7)),
                "linear_velocity": (STATE_LINEAR_VELOCITY, slice(7, 10)),
                "angular_velocity": (STATE_ANGULAR_VELOCITY, slice(10, 13))}


def compute_hull(points):
    """ConvexHullShape.ComputeDescription on the host (bepu_compute_hull):
    returns (surface vertices [n,3] recentred on the centroid, volume
    distribution [3,3], centroid [3], volume)."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    fp = ctypes.POINTER(ctypes.c_float)
    L = _lib.load()
    n = ctypes.c_uint32()
    verts = np.zeros((len(p), 3), np.float32)
    vd = np.zeros((3, 3), np.float32)
    center = np.zeros(3, np.float32)
    vol = ctypes.c_float()
    check(L.bepu_compute_hull(p.ctypes.data_as(fp), len(p), verts.ctypes.data_as(fp), len(p), ctypes.byref(n),
                              vd.ctypes.data_as(fp), center.ctypes.data_as(fp), ctypes.byref(vol)), "compute_hull")
    return verts[:n.value].copy(), vd, center, vol.value


def compute_hull_minimum_radius(points):
    """InertiaHelper.ComputeMinimumRadius of the hull of `points` about its
    centroid (bepu_compute_hull_minimum_radius), without the margin."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    r = ctypes.c_float()
    check(_lib.load().bepu_compute_hull_minimum_radius(p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(p),
                                                       ctypes.byref(r)), "compute_hull_minimum_radius")
    return r.value


class Material:
    """BEPUphysics/Materials/Material.cs (defaults 1.0 / 0.8 / 0.0)."""

    def __init__(self, static_friction=1.0, kinetic_friction=0.8, bounciness=0.0):
        self.StaticFriction = float(static_friction)
        self.KineticFriction = float(kinetic_friction)
        self.Bounciness = float(bounciness)


class Entity:
    shape = None

    def __init__(self, position=(0, 0, 0), mass=None):
        self._space = None
        self._id = None
        self._pending = {
            "position": tuple(float(x) for x in position),
            "orientation": (0.0, 0.0, 0.0, 1.0),
            "linear_velocity": (0.0, 0.0, 0.0),
            "angular_velocity": (0.0, 0.0, 0.0),
        }
        if mass is not None:
            m = float(mass)
            if m != m or m in (float("inf"), float("-inf")):
                # EntityBase.BecomeDynamic throws InvalidOperationException (EntityBase.cs:879-880)
                raise ValueError("Cannot use a mass of %r for a dynamic entity." % mass)
        self.mass = 0.0 if mass is None else float(mass)
        self.Material = Material()
        self.IsAlwaysActive = True
        self.AllowStabilization = True

    @property
    def IsDynamic(self):
        return self.mass > 0

    @property
    def Space(self):
        return self._space

    @property
    def Id(self):
        return self._id

    def _state(self):
        if self._space is None or self._id is None or not self._space._uploaded(self._id):
            return None
        return self._space.motion_states()[self._id]

    def _get(self, key, sl):
        st = self._state()
        return tuple(float(x) for x in st[sl]) if st is not None else self._pending[key]

    def _set(self, key, value):
        value = tuple(float(x) for x in value)
        if self._space is not None and self._id is not None:
            # The entity's setter on the world (staged or on the device).
            mask, sl = _STATE_FIELD[key]
            row = np.zeros(13, np.float32)
            row[sl] = value
            self._space.set_motion_states([self._id], row[None, :], mask)
        self._pending[key] = value

    Position = property(lambda s: s._get("position", slice(0, 3)), lambda s, v: s._set("position", v))
    Orientation = property(lambda s: s._get("orientation", slice(3, 7)), lambda s, v: s._set("orientation", v))
    LinearVelocity = property(lambda s: s._get("linear_velocity", slice(7, 10)),
                              lambda s, v: s._set("linear_velocity", v))
    AngularVelocity = property(lambda s: s._get("angular_velocity", slice(10, 13)),
                               lambda s, v: s._set("angular_velocity", v))

    @property
    def MotionState(self):
        """(position, orientation xyzw, linear velocity, angular velocity)."""
        return self.Position + self.Orientation + self.LinearVelocity + self.AngularVelocity

    @MotionState.setter
    def MotionState(self, st):
        st = [float(x) for x in st]
        assert len(st) == 13
        if self._space is not None and self._id is not None:
            self._space.set_motion_states([self._id], np.asarray(st, np.float32)[None, :], STATE_ALL)
        for key, (_, sl) in _STATE_FIELD.items():
            self._pending[key] = tuple(st[sl])

    def _motion_state(self):
        ms = _lib.MotionState()
        ms.position[:] = self._pending["position"]
        ms.orientation[:] = self._pending["orientation"]
        ms.linear_velocity[:] = self._pending["linear_velocity"]
        ms.angular_velocity[:] = self._pending["angular_velocity"]
        return ms

    def _flags(self):
        f = 1 if self.IsAlwaysActive else 0
        if not self.AllowStabilization:
            f |= 2
        return f

    def _material(self):
        m = self.Material
        return _lib.Material(m.StaticFriction, m.KineticFriction, m.Bounciness)


class Box(Entity):
    """Entities/Prefabs/Box.cs: width/height/length are full extents."""
    shape = SHAPE_BOX

    def __init__(self, position, width, height, length, mass=None):
        super().__init__(position, mass)
        self.Width, self.Height, self.Length = float(width), float(height), float(length)

    def _add(self, lib, h, out):
        return lib.bepu_add_box(h, ctypes.byref(self._motion_state()), self.Width, self.Height, self.Length,
                                self.mass, ctypes.byref(self._material()), self._flags(), ctypes.byref(out))


class Sphere(Entity):
    """Entities/Prefabs/Sphere.cs."""
    shape = SHAPE_SPHERE

    def __init__(self, position, radius, mass=None):
        super().__init__(position, mass)
        self.Radius = float(radius)

    def _add(self, lib, h, out):
        return lib.bepu_add_sphere(h, ctypes.byref(self._motion_state()), self.Radius, self.mass,
                                   ctypes.byref(self._material()), self._flags(), ctypes.byref(out))


class Capsule(Entity):
    """Entities/Prefabs/Capsule.cs: a segment of `length` along local y swept by `radius`."""
    shape = SHAPE_CAPSULE

    def __init__(self, position, length, radius, mass=None):
        super().__init__(position, mass)
        self.Length, self.Radius = float(length), float(radius)

    def _add(self, lib, h, out):
        return lib.bepu_add_capsule(h, ctypes.byref(self._motion_state()), self.Length, self.Radius, self.mass,
                                    ctypes.byref(self._material()), self._flags(), ctypes.byref(out))


class ConvexHull(Entity):
    """Entities/Prefabs/ConvexHull.cs: the hull of `points`; the body sits at
    `position` and its shape is recentred on the hull's volume centroid."""
    shape = SHAPE_CONVEX_HULL

    def __init__(self, position, points, mass=None):
        super().__init__(position, mass)
        self.Points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)

    def _add(self, lib, h, out):
        fp = ctypes.POINTER(ctypes.c_float)
        sid = ctypes.c_uint32()
        st = lib.bepu_create_hull_shape(h, self.Points.ctypes.data_as(fp), len(self.Points), ctypes.byref(sid), None)
        if st != 0:
            return st
        return lib.bepu_add_convex_hull(h, ctypes.byref(self._motion_state()), sid.value, self.mass,
                                        ctypes.byref(self._material()), self._flags(), ctypes.byref(out))


class StaticMesh:
    """BroadPhaseEntries/StaticMesh.cs: a static triangle mesh (not an entity).
    `vertices` [n, 3], `indices` three per triangle, `transform` the
    AffineTransform (3x3 linear part as Matrix3x3 rows, translation) or None;
    Sidedness 0 DoubleSided (default), 1 Clockwise, 2 Counterclockwise."""

    def __init__(self, vertices, indices, transform=None):
        self.Vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.Indices = np.ascontiguousarray(indices, np.int32).reshape(-1)
        self.Transform = transform
        self.Sidedness = 0
        self.ImproveBoundaryBehavior = True
        self.Material = Material()
        self._space = None
        self._id = None

    @property
    def Id(self):
        return self._id

    def _add(self, lib, h, out):
        sid = ctypes.c_uint32()
        st = lib.bepu_create_mesh_shape(h, *_mesh_args(self.Vertices, self.Indices, self.Transform), ctypes.byref(sid))
        if st != 0:
            return st
        m = self.Material
        mat = _lib.Material(m.StaticFriction, m.KineticFriction, m.Bounciness)
        return lib.bepu_add_static_mesh(h, sid.value, int(self.Sidedness), 1 if self.ImproveBoundaryBehavior else 0,
                                        ctypes.byref(mat), ctypes.byref(out))


class Terrain:
    """BroadPhaseEntries/Terrain.cs: a static heightfield (not an entity).
    `heights` [width, length] (heights[column, row]; columns along local x,
    rows along local z), `transform` the AffineTransform (3x3 linear part as
    Matrix3x3 rows, translation); QuadTriangleOrganization 0
    BottomLeftUpperRight (default), 1 BottomRightUpperLeft.  Its sidedness
    follows the transform's winding."""

    def __init__(self, heights, transform, organization=0):
        self.Heights = np.ascontiguousarray(heights, np.float32)
        self.Transform = transform
        self.QuadTriangleOrganization = int(organization)
        self.ImproveBoundaryBehavior = True
        self.Material = Material()
        self._space = None
        self._id = None

    @property
    def Id(self):
        return self._id

    def _add(self, lib, h, out):
        sid = ctypes.c_uint32()
        st = lib.bepu_create_terrain_shape(h, *_terrain_args(self.Heights, self.QuadTriangleOrganization,
                                                             self.Transform), ctypes.byref(sid))
        if st != 0:
            return st
        m = self.Material
        mat = _lib.Material(m.StaticFriction, m.KineticFriction, m.Bounciness)
        return lib.bepu_add_terrain(h, sid.value, 1 if self.ImproveBoundaryBehavior else 0, ctypes.byref(mat),
                                    ctypes.byref(out))


def _terrain_args(heights, organization, transform):
    hh = np.ascontiguousarray(heights, np.float32)
    if hh.ndim != 2:
        raise ValueError("terrain heights must be [width, length]")
    t = np.concatenate([np.asarray(transform[0], np.float32).reshape(9),
                        np.asarray(transform[1], np.float32).reshape(3)]).astype(np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    _terrain_args.keep = (hh, t)
    return hh.ctypes.data_as(fp), hh.shape[0], hh.shape[1], int(organization), t.ctypes.data_as(fp)


def _mesh_args(vertices, indices, transform):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(indices, np.int32).reshape(-1)
    fp = ctypes.POINTER(ctypes.c_float)
    t = None
    if transform is not None:
        t = np.concatenate([np.asarray(transform[0], np.float32).reshape(9),
                            np.asarray(transform[1], np.float32).reshape(3)]).astype(np.float32)
    # the arrays must outlive the call: keep them on the function object
    _mesh_args.keep = (v, ix, t)
    return (v.ctypes.data_as(fp), len(v), ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(ix),
            t.ctypes.data_as(fp) if t is not None else None)


class _Solver:
    def __init__(self, space):
        self._space = space

    @property
    def IterationLimit(self):
        return self._space._settings.iteration_limit

    @IterationLimit.setter
    def IterationLimit(self, v):
        self._space._configure("iteration_limit", max(int(v), 0))


class _ForceUpdater:
    def __init__(self, space):
        self._space = space

    @property
    def Gravity(self):
        return tuple(self._space._settings.gravity)

    @Gravity.setter
    def Gravity(self, g):
        self._space._configure("gravity", tuple(float(x) for x in g))


class _TimeStepSettings:
    def __init__(self, space):
        self._space = space

    @property
    def TimeStepDuration(self):
        return self._space._settings.dt

    @TimeStepDuration.setter
    def TimeStepDuration(self, v):
        self._space._configure("dt", float(v))


class Space:
    """A simulation world whose step runs on the GPU."""

    def __init__(self, capacity=1024, settings=None, pair_capacity=0):
        """`capacity` bodies; `pair_capacity` broadphase pairs (0: max(1024, 4 * capacity))."""
        self._lib = _lib.load()
        self._settings = settings if settings is not None else _lib.default_settings()
        self._capacity = int(capacity)
        self._pair_capacity = int(pair_capacity)
        self._h = None
        self._entities = []
        self._mesh_ids = []
        self._n_uploaded = 0
        self._ms_cache = None
        self.Solver = _Solver(self)
        self.ForceUpdater = _ForceUpdater(self)
        self.TimeStepSettings = _TimeStepSettings(self)

    # -- lifetime ----------------------------------------------------------
    def _handle(self):
        if self._h is None:
            h = ctypes.c_void_p()
            check(self._lib.bepu_world_create_ex(ctypes.byref(self._settings), self._capacity, self._pair_capacity,
                                                 ctypes.byref(h)), "bepu_world_create")
            self._h = h
        return self._h

    def _configure(self, field, value):
        setattr(self._settings, field, value)
        if self._h is not None:
            check(self._lib.bepu_set_settings(self._h, ctypes.byref(self._settings)), "set_settings")

    def set_settings(self, settings):
        """Replaces the whole settings block; later steps use it."""
        self._settings = settings
        if self._h is not None:
            check(self._lib.bepu_set_settings(self._h, ctypes.byref(settings)), "set_settings")

    def close(self):
        if self._h is not None:
            self._lib.bepu_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- Space.Add / Update -------------------------------------------------
    def Add(self, entity):
        if entity._space is not None:
            raise ValueError("The object belongs to a Space already; cannot add it again.")
        out = ctypes.c_uint32()
        st = entity._add(self._lib, self._handle(), out)
        if st != 0:
            raise BepuError(st, "Space.Add")
        entity._space = self
        entity._id = out.value
        self._entities.append(entity)
        self._ms_cache = None
        return entity

    def Remove(self, entity):
        """Space.Remove(entity): an Entity of this space, or a body id."""
        if isinstance(entity, Entity):
            if entity._space is not self:
                raise ValueError("The object does not belong to this Space; cannot remove it.")
            bid = entity._id
        else:
            bid = int(entity)
        st = self._lib.bepu_remove(self._handle(), bid)
        if st == 3:
            raise ValueError("Body %d does not belong to this Space; cannot remove it." % bid)
        check(st, "Space.Remove")
        if isinstance(entity, Entity):
            entity._space = None
            entity._id = None
        self._ms_cache = None

    def set_motion_states(self, ids, states, fields=STATE_ALL):
        """Entity state setters for the listed bodies ([n, 13] states; `fields`
        a STATE_* mask selecting the parts written)."""
        ids = np.ascontiguousarray(ids, np.uint32)
        st = np.ascontiguousarray(states, np.float32).reshape(-1, 13)
        check(self._lib.bepu_set_motion_states(self._handle(), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                               st.ctypes.data_as(ctypes.POINTER(_lib.MotionState)), len(ids),
                                               int(fields)), "set_motion_states")
        self._ms_cache = None

    def activity(self):
        """Entity.ActivityInformation.IsActive of every body (bool array)."""
        n = ctypes.c_uint32()
        check(self._lib.bepu_get_activity(self._handle(), None, 0, ctypes.byref(n)), "get_activity")
        out = np.zeros(max(n.value, 1), np.uint8)
        check(self._lib.bepu_get_activity(self._handle(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                          n.value, ctypes.byref(n)), "get_activity")
        return out[:n.value].astype(bool)

    def add_ball_socket(self, a, b, anchor):
        """Space.Add(new BallSocketJoint(a, b, anchor)): joins bodies a and b
        (ids, or entities of this space) at the world point `anchor`; returns
        the joint id."""
        a = a.Id if isinstance(a, Entity) else int(a)
        b = b.Id if isinstance(b, Entity) else int(b)
        p = (ctypes.c_float * 3)(*[float(x) for x in anchor])
        jid = ctypes.c_uint32()
        check(self._lib.bepu_add_ball_socket(self._handle(), a, b, p, ctypes.byref(jid)), "add_ball_socket")
        return jid.value

    def set_joint_settings(self, joint, stiffness=600000.0, damping=90000.0, max_corrective=3.4028234663852886e38):
        """SpringSettings (stiffness, damping) and MaxCorrectiveVelocity of a joint."""
        check(self._lib.bepu_set_joint_settings(self._handle(), int(joint), float(stiffness), float(damping),
                                                float(max_corrective)), "set_joint_settings")

    def joint_impulses(self):
        """BallSocketJoint.TotalImpulse of every joint, [n, 3]."""
        n = ctypes.c_uint32()
        check(self._lib.bepu_get_joint_impulses(self._handle(), None, 0, ctypes.byref(n)), "get_joint_impulses")
        out = np.zeros((max(n.value, 1), 3), np.float32)
        check(self._lib.bepu_get_joint_impulses(self._handle(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                n.value, ctypes.byref(n)), "get_joint_impulses")
        return out[:n.value]

    POSITION_UPDATE_MODES = {"discrete": 0, "passive": 1, "continuous": 2}

    def set_position_update_mode(self, ids, mode):
        """Entity.PositionUpdateMode for each listed body ("discrete",
        "passive", "continuous" or 0..2); from the next step on."""
        m = self.POSITION_UPDATE_MODES.get(mode, mode)
        a = np.ascontiguousarray(ids, np.uint32)
        check(self._lib.bepu_set_position_update_mode(self._handle(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                      len(a), int(m)), "set_position_update_mode")

    def activate(self, ids):
        """Entity.ActivityInformation.Activate() for each listed body."""
        ids = np.ascontiguousarray(ids, np.uint32)
        check(self._lib.bepu_activate(self._handle(), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(ids)),
              "activate")

    def faulted_worlds(self):
        """Sub-worlds evicted by the non-finite guard (still live), sorted."""
        n = ctypes.c_uint32()
        u32p = ctypes.POINTER(ctypes.c_uint32)
        check(self._lib.bepu_get_faulted_worlds(self._handle(), None, 0, ctypes.byref(n)), "get_faulted_worlds")
        out = np.zeros(max(n.value, 1), np.uint32)
        check(self._lib.bepu_get_faulted_worlds(self._handle(), out.ctypes.data_as(u32p), n.value, ctypes.byref(n)),
              "get_faulted_worlds")
        return out[:n.value].tolist()

    def clear_error(self):
        """Acknowledges a sticky overflow / barrier error so stepping resumes."""
        check(self._lib.bepu_clear_error(self._handle()), "clear_error")

    def Update(self, steps=1):
        check(self._lib.bepu_step(self._handle(), int(steps)), "Space.Update")
        self._n_uploaded = len(self._entities)
        self._ms_cache = None

    add, update, remove = Add, Update, Remove

    @property
    def Entities(self):
        return list(self._entities)

    def _uploaded(self, idx):
        return idx is not None and idx < self._n_uploaded

    def begin_world(self):
        """Starts a new independent sub-world (the analogue of another Space):
        bodies added afterwards belong to it and never pair with other worlds."""
        out = ctypes.c_uint32()
        check(self._lib.bepu_begin_world(self._handle(), ctypes.byref(out)), "begin_world")
        return out.value

    def add_world(self, scene):
        """Adds `scene` as a new sub-world; returns (world id, body ids)."""
        wid = self.begin_world()
        return wid, self.add_scene(scene)

    def insert_world(self, scene):
        """Inserts `scene` as a new sub-world between steps (continuous
        batching); returns (world id, first body id).  The scene array's rows
        have the layout of bepu_body_desc."""
        scene = np.ascontiguousarray(scene)
        assert scene.dtype.itemsize == ctypes.sizeof(_lib.BodyDesc)
        wid, first = ctypes.c_uint32(), ctypes.c_uint32()
        ptr = scene.ctypes.data_as(ctypes.POINTER(_lib.BodyDesc))
        check(self._lib.bepu_insert_world(self._handle(), ptr, len(scene), ctypes.byref(wid), ctypes.byref(first)),
              "insert_world")
        self._ms_cache = None
        return wid.value, first.value

    def insert_worlds(self, scene_list):
        """Inserts several scenes as new sub-worlds in one call; returns a list
        of (world id, first body id) for those that fit (in order)."""
        if not scene_list:
            return []
        packed = np.ascontiguousarray(np.concatenate(scene_list))
        counts = np.array([len(x) for x in scene_list], np.uint32)
        k = len(scene_list)
        ids, firsts, done = np.zeros(k, np.uint32), np.zeros(k, np.uint32), ctypes.c_uint32()
        u32p = ctypes.POINTER(ctypes.c_uint32)
        st = self._lib.bepu_insert_worlds(self._handle(), packed.ctypes.data_as(ctypes.POINTER(_lib.BodyDesc)),
                                          counts.ctypes.data_as(u32p), k, ids.ctypes.data_as(u32p),
                                          firsts.ctypes.data_as(u32p), ctypes.byref(done))
        if st not in (0, 7):
            check(st, "insert_worlds")
        self._ms_cache = None
        return list(zip(ids[:done.value].tolist(), firsts[:done.value].tolist()))

    def remove_worlds(self, wids):
        a = np.ascontiguousarray(wids, np.uint32)
        check(self._lib.bepu_remove_worlds(self._handle(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(a)),
              "remove_worlds")

    def motion_states_ranges(self, firsts, counts):
        """Motion states of several body ranges in one readback: a list of (count_i, 13) arrays."""
        f = np.ascontiguousarray(firsts, np.uint32)
        c = np.ascontiguousarray(counts, np.uint32)
        total = int(c.sum())
        out = (_lib.MotionState * max(total, 1))()
        u32p = ctypes.POINTER(ctypes.c_uint32)
        check(self._lib.bepu_get_motion_states_ranges(self._handle(), f.ctypes.data_as(u32p), c.ctypes.data_as(u32p),
                                                      len(f), out), "motion_states_ranges")
        flat = np.frombuffer(out, dtype=np.float32, count=13 * total).reshape(total, 13).copy()
        return np.split(flat, np.cumsum(c)[:-1].astype(np.int64)) if len(c) else []

    def request_motion_states_ranges(self, firsts, counts):
        """Queues a readback of several body ranges behind the queued steps;
        returns a ticket for collect_motion_states (which waits for it)."""
        f = np.ascontiguousarray(firsts, np.uint32)
        c = np.ascontiguousarray(counts, np.uint32)
        t = ctypes.c_uint32()
        u32p = ctypes.POINTER(ctypes.c_uint32)
        check(self._lib.bepu_request_motion_states_ranges(self._handle(), f.ctypes.data_as(u32p),
                                                          c.ctypes.data_as(u32p), len(f), ctypes.byref(t)),
              "request_motion_states_ranges")
        return t.value, c

    def collect_motion_states(self, request):
        """The states of a requested readback: a list of (count_i, 13) arrays."""
        ticket, c = request
        total = int(c.sum())
        out = (_lib.MotionState * max(total, 1))()
        check(self._lib.bepu_collect_motion_states(self._handle(), ticket, out), "collect_motion_states")
        flat = np.frombuffer(out, dtype=np.float32, count=13 * total).reshape(total, 13).copy()
        return np.split(flat, np.cumsum(c)[:-1].astype(np.int64)) if len(c) else []

    def remove_world(self, wid):
        check(self._lib.bepu_remove_world(self._handle(), wid), "remove_world")

    def synchronize(self):
        check(self._lib.bepu_synchronize(self._handle()), "synchronize")

    def motion_states_range(self, first, count):
        """(count, 13) motion states of bodies [first, first + count)."""
        out = (_lib.MotionState * max(count, 1))()
        check(self._lib.bepu_get_motion_states_range(self._handle(), first, count, out), "motion_states_range")
        return np.frombuffer(out, dtype=np.float32, count=13 * count).reshape(count, 13).copy()

    def world_range(self, wid):
        """(first body id, body count) of a sub-world."""
        first, count = ctypes.c_uint32(), ctypes.c_uint32()
        check(self._lib.bepu_get_world_range(self._handle(), wid, ctypes.byref(first), ctypes.byref(count)),
              "world_range")
        return first.value, count.value

    def create_hull_shape(self, points):
        """Registers a hull shape (bepu_create_hull_shape); returns (shape id, centroid)."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        fp = ctypes.POINTER(ctypes.c_float)
        sid = ctypes.c_uint32()
        center = np.zeros(3, np.float32)
        check(self._lib.bepu_create_hull_shape(self._handle(), p.ctypes.data_as(fp), len(p), ctypes.byref(sid),
                                               center.ctypes.data_as(fp)), "create_hull_shape")
        return sid.value, center

    def create_mesh_shape(self, vertices, indices, transform=None):
        """Registers a static mesh shape (bepu_create_mesh_shape); returns its id."""
        sid = ctypes.c_uint32()
        check(self._lib.bepu_create_mesh_shape(self._handle(), *_mesh_args(vertices, indices, transform),
                                               ctypes.byref(sid)), "create_mesh_shape")
        return sid.value

    def create_terrain_shape(self, heights, organization, transform):
        """Registers a terrain shape with its transform (bepu_create_terrain_shape); returns its id."""
        sid = ctypes.c_uint32()
        check(self._lib.bepu_create_terrain_shape(self._handle(), *_terrain_args(heights, organization, transform),
                                                  ctypes.byref(sid)), "create_terrain_shape")
        return sid.value

    def save_snapshot(self):
        """The whole world's state (bepu_save_snapshot) as a uint8 array."""
        n = ctypes.c_uint64()
        check(self._lib.bepu_snapshot_size(self._handle(), ctypes.byref(n)), "snapshot_size")
        buf = np.empty(n.value, np.uint8)
        check(self._lib.bepu_save_snapshot(self._handle(), buf.ctypes.data, n.value, ctypes.byref(n)),
              "save_snapshot")
        return buf

    def restore_snapshot(self, snapshot):
        """Loads a snapshot of this world (same structure) back (bepu_restore_snapshot)."""
        buf = np.ascontiguousarray(snapshot, np.uint8)
        check(self._lib.bepu_restore_snapshot(self._handle(), buf.ctypes.data, buf.size), "restore_snapshot")
        self._ms_cache = None

    def register_meshes(self, meshes):
        """Registers mesh and terrain shapes in order ((vertices, indices,
        transform) or a scenes.TerrainDesc each); scene rows of static meshes
        and terrains name them by index in dims[0]."""
        from .scenes import TerrainDesc
        self._mesh_ids = [self.create_terrain_shape(m.heights, m.organization, m.transform)
                          if isinstance(m, TerrainDesc) else self.create_mesh_shape(*m) for m in meshes]
        return self._mesh_ids

    def register_hulls(self, hull_points):
        """Registers hull shapes in order; scene rows of hull bodies name them
        by index in dims[0], so register a scene's hull list before its bodies."""
        return [self.create_hull_shape(p)[0] for p in hull_points]

    def add_scene(self, scene):
        """Adds every row of a bepuphysics1_amd.scenes array; returns the ids.
        Hull rows name a registered hull shape in dims[0]."""
        lib, h = self._lib, self._handle()
        ids = np.empty(len(scene), np.uint32)
        ms = _lib.MotionState()
        mat = _lib.Material()
        out = ctypes.c_uint32()
        for i in range(len(scene)):
            r = scene[i]
            s = r["state"]
            ms.position[:] = s[0:3].tolist()
            ms.orientation[:] = s[3:7].tolist()
            ms.linear_velocity[:] = s[7:10].tolist()
            ms.angular_velocity[:] = s[10:13].tolist()
            mat.static_friction, mat.kinetic_friction, mat.bounciness = [float(x) for x in r["material"]]
            d = r["dims"]
            if int(r["shape"]) == SHAPE_BOX:
                st = lib.bepu_add_box(h, ctypes.byref(ms), float(d[0]), float(d[1]), float(d[2]), float(r["mass"]),
                                      ctypes.byref(mat), int(r["flags"]), ctypes.byref(out))
            elif int(r["shape"]) == SHAPE_SPHERE:
                st = lib.bepu_add_sphere(h, ctypes.byref(ms), float(d[0]), float(r["mass"]), ctypes.byref(mat),
                                         int(r["flags"]), ctypes.byref(out))
            elif int(r["shape"]) == SHAPE_CAPSULE:
                st = lib.bepu_add_capsule(h, ctypes.byref(ms), float(d[0]), float(d[1]), float(r["mass"]),
                                          ctypes.byref(mat), int(r["flags"]), ctypes.byref(out))
            elif int(r["shape"]) == SHAPE_CONVEX_HULL:
                st = lib.bepu_add_convex_hull(h, ctypes.byref(ms), int(d[0]), float(r["mass"]), ctypes.byref(mat),
                                              int(r["flags"]), ctypes.byref(out))
            elif int(r["shape"]) == SHAPE_STATIC_MESH:
                st = lib.bepu_add_static_mesh(h, self._mesh_ids[int(d[0])], int(d[1]), 1 if d[2] else 0,
                                              ctypes.byref(mat), ctypes.byref(out))
            elif int(r["shape"]) == SHAPE_TERRAIN:
                st = lib.bepu_add_terrain(h, self._mesh_ids[int(d[0])], 1 if d[1] else 0, ctypes.byref(mat),
                                          ctypes.byref(out))
            else:
                raise NotImplementedError("shape %d" % int(r["shape"]))
            if st != 0:
                raise BepuError(st, "Space.Add")
            ids[i] = out.value
            self._entities.append(None)
        self._ms_cache = None
        return ids

    # -- readback ---------------------------------------------------------
    def _count(self, fn):
        n = ctypes.c_uint32()
        fn(self._handle(), None, 0, ctypes.byref(n))
        return n.value

    def motion_states(self):
        """MotionState of every body as an [N, 13] float32 array (position, orientation xyzw, v, w)."""
        if self._ms_cache is None:
            n = self._count(self._lib.bepu_get_motion_states)
            buf = (_lib.MotionState * max(n, 1))()
            check(self._lib.bepu_get_motion_states(self._handle(), buf, n, ctypes.byref(ctypes.c_uint32())),
                  "get_motion_states")
            self._ms_cache = np.frombuffer(buf, np.float32).reshape(-1, 13)[:n].copy()
        return self._ms_cache

    def pairs(self):
        """Sorted pair keys of the last step's broadphase overlaps."""
        n = self._count(self._lib.bepu_get_pairs)
        out = np.zeros(max(n, 1), np.uint64)
        check(self._lib.bepu_get_pairs(self._handle(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n,
                                       ctypes.byref(ctypes.c_uint32())), "get_pairs")
        return out[:n]

    def manifolds(self):
        n = self._count(self._lib.bepu_get_manifolds)
        buf = (_lib.Manifold * max(n, 1))()
        check(self._lib.bepu_get_manifolds(self._handle(), buf, n, ctypes.byref(ctypes.c_uint32())),
              "get_manifolds")
        return list(buf)[:n]

    def impulses(self):
        n = self._count(self._lib.bepu_get_impulses)
        buf = (_lib.Impulses * max(n, 1))()
        check(self._lib.bepu_get_impulses(self._handle(), buf, n, ctypes.byref(ctypes.c_uint32())),
              "get_impulses")
        return list(buf)[:n]

    def stats(self):
        s = _lib.StepStats()
        check(self._lib.bepu_get_stats(self._handle(), ctypes.byref(s)), "get_stats")
        return s

    BROADPHASES = {"grid": 0, "tree": 1}

    def set_broadphase(self, kind):
        """Space.BroadPhase setter (Space.cs:110-134): "grid" (default) or "tree"."""
        check(self._lib.bepu_set_broadphase(self._handle(), self.BROADPHASES[kind]), "set_broadphase")

    TUNING_FIELDS = ("narrowphase_block", "convex_lds", "convex_side_stream", "small_island_max", "step_graph")

    def tuning(self):
        """The launch configuration (bepu_tuning) as a dict."""
        t = _lib.Tuning()
        check(self._lib.bepu_get_tuning(self._handle(), ctypes.byref(t)), "get_tuning")
        return {f: int(getattr(t, f)) for f in self.TUNING_FIELDS}

    def set_tuning(self, **values):
        """Sets some fields of the launch configuration; results do not depend on it."""
        t = _lib.Tuning()
        check(self._lib.bepu_get_tuning(self._handle(), ctypes.byref(t)), "get_tuning")
        for f, v in values.items():
            if f not in self.TUNING_FIELDS:
                raise KeyError(f)
            setattr(t, f, int(v))
        check(self._lib.bepu_set_tuning(self._handle(), ctypes.byref(t)), "set_tuning")

    def autotune(self, trial_steps=2):
        """Searches the launch configuration on this world's own steps (they
        advance the simulation) and keeps the fastest; returns it with the
        chosen configuration's time per step and the steps run."""
        t = _lib.Tuning()
        n = ctypes.c_uint32()
        check(self._lib.bepu_autotune(self._handle(), int(trial_steps), ctypes.byref(t), ctypes.byref(n)), "autotune")
        out = {f: int(getattr(t, f)) for f in self.TUNING_FIELDS}
        out["ms_per_step"] = float(t.ms_per_step)
        out["steps"] = int(n.value)
        return out

    def set_timing(self, enable=True):
        check(self._lib.bepu_set_timing(self._handle(), 1 if enable else 0), "set_timing")

    def stage_times(self):
        ms = (ctypes.c_float * 16)()
        n = ctypes.c_uint32()
        names = ctypes.c_char_p()
        check(self._lib.bepu_get_stage_times(self._handle(), ms, 16, ctypes.byref(n), ctypes.byref(names)),
              "get_stage_times")
        return dict(zip(names.value.decode().split(","), list(ms)[:n.value]))
