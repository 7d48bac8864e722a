"""ctypes binding of libbepu_gpu.so (the C ABI declared in include/bepu_gpu.h).

The library is built in-tree by ``python -m bepuphysics1_amd.build``.  There
is no fallback: if the library is missing or no HIP device is present, calls
fail loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libbepu_gpu.so")

BEPU_OK = 0
STATUS_NAMES = {
    0: "ok", 1: "invalid argument", 2: "already added", 3: "not found", 4: "out of device memory",
    5: "device error", 6: "non-finite state", 7: "capacity exceeded", 8: "unsupported",
}


class BepuError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__("%s failed: %s (%d)" % (what, STATUS_NAMES.get(status, "?"), status))


class Settings(ctypes.Structure):
    _fields_ = [
        ("gravity", ctypes.c_float * 3),
        ("dt", ctypes.c_float),
        ("iteration_limit", ctypes.c_int32),
        ("min_iteration_count", ctypes.c_int32),
        ("min_impulse", ctypes.c_float),
        ("allowed_penetration", ctypes.c_float),
        ("default_margin", ctypes.c_float),
        ("max_contact_distance", ctypes.c_float),
        ("contact_invalidation_length_sq", ctypes.c_float),
        ("contact_min_separation_sq", ctypes.c_float),
        ("bounce_velocity_threshold", ctypes.c_float),
        ("max_penetration_recovery_speed", ctypes.c_float),
        ("penetration_recovery_stiffness", ctypes.c_float),
        ("static_friction_velocity_threshold", ctypes.c_float),
        ("twist_friction_factor", ctypes.c_float),
        ("softness", ctypes.c_float),
        ("use_stabilization", ctypes.c_int32),
        ("velocity_lower_limit", ctypes.c_float),
        ("low_velocity_time_minimum", ctypes.c_float),
        ("linear_damping", ctypes.c_float),
        ("angular_damping", ctypes.c_float),
        ("max_deactivation_attempts", ctypes.c_int32), ("core_shape_scaling", ctypes.c_float),
    ]


class MotionState(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_float * 3),
        ("orientation", ctypes.c_float * 4),
        ("linear_velocity", ctypes.c_float * 3),
        ("angular_velocity", ctypes.c_float * 3),
    ]


class Material(ctypes.Structure):
    _fields_ = [("static_friction", ctypes.c_float), ("kinetic_friction", ctypes.c_float),
                ("bounciness", ctypes.c_float)]


class Contact(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("normal", ctypes.c_float * 3),
                ("penetration_depth", ctypes.c_float), ("id", ctypes.c_int32)]


class Manifold(ctypes.Structure):
    _fields_ = [("key", ctypes.c_uint64), ("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("count", ctypes.c_int32), ("kind", ctypes.c_int32), ("contacts", Contact * 4)]


class Impulses(ctypes.Structure):
    _fields_ = [("key", ctypes.c_uint64), ("color", ctypes.c_int32), ("count", ctypes.c_int32),
                ("normal", ctypes.c_float * 4), ("friction", ctypes.c_float * 4),
                ("sliding", ctypes.c_float * 2), ("twist", ctypes.c_float), ("pad", ctypes.c_float)]


class Tuning(ctypes.Structure):
    _fields_ = [("narrowphase_block", ctypes.c_uint32), ("convex_lds", ctypes.c_uint32),
                ("convex_side_stream", ctypes.c_uint32), ("small_island_max", ctypes.c_uint32),
                ("step_graph", ctypes.c_uint32), ("ms_per_step", ctypes.c_float)]


class StepStats(ctypes.Structure):
    _fields_ = [("bodies", ctypes.c_uint32), ("pairs", ctypes.c_uint32), ("manifolds", ctypes.c_uint32),
                ("contacts", ctypes.c_uint32), ("colors", ctypes.c_uint32), ("steps", ctypes.c_uint32),
                ("max_penetration", ctypes.c_float), ("global_manifolds", ctypes.c_uint32),
                ("convex_pairs", ctypes.c_uint32 * 3), ("mesh_pairs", ctypes.c_uint32),
                ("mesh_testers", ctypes.c_uint32)]


# Every entry point of include/bepu_gpu.h: name -> (restype, argtypes)
class BodyDesc(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("dims", ctypes.c_float * 3), ("mass", ctypes.c_float),
                ("state", MotionState), ("material", Material), ("flags", ctypes.c_uint32)]


_vp = ctypes.c_void_p
_u32p = ctypes.POINTER(ctypes.c_uint32)
EXPORTS = {
    "bepu_settings_default": (None, [ctypes.POINTER(Settings)]),
    "bepu_world_create": (ctypes.c_int, [ctypes.POINTER(Settings), ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "bepu_world_destroy": (ctypes.c_int, [_vp]),
    "bepu_add_box": (ctypes.c_int, [_vp, ctypes.POINTER(MotionState), ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, ctypes.POINTER(Material), ctypes.c_uint32,
                                    _u32p]),
    "bepu_add_sphere": (ctypes.c_int, [_vp, ctypes.POINTER(MotionState), ctypes.c_float, ctypes.c_float,
                                       ctypes.POINTER(Material), ctypes.c_uint32, _u32p]),
    "bepu_add_capsule": (ctypes.c_int, [_vp, ctypes.POINTER(MotionState), ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.POINTER(Material), ctypes.c_uint32, _u32p]),
    "bepu_compute_hull": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, _u32p,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float)]),
    "bepu_create_hull_shape": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, _u32p,
                                              ctypes.POINTER(ctypes.c_float)]),
    "bepu_add_convex_hull": (ctypes.c_int, [_vp, ctypes.POINTER(MotionState), ctypes.c_uint32, ctypes.c_float,
                                            ctypes.POINTER(Material), ctypes.c_uint32, _u32p]),
    "bepu_create_mesh_shape": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32,
                                              ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32,
                                              ctypes.POINTER(ctypes.c_float), _u32p]),
    "bepu_add_static_mesh": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.POINTER(Material), _u32p]),
    "bepu_create_terrain_shape": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32,
                                                 ctypes.c_uint32, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                                 _u32p]),
    "bepu_add_terrain": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_int32, ctypes.POINTER(Material), _u32p]),
    "bepu_snapshot_size": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "bepu_save_snapshot": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "bepu_restore_snapshot": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    "bepu_step": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "bepu_get_motion_states": (ctypes.c_int, [_vp, ctypes.POINTER(MotionState), ctypes.c_uint32, _u32p]),
    "bepu_get_pairs": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32, _u32p]),
    "bepu_get_manifolds": (ctypes.c_int, [_vp, ctypes.POINTER(Manifold), ctypes.c_uint32, _u32p]),
    "bepu_get_impulses": (ctypes.c_int, [_vp, ctypes.POINTER(Impulses), ctypes.c_uint32, _u32p]),
    "bepu_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(StepStats)]),
    "bepu_get_stage_times": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, _u32p,
                                            ctypes.POINTER(ctypes.c_char_p)]),
    "bepu_set_timing": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "bepu_set_broadphase": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "bepu_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(Tuning)]),
    "bepu_set_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(Tuning)]),
    "bepu_autotune": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(Tuning), _u32p]),
    "bepu_begin_world": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32)]),
    "bepu_insert_world": (ctypes.c_int, [_vp, ctypes.POINTER(BodyDesc), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "bepu_remove_world": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "bepu_synchronize": (ctypes.c_int, [_vp]),
    "bepu_insert_worlds": (ctypes.c_int, [_vp, ctypes.POINTER(BodyDesc), ctypes.POINTER(ctypes.c_uint32),
                                          ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                          ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "bepu_remove_worlds": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]),
    "bepu_get_motion_states_ranges": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32),
                                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                                                     ctypes.POINTER(MotionState)]),
    "bepu_get_motion_states_range": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32,
                                                    ctypes.POINTER(MotionState)]),
    "bepu_get_world_range": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(ctypes.c_uint32)]),
    "bepu_world_create_ex": (ctypes.c_int, [ctypes.POINTER(Settings), ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.POINTER(_vp)]),
    "bepu_clear_error": (ctypes.c_int, [_vp]),
    "bepu_remove": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "bepu_set_motion_states": (ctypes.c_int, [_vp, _u32p, ctypes.POINTER(MotionState), ctypes.c_uint32,
                                              ctypes.c_uint32]),
    "bepu_get_settings": (ctypes.c_int, [_vp, ctypes.POINTER(Settings)]),
    "bepu_set_settings": (ctypes.c_int, [_vp, ctypes.POINTER(Settings)]),
    "bepu_get_motion_states_ranges_noalloc": (ctypes.c_int, [_vp, _u32p, _u32p, ctypes.c_uint32,
                                                             ctypes.POINTER(MotionState)]),
    "bepu_request_motion_states_ranges": (ctypes.c_int, [_vp, _u32p, _u32p, ctypes.c_uint32, _u32p]),
    "bepu_collect_motion_states": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(MotionState)]),
    "bepu_get_faulted_worlds": (ctypes.c_int, [_vp, _u32p, ctypes.c_uint32, _u32p]),
    "bepu_get_activity": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32, _u32p]),
    "bepu_activate": (ctypes.c_int, [_vp, _u32p, ctypes.c_uint32]),
    "bepu_set_position_update_mode": (ctypes.c_int, [_vp, _u32p, ctypes.c_uint32, ctypes.c_int32]),
    "bepu_add_ball_socket": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                                            _u32p]),
    "bepu_set_joint_settings": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "bepu_get_joint_impulses": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, _u32p]),
    "bepu_compute_hull_minimum_radius": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.c_uint32,
                                                        ctypes.POINTER(ctypes.c_float)]),
    "bepu_device_available": (ctypes.c_int32, []),
    "bepu_status_string": (ctypes.c_char_p, [ctypes.c_int]),
}

_lib = None


def load():
    """Loads the library (no compute call is made)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libbepu_gpu.so is not built (%s); run `python -m bepuphysics1_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what):
    if status != BEPU_OK:
        raise BepuError(status, what)


def default_settings():
    s = Settings()
    load().bepu_settings_default(ctypes.byref(s))
    return s
