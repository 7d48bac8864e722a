"""MI355X-native (gfx950) implementation of BEPUphysics v1's per-step
rigid-body path: broadphase, convex narrowphase, sequential-impulse contact
solver and position integration, behind a C ABI (include/bepu_gpu.h).
"""
from . import scenes
from ._lib import BepuError, load as load_library
from .space import (Box, Capsule, ConvexHull, Entity, Material, Space, Sphere, StaticMesh, Terrain, compute_hull,
                    compute_hull_minimum_radius)

__all__ = ["Space", "Entity", "Box", "Sphere", "Capsule", "ConvexHull", "StaticMesh", "Terrain", "Material", "BepuError", "scenes",
           "load_library", "compute_hull", "compute_hull_minimum_radius"]
