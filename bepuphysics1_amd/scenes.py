"""Seeded synthetic scenes for the step path.

A scene is a numpy structured array, one row per body, in add order:

    shape     int32    BEPU_SHAPE_BOX (0) / SPHERE (1) / CONVEX_HULL (2) / CAPSULE (3)
    dims      f32[3]   box (width, height, length); sphere (radius, 0, 0);
                       capsule (length, radius, 0); hull (index into hull_library(), 0, 0);
                       static mesh (index into the scene's mesh list, sidedness,
                       1 to improve boundary behaviour)
    mass      f32      <= 0 makes a kinematic body (Entity(shape, mass), EntityBase.cs:632-645)
    state     f32[13]  MotionState: position xyz, orientation xyzw, linear v, angular v
    material  f32[3]   static friction, kinetic friction, bounciness
    flags     uint32   BEPU_FLAG_*

The same array feeds the GPU library (through ``Space``) and the oracle, so
both run on identical inputs.  Hull bodies share the shapes of one seeded
library (``hull_library()``): register it with ``Space.register_hulls``
before adding a scene that has hulls.  Scenes follow the reference's demo scenes
(BEPUphysicsDemos/Demos/PyramidDemo.cs:17-37, BoxBoxTestDemo.cs:18-51) at the
sizes of the benchmark workloads.
"""
import numpy as np

SHAPE_BOX = 0
SHAPE_SPHERE = 1
SHAPE_CONVEX_HULL = 2
SHAPE_CAPSULE = 3
SHAPE_STATIC_MESH = 4
SHAPE_TERRAIN = 5
FLAG_ALWAYS_ACTIVE = 1
FLAG_NO_STABILIZATION = 2
FLAG_CCD_PASSIVE = 4      # PositionUpdateMode.Passive
FLAG_CCD_CONTINUOUS = 8   # PositionUpdateMode.Continuous

BODY_DTYPE = np.dtype([
    ("shape", np.int32),
    ("dims", np.float32, 3),
    ("mass", np.float32),
    ("state", np.float32, 13),
    ("material", np.float32, 3),
    ("flags", np.uint32),
])

IDENTITY = (0.0, 0.0, 0.0, 1.0)

HULL_LIBRARY_SIZE = 256
_HULL_LIBRARY = None


def hull_points(rng, n=16, scale=(0.5, 1.5)):
    """n random points on the unit sphere scaled by a factor uniform in `scale`."""
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return (p * rng.uniform(*scale)).astype(np.float32)


def hull_library():
    """The seeded hull shapes scenes refer to: 16 random points on the unit
    sphere scaled by [0.5, 1.5] each (the hulls of the stress workload)."""
    global _HULL_LIBRARY
    if _HULL_LIBRARY is None:
        rng = np.random.default_rng(0x4855)
        _HULL_LIBRARY = [hull_points(rng) for _ in range(HULL_LIBRARY_SIZE)]
    return _HULL_LIBRARY


def has_meshes(scene):
    return bool(np.isin(np.asarray(scene["shape"]), (SHAPE_STATIC_MESH, SHAPE_TERRAIN)).any())


class TerrainDesc(tuple):
    """A terrain in the mesh list that goes with a scene: heights[width,
    length] (column x along the first axis, row z along the second), the
    QuadTriangleOrganization (0 BottomLeftUpperRight, 1 BottomRightUpperLeft)
    and the affine transform (3x3 linear part as Matrix3x3 rows, translation)."""
    __slots__ = ()

    def __new__(cls, heights, organization, transform):
        lin = np.asarray(transform[0], np.float32).reshape(3, 3)
        tr = np.asarray(transform[1], np.float32).reshape(3)
        return tuple.__new__(cls, (np.ascontiguousarray(heights, np.float32), int(organization), (lin, tr)))

    heights = property(lambda self: self[0])
    organization = property(lambda self: self[1])
    transform = property(lambda self: self[2])


def terrain(k, improve_boundary=True, material=(1.0, 1.0, 0.0)):
    """A Terrain row: `k` indexes the mesh list entry (a TerrainDesc) it uses."""
    return _row(SHAPE_TERRAIN, (float(k), 1.0 if improve_boundary else 0.0, 0.0), 0.0, (0.0, 0.0, 0.0),
                material=material, flags=0)


def heightfield_terrain(width=17, length=17, spacing=2.0, amplitude=0.0, seed=0, center=(0.0, 0.0), y=0.0,
                        organization=0, linear=None):
    """A width x length heightfield with `spacing` between samples, centred
    on `center` (x, z), heights y plus a seeded smooth bump field; `linear`
    (3x3) replaces the plain scaling.  As a TerrainDesc."""
    rng = np.random.default_rng(seed)
    gx, gz = np.meshgrid(np.arange(width) * spacing, np.arange(length) * spacing, indexing="ij")
    h = np.full(gx.shape, 0.0)
    if amplitude:
        size = spacing * max(width, length)
        for _ in range(4):
            k = rng.uniform(0.5, 2.0, 2) * 2 * np.pi / size
            ph = rng.uniform(0, 2 * np.pi, 2)
            h += amplitude / 4 * np.sin(k[0] * gx + ph[0]) * np.cos(k[1] * gz + ph[1])
    lin = np.diag([spacing, 1.0, spacing]) if linear is None else np.asarray(linear, np.float64)
    tr = (center[0] - (width - 1) * spacing / 2, y, center[1] - (length - 1) * spacing / 2)
    return TerrainDesc(h, organization, (lin, tr))


def terrain_pile(n=64, seed=0, shapes=(SHAPE_BOX, SHAPE_SPHERE, SHAPE_CONVEX_HULL, SHAPE_CAPSULE), drop=(1.0, 6.0),
                 spacing=2.0, amplitude=1.0, organization=0, improve_boundary=True):
    """A random pile of `shapes` dropped onto a terrain (the reference's
    TerrainDemo).  Returns (scene, meshes)."""
    rng = np.random.default_rng(seed)
    foot = 2.0 * np.sqrt(max(n, 1)) + 2.0
    samples = int(np.ceil((foot + 8.0) / spacing)) + 1
    meshes = [heightfield_terrain(samples, samples, spacing, amplitude, seed=seed, organization=organization)]
    rows = [terrain(0, improve_boundary)]
    for _ in range(n):
        s = int(rng.choice(shapes))
        pos = (rng.uniform(-foot / 2, foot / 2), amplitude + rng.uniform(*drop), rng.uniform(-foot / 2, foot / 2))
        rows.append(_row(s, _shape_dims(rng, s), 1.0, pos, quat=tuple(random_quaternions(rng, 1)[0])))
    return make(rows), meshes


def static_mesh(mesh, sidedness=0, improve_boundary=True, material=(1.0, 1.0, 0.0)):
    """A StaticMesh row: `mesh` indexes the mesh list that goes with the scene
    (each entry (vertices [n, 3], indices [3k], transform or None), as
    heightfield_mesh returns them)."""
    return _row(SHAPE_STATIC_MESH, (float(mesh), float(sidedness), 1.0 if improve_boundary else 0.0), 0.0,
                (0.0, 0.0, 0.0), material=material, flags=0)


def heightfield_mesh(cells=16, size=32.0, amplitude=0.0, seed=0, center=(0.0, 0.0), y=0.0, transform=None):
    """A cells x cells grid of quads over a size x size square, two triangles
    per quad (counterclockwise seen from above), heights y plus a seeded
    smooth bump field of the given amplitude; as (vertices, indices, transform)."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(-size / 2, size / 2, cells + 1)
    gx, gz = np.meshgrid(xs + center[0], xs + center[1])
    h = np.full_like(gx, y)
    if amplitude:
        for _ in range(4):
            k = rng.uniform(0.5, 2.0, 2) * 2 * np.pi / size
            ph = rng.uniform(0, 2 * np.pi, 2)
            h += amplitude / 4 * np.sin(k[0] * gx + ph[0]) * np.cos(k[1] * gz + ph[1])
    verts = np.stack([gx.ravel(), h.ravel(), gz.ravel()], 1).astype(np.float32)
    idx = []
    n = cells + 1
    for j in range(cells):
        for i in range(cells):
            a = j * n + i
            b, c, d = a + 1, a + n, a + n + 1
            idx += [a, c, b, b, c, d]
    return verts, np.asarray(idx, np.int32), transform


def mesh_pile(n=64, seed=0, shapes=(SHAPE_BOX, SHAPE_SPHERE, SHAPE_CONVEX_HULL, SHAPE_CAPSULE), drop=(1.0, 6.0),
              cells=16, amplitude=1.0, sidedness=0, improve_boundary=True):
    """A random pile of `shapes` dropped onto a static heightfield mesh (the
    reference's StaticMeshDemo / TerrainDemo ground).  Returns (scene, meshes)."""
    rng = np.random.default_rng(seed)
    foot = 2.0 * np.sqrt(max(n, 1)) + 2.0
    size = foot + 8.0
    meshes = [heightfield_mesh(cells, size, amplitude, seed=seed)]
    rows = [static_mesh(0, sidedness, improve_boundary)]
    for _ in range(n):
        s = int(rng.choice(shapes))
        pos = (rng.uniform(-foot / 2, foot / 2), amplitude + rng.uniform(*drop), rng.uniform(-foot / 2, foot / 2))
        rows.append(_row(s, _shape_dims(rng, s), 1.0, pos, quat=tuple(random_quaternions(rng, 1)[0])))
    return make(rows), meshes


def has_hulls(scene):
    return bool((np.asarray(scene["shape"]) == SHAPE_CONVEX_HULL).any())


def _shape_dims(rng, s):
    if s == SHAPE_BOX:
        return 2 * rng.uniform(0.25, 1.0, 3)
    if s == SHAPE_SPHERE:
        return (rng.uniform(0.25, 1.0), 0.0, 0.0)
    if s == SHAPE_CAPSULE:
        return (rng.uniform(0.5, 2.0), rng.uniform(0.2, 0.6), 0.0)
    return (float(rng.integers(0, HULL_LIBRARY_SIZE)), 0.0, 0.0)


def _row(shape, dims, mass, pos, quat=IDENTITY, vel=(0, 0, 0), angvel=(0, 0, 0),
         material=(1.0, 0.8, 0.0), flags=FLAG_ALWAYS_ACTIVE):
    r = np.zeros((), BODY_DTYPE)
    r["shape"] = shape
    r["dims"] = dims
    r["mass"] = mass
    r["state"] = list(pos) + list(quat) + list(vel) + list(angvel)
    r["material"] = material
    r["flags"] = flags
    return r


def make(rows):
    out = np.zeros(len(rows), BODY_DTYPE)
    for i, r in enumerate(rows):
        out[i] = r
    return out


def ground(width=50.0, height=1.0, length=50.0, top=0.0, center=(0.0, 0.0), material=(1.0, 1.0, 0.0)):
    """Static (kinematic) ground box whose top face sits at y = top."""
    return _row(SHAPE_BOX, (width, height, length), 0.0, (center[0], top - height / 2, center[1]),
                material=material, flags=0)


def pyramid(rows=10, gap=0.01, friction=0.5):
    """Pyramid of unit boxes, mass 1, on a 50x1x50 static ground.

    Row r holds rows - r boxes; interaction friction 0.5 (static and kinetic)
    and bounciness 0 come from box material (0.5, 0.5, 0) times ground material
    (1, 1, 0) under the product blend (MaterialManager.cs:77-85).
    """
    out = [ground()]
    for r in range(rows):
        n = rows - r
        for j in range(n):
            x = (j - (n - 1) / 2.0) * (1.0 + gap)
            out.append(_row(SHAPE_BOX, (1.0, 1.0, 1.0), 1.0, (x, 0.5 + r, 0.0),
                            material=(friction, friction, 0.0)))
    return make(out)


def towers(nx=64, nz=64, height=20, spacing=2.0, jitter=0.01, half_extent=0.5, seed=0):
    """nx x nz grid of box towers (height boxes each) on one static ground.

    Every box has half extent 0.5, mass 1 and a seeded horizontal jitter in
    [-jitter, jitter].  At spacing 2.0 no two towers touch, so every tower
    is an independent 20-deep contact chain.
    """
    rng = np.random.default_rng(seed)
    size = max(nx, nz) * spacing + 10.0
    n = nx * nz * height
    out = np.zeros(n + 1, BODY_DTYPE)
    out[0] = ground(size, 1.0, size)
    ix, iz, lv = np.meshgrid(np.arange(nx), np.arange(nz), np.arange(height), indexing="ij")
    ix, iz, lv = ix.ravel(), iz.ravel(), lv.ravel()
    jx = rng.uniform(-jitter, jitter, n).astype(np.float32)
    jz = rng.uniform(-jitter, jitter, n).astype(np.float32)
    b = out[1:]
    b["shape"] = SHAPE_BOX
    b["dims"] = 2 * half_extent
    b["mass"] = 1.0
    st = np.zeros((n, 13), np.float32)
    st[:, 0] = (ix - (nx - 1) / 2.0) * spacing + jx
    st[:, 1] = half_extent + lv * (2 * half_extent)
    st[:, 2] = (iz - (nz - 1) / 2.0) * spacing + jz
    st[:, 6] = 1.0
    b["state"] = st
    b["material"] = (1.0, 0.8, 0.0)
    b["flags"] = FLAG_ALWAYS_ACTIVE
    return out


def random_quaternions(rng, n):
    """Uniform random unit quaternions (Shoemake's method), xyzw."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    return q.astype(np.float32)


def random_pile(n, seed=0, shapes=(SHAPE_BOX, SHAPE_SPHERE), drop=(1.0, 10.0), with_ground=True):
    """A random pile of n bodies dropped onto a ground, mixed shapes.

    Boxes have half extents in [0.25, 1], spheres radius in [0.25, 1]; bodies
    are spread over a footprint that grows with sqrt(n) and dropped from
    heights uniform in `drop`.
    """
    rng = np.random.default_rng(seed)
    foot = 2.0 * np.sqrt(max(n, 1)) + 2.0
    out = np.zeros(n + (1 if with_ground else 0), BODY_DTYPE)
    k = 0
    if with_ground:
        out[0] = ground(foot + 20.0, 1.0, foot + 20.0)
        k = 1
    kinds = rng.choice(np.array(shapes), size=n)
    q = random_quaternions(rng, n)
    for i in range(n):
        s = int(kinds[i])
        dims = _shape_dims(rng, s)
        pos = (rng.uniform(-foot / 2, foot / 2), rng.uniform(*drop), rng.uniform(-foot / 2, foot / 2))
        out[k + i] = _row(s, dims, 1.0, pos, quat=q[i])
    return out


STRESS_MIX = (SHAPE_BOX, SHAPE_BOX, SHAPE_SPHERE, SHAPE_CONVEX_HULL)


def stress(n=32768, edge=None, seed=0, shapes=(SHAPE_BOX, SHAPE_SPHERE)):
    """Bodies placed uniformly at random in a cube of side `edge` with random
    orientations, no ground, shapes in equal shares of `shapes` (STRESS_MIX:
    50% boxes, 25% spheres, 25% hulls).  Box half extents and sphere radii in
    [0.25, 1]; hulls from hull_library().  The default edge gives a mean AABB
    degree of about 4 (74 for boxes and spheres, 71.2 with hulls, at 32,768
    bodies, scaled with the cube root of n)."""
    rng = np.random.default_rng(seed)
    if edge is None:
        # Mean AABB degree about 4 (about 2 pairs per body) for either mix.
        base = 71.2 if SHAPE_CONVEX_HULL in shapes else 74.0
        edge = base * (n / 32768.0) ** (1.0 / 3.0)
    out = np.zeros(n, BODY_DTYPE)
    kinds = np.array(shapes)[np.arange(n) % len(shapes)]
    rng.shuffle(kinds)
    q = random_quaternions(rng, n)
    pos = rng.uniform(0.0, edge, (n, 3)).astype(np.float32)
    for i in range(n):
        s = int(kinds[i])
        out[i] = _row(s, _shape_dims(rng, s), 1.0, pos[i], quat=q[i])
    return out


def mixed_sizes(n=2048, seed=0, edge=200.0, offset=(0.0, 0.0, 0.0), static_fraction=0.1, max_half=10.0):
    """Boxes and spheres whose sizes span more than two orders of magnitude
    (half extents / radii log-uniform in [0.05, max_half]) at random positions and
    orientations, a tenth of them static: exercises every level of the grid
    broadphase and the kinematic-kinematic rule.  `offset` moves the scene
    away from the origin."""
    rng = np.random.default_rng(seed)
    out = np.zeros(n, BODY_DTYPE)
    q = random_quaternions(rng, n)
    pos = (rng.uniform(0.0, edge, (n, 3)) + np.asarray(offset, np.float64)).astype(np.float32)
    for i in range(n):
        size = np.exp(rng.uniform(np.log(0.05), np.log(max_half), 3))
        mass = 0.0 if rng.uniform() < static_fraction else 1.0
        if i % 2 == 0:
            out[i] = _row(SHAPE_BOX, 2 * size, mass, pos[i], quat=q[i])
        else:
            out[i] = _row(SHAPE_SPHERE, (size[0], 0.0, 0.0), mass, pos[i], quat=q[i])
    return out


def batched_worlds(n_worlds=4096, seed=0, nmin=8, nmax=256, shapes=(SHAPE_BOX, SHAPE_SPHERE), drop=(1.0, 10.0)):
    """Independent worlds for the batched workload: world k is a random pile
    (its own ground) of N_k bodies, N_k log-uniform in [nmin, nmax], dropped
    from heights uniform in `drop`.  Returns a list of scene arrays."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_worlds)).astype(np.int64)
    seeds = rng.integers(0, 2**31 - 1, n_worlds)
    return [random_pile(int(n), seed=int(s), shapes=shapes, drop=drop) for n, s in zip(sizes, seeds)]


def random_pile_fast(n, rng, drop=(1.0, 10.0), sphere_fraction=0.5, hull_fraction=0.0):
    """Vectorised random pile (same distribution family as random_pile):
    boxes with half extents in [0.25, 1], spheres with radius in [0.25, 1]
    and hulls from hull_library(), uniform random orientations, over a
    sqrt(n) footprint, dropped from heights uniform in `drop`, plus the
    world's own ground."""
    foot = 2.0 * np.sqrt(max(n, 1)) + 2.0
    out = np.zeros(n + 1, BODY_DTYPE)
    out[0] = ground(foot + 20.0, 1.0, foot + 20.0)
    body = out[1:]
    u = rng.uniform(size=n)
    sph = u < sphere_fraction
    hul = (u >= sphere_fraction) & (u < sphere_fraction + hull_fraction)
    body["shape"] = np.where(sph, SHAPE_SPHERE, np.where(hul, SHAPE_CONVEX_HULL, SHAPE_BOX))
    dims = 2 * rng.uniform(0.25, 1.0, (n, 3))
    rad = rng.uniform(0.25, 1.0, n)
    dims[sph, 0] = rad[sph]
    dims[sph, 1:] = 0.0
    hid = rng.integers(0, HULL_LIBRARY_SIZE, n)
    dims[hul, 0] = hid[hul]
    dims[hul, 1:] = 0.0
    body["dims"] = dims
    body["mass"] = 1.0
    st = np.zeros((n, 13), np.float32)
    st[:, 0] = rng.uniform(-foot / 2, foot / 2, n)
    st[:, 1] = rng.uniform(drop[0], drop[1], n)
    st[:, 2] = rng.uniform(-foot / 2, foot / 2, n)
    st[:, 3:7] = random_quaternions(rng, n)
    body["state"] = st
    body["material"] = (1.0, 0.8, 0.0)
    body["flags"] = FLAG_ALWAYS_ACTIVE
    return out


def shooting_gallery(n=64, seed=0, speed=(120.0, 480.0), shapes=(SHAPE_BOX, SHAPE_SPHERE, SHAPE_CONVEX_HULL,
                                                                  SHAPE_CAPSULE), modes=(0.6, 0.2)):
    """Fast projectiles fired at thin walls (continuous collision detection):
    a ground, a 0.2-thick kinematic wall across x = 0 and another at x = 12,
    and n projectiles between them with speeds uniform in `speed` (2-8 m per
    step at 60 Hz) along -x or +x with small sideways components, half of
    them flying at each wall.  Projectile sizes are small next to the travel
    per step (box half extents and radii in [0.08, 0.25]; hulls from
    hull_library()); a fraction modes[0] is continuous, modes[1] passive and
    the rest discrete.  Continuous projectiles stop at the walls; the
    discrete ones tunnel."""
    rng = np.random.default_rng(seed)
    rows = [ground(60.0, 1.0, 60.0),
            _row(SHAPE_BOX, (0.2, 8.0, 16.0), 0.0, (0.0, 4.0, 0.0), flags=0),
            _row(SHAPE_BOX, (0.2, 8.0, 16.0), 0.0, (12.0, 4.0, 0.0), flags=0)]
    q = random_quaternions(rng, n)
    for i in range(n):
        s = int(rng.choice(np.array(shapes)))
        if s == SHAPE_SPHERE:
            dims = (rng.uniform(0.08, 0.25), 0.0, 0.0)
        elif s == SHAPE_BOX:
            dims = tuple(2 * rng.uniform(0.08, 0.25, 3))
        elif s == SHAPE_CAPSULE:
            dims = (rng.uniform(0.1, 0.4), rng.uniform(0.08, 0.2), 0.0)
        else:
            dims = (float(rng.integers(0, HULL_LIBRARY_SIZE)), 0.0, 0.0)
        left = i % 2 == 0
        x = rng.uniform(2.0, 10.0)
        v = rng.uniform(*speed) * (-1.0 if left else 1.0)
        vel = (v, rng.uniform(-10.0, 10.0), rng.uniform(-10.0, 10.0))
        u = rng.uniform()
        mode = FLAG_CCD_CONTINUOUS if u < modes[0] else (FLAG_CCD_PASSIVE if u < modes[0] + modes[1] else 0)
        rows.append(_row(s, dims, 1.0, (x, rng.uniform(1.0, 7.0), rng.uniform(-6.0, 6.0)), quat=q[i], vel=vel,
                         angvel=tuple(rng.uniform(-5, 5, 3)), flags=FLAG_ALWAYS_ACTIVE | mode))
    return make(rows)


def chains(n_chains=6, links=8, seed=0, shapes=(SHAPE_BOX, SHAPE_SPHERE, SHAPE_CONVEX_HULL, SHAPE_CAPSULE),
           flags=FLAG_ALWAYS_ACTIVE, spacing=3.0):
    """Chains of `links` bodies hung from kinematic anchors by ball-socket
    joints (BallSocketJoint), started sideways so that they swing down onto a
    ground and a few loose boxes.  Returns (scene, joints) where joints lists
    (body a, body b, anchor xyz) in the order to add them."""
    rng = np.random.default_rng(seed)
    rows = [ground(80.0, 1.0, 80.0)]
    joints = []
    half = 0.75  # half the distance between joints along a chain (hulls are up to 1.5 across)
    top = links * 2 * half + 2.0
    for c in range(n_chains):
        x0 = (c - n_chains / 2) * spacing
        anchor_id = len(rows)
        rows.append(_row(SHAPE_BOX, (0.3, 0.3, 0.3), 0.0, (x0, top, 0.0), flags=0))
        prev, prev_end = anchor_id, np.array([x0, top, 0.0])
        for k in range(links):
            s = int(rng.choice(np.array(shapes)))
            # lay the chain out sideways from the anchor
            centre = prev_end + np.array([half, 0.0, 0.0])
            if s == SHAPE_SPHERE:
                dims = (0.25, 0.0, 0.0)
            elif s == SHAPE_BOX:
                dims = (0.5, 0.3, 0.3)
            elif s == SHAPE_CAPSULE:
                dims = (0.3, 0.15, 0.0)
            else:
                dims = (float(rng.integers(0, HULL_LIBRARY_SIZE)), 0.0, 0.0)
            bid = len(rows)
            rows.append(_row(s, dims, float(rng.uniform(0.5, 2.0)), tuple(centre), flags=flags))
            joints.append((prev, bid, tuple(float(v) for v in prev_end)))
            prev, prev_end = bid, centre + (centre - prev_end)
    for k in range(4):
        rows.append(_row(SHAPE_BOX, (0.8, 0.8, 0.8), 1.0, ((k - 2) * 2.0, 0.4, 3.0), flags=flags))
    return make(rows), joints
