"""Record layouts shared by the host code and the C ABI.

Every array crossing the boundary uses the reference's own in-memory record
layout (Jolt/Physics/SoftBody/SoftBodySharedSettings.h), so a Jolt integration
can hand ``settings.mEdgeConstraints.data()`` and friends straight through:

=============  =====  =====================================================
record         bytes  reference
=============  =====  =====================================================
Vertex         28     SoftBodySharedSettings.h:127-138 (pos, vel, invMass)
Face           16     SoftBodySharedSettings.h:141-154
Edge           16     SoftBodySharedSettings.h:157-171
DihedralBend   24     SoftBodySharedSettings.h:191-205
Volume         24     SoftBodySharedSettings.h:208-222
LRA            12     SoftBodySharedSettings.h:289-303
UpdateGroup    28     SoftBodySharedSettings.h:374-383 (end indices)
=============  =====  =====================================================
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

VERTEX = np.dtype([("pos", "<f4", 3), ("vel", "<f4", 3), ("inv_mass", "<f4")])
FACE = np.dtype([("v", "<u4", 3), ("material", "<u4")])
EDGE = np.dtype([("v", "<u4", 2), ("rest", "<f4"), ("compliance", "<f4")])
BEND = np.dtype([("v", "<u4", 4), ("compliance", "<f4"), ("initial_angle", "<f4")])
VOLUME = np.dtype([("v", "<u4", 4), ("six_rest_volume", "<f4"), ("compliance", "<f4")])
LRA = np.dtype([("v", "<u4", 2), ("max_distance", "<f4")])
GROUP = np.dtype([("edge_end", "<u4"), ("lra_end", "<u4"), ("rod_ss_end", "<u4"),
                  ("rod_bt_end", "<u4"), ("bend_end", "<u4"), ("volume_end", "<u4"),
                  ("skinned_end", "<u4")])

# Skinned (SoftBodySharedSettings.h:252-285) and InvBind (:225-236, Mat44 16-byte aligned)
SKINNED = np.dtype([("vertex", "<u4"), ("weights", [("inv_bind", "<u4"), ("weight", "<f4")], 4),
                    ("max_distance", "<f4"), ("back_stop_distance", "<f4"), ("back_stop_radius", "<f4"),
                    ("normal_info", "<u4")])
INV_BIND = np.dtype([("joint", "<u4"), ("pad", "<u4", 3), ("m", "<f4", 16)])     # m: column major
assert SKINNED.itemsize == 52 and INV_BIND.itemsize == 80
# Cosserat rods in the reference's memory layout (SoftBodySharedSettings.h:312-342)
ROD_SS = np.dtype([("vertex", "<u4", 2), ("length", "<f4"), ("inv_mass", "<f4"), ("compliance", "<f4"),
                   ("pad", "<u4", 3), ("bishop", "<f4", 4)])
ROD_BT = np.dtype([("rod", "<u4", 2), ("compliance", "<f4"), ("pad", "<u4"), ("omega0", "<f4", 4)])
assert ROD_SS.itemsize == 48 and ROD_BT.itemsize == 32

assert VERTEX.itemsize == 28 and FACE.itemsize == 16 and EDGE.itemsize == 16
assert BEND.itemsize == 24 and VOLUME.itemsize == 24 and LRA.itemsize == 12
assert GROUP.itemsize == 28

# EBendType / ELRAType (SoftBodySharedSettings.h:21-34)
BEND_NONE, BEND_DISTANCE, BEND_DIHEDRAL = 0, 1, 2
LRA_NONE, LRA_EUCLIDEAN, LRA_GEODESIC = 0, 1, 2

# Collider shape types of the C ABI
SHAPE_PLANE, SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CYLINDER, SHAPE_TAPERED_CAPSULE = 0, 1, 2, 3, 4, 5
SHAPE_CONVEX_HULL = 6
SHAPE_TAPERED_CYLINDER = 7      # device params {top, bottom, top radius, bottom radius}
SHAPE_MESH = 8                  # triangle tree (MeshShape or TriangleShape)
SHAPE_COMPOUND = 100            # compound / decorated body: `shape` builds it, `parts` describe its leaves

# xpbd_tree_node: the four children's bounds per axis and their references
TREE_NODE = np.dtype([("min", "<f4", (3, 4)), ("max", "<f4", (3, 4)), ("child", "<u4", 4)])
TREE_NODE_REF, TREE_LEAF_REF, TREE_EMPTY_REF = 0x00000000, 0x40000000, 0x80000000
TREE_HEIGHT_FIELD = 1

# placement steps of a compound's leaf (xpbd_part_op)
PART_OP = np.dtype([("kind", "<u4"), ("v", "<f4", 6)])
OP_CULL, OP_TRANSLATE, OP_ROTATE, OP_OFFSET = 1, 2, 3, 4
# EMotionType (MotionType.h:10-15)
MOTION_STATIC, MOTION_KINEMATIC, MOTION_DYNAMIC = 0, 1, 2


def _empty(dtype):
    return np.zeros(0, dtype=dtype)


@dataclass
class SettingsArrays:
    """All arrays of one SoftBodySharedSettings (after Optimize)."""

    vertices: np.ndarray
    faces: np.ndarray = field(default_factory=lambda: _empty(FACE))
    edges: np.ndarray = field(default_factory=lambda: _empty(EDGE))
    bends: np.ndarray = field(default_factory=lambda: _empty(BEND))
    volumes: np.ndarray = field(default_factory=lambda: _empty(VOLUME))
    lras: np.ndarray = field(default_factory=lambda: _empty(LRA))
    groups: np.ndarray = field(default_factory=lambda: _empty(GROUP))
    skinned: np.ndarray = field(default_factory=lambda: _empty(SKINNED))
    skinned_normals: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    inv_binds: np.ndarray = field(default_factory=lambda: _empty(INV_BIND))
    rod_stretch_shear: np.ndarray = field(default_factory=lambda: _empty(ROD_SS))
    rod_bend_twist: np.ndarray = field(default_factory=lambda: _empty(ROD_BT))

    def counts(self):
        return dict(V=len(self.vertices), F=len(self.faces), E=len(self.edges), B=len(self.bends),
                    T=len(self.volumes), L=len(self.lras), G=len(self.groups))

    def num_constraints(self):
        return (len(self.edges) + len(self.bends) + len(self.volumes) + len(self.lras)
                + len(self.rod_stretch_shear) + len(self.rod_bend_twist))

    def copy(self):
        return SettingsArrays(*(np.array(a, copy=True) for a in (
            self.vertices, self.faces, self.edges, self.bends, self.volumes, self.lras, self.groups,
            self.skinned, self.skinned_normals, self.inv_binds)))


@dataclass
class BodyParams:
    """SoftBodyCreationSettings fields used by the update (SoftBodyCreationSettings.h:61-72)."""

    num_iterations: int = 5
    linear_damping: float = 0.1
    max_linear_velocity: float = 500.0
    restitution: float = 0.0
    friction: float = 0.2
    pressure: float = 0.0
    gravity_factor: float = 1.0
    vertex_radius: float = 0.0
    update_position: bool = True
    make_rotation_identity: bool = True
    allow_sleeping: bool = True

    def as_floats(self):
        return np.array([self.num_iterations, self.linear_damping, self.max_linear_velocity,
                         self.restitution, self.friction, self.pressure, self.gravity_factor,
                         self.vertex_radius, float(self.update_position),
                         float(self.make_rotation_identity), float(self.allow_sleeping)],
                        dtype=np.float32)


@dataclass
class ConvexHull:
    """A convex hull as ConvexHullShape holds it (ConvexHullShape.h:114-141): points
    relative to the centre of mass, one plane per face (normal xyz, constant), each
    face's first entry and vertex count in `vertex_index` (counter-clockwise), and the
    centre of mass in the shape's frame."""

    points: np.ndarray          # (P, 3) float32
    planes: np.ndarray          # (F, 4) float32
    face_first: np.ndarray      # (F,) uint32
    face_count: np.ndarray      # (F,) uint32
    vertex_index: np.ndarray    # (sum of face_count,) uint32
    center_of_mass: tuple = (0.0, 0.0, 0.0)

    @staticmethod
    def from_polytope(points, faces):
        """Hull data for a convex polytope given by its vertices and counter-clockwise
        faces (lists of vertex indices), in float64 then rounded: centre of mass of the
        solid by a tetrahedron fan, plane of each face through its centroid with the
        normalised Newell normal.  The reference derives these with its own hull
        builder (ConvexHullShape.cpp:50-180), so values agree to rounding, not bits."""
        pts = np.asarray(points, np.float64).reshape(-1, 3)
        faces = [list(map(int, f)) for f in faces]
        vol = 0.0
        acc = np.zeros(3)
        o = pts[faces[0][0]]
        for f in faces:
            for k in range(1, len(f) - 1):
                a, b, c = pts[f[0]] - o, pts[f[k]] - o, pts[f[k + 1]] - o
                v6 = float(np.dot(a, np.cross(b, c)))
                vol += v6
                acc += v6 * (a + b + c) / 4.0
        com = o + acc / vol
        rel = pts - com
        planes, first, count, index = [], [], [], []
        for f in faces:
            q = rel[f]
            n = np.zeros(3)
            for k in range(len(f)):
                a, b = q[k], q[(k + 1) % len(f)]
                n += np.array([(a[1] - b[1]) * (a[2] + b[2]), (a[2] - b[2]) * (a[0] + b[0]),
                               (a[0] - b[0]) * (a[1] + b[1])])
            n /= np.linalg.norm(n)
            planes.append([n[0], n[1], n[2], -float(np.dot(n, q.mean(axis=0)))])
            first.append(len(index))
            count.append(len(f))
            index.extend(f)
        return ConvexHull(rel.astype(np.float32), np.array(planes, np.float32), np.array(first, np.uint32),
                          np.array(count, np.uint32), np.array(index, np.uint32), tuple(float(x) for x in com))


@dataclass
class TriangleTree:
    """A mesh's bounding-volume tree in the form the device walks it (xpbd_triangle_tree):
    root reference, nodes (TREE_NODE), triangles (T, 9) float32 in leaf order."""

    root: int
    nodes: np.ndarray
    triangles: np.ndarray
    flags: int = 0              # TREE_HEIGHT_FIELD: a height field's hierarchy


@dataclass
class ShapeSpec:
    """A shape of a compound or decorated collider, as the reference's shape classes build it:
    kind "basic" (`type`, `params` as Collider takes them), "hull" (`params`: points xyz),
    "triangle" (`params`: three vertices xyz),
    "scaled" (`inner`, `vec` = scale), "rotated_translated" (`inner`, `vec` = position, `rot`),
    "offset_com" (`inner`, `vec` = offset), "static_compound" / "mutable_compound"
    (`children`: [(ShapeSpec, position, rotation)])."""

    kind: str
    type: int = 0
    params: tuple = ()
    inner: "ShapeSpec | None" = None
    vec: tuple = ()
    rot: tuple = (0.0, 0.0, 0.0, 1.0)
    children: list = field(default_factory=list)

    @staticmethod
    def basic(type, *params):
        return ShapeSpec("basic", type=type, params=tuple(params))

    @staticmethod
    def hull(points):
        return ShapeSpec("hull", params=tuple(float(x) for x in np.asarray(points, np.float32).ravel()))

    @staticmethod
    def triangle(v9):
        return ShapeSpec("triangle", params=tuple(float(x) for x in np.asarray(v9, np.float32).ravel()))

    @staticmethod
    def scaled(inner, scale):
        return ShapeSpec("scaled", inner=inner, vec=tuple(scale))

    @staticmethod
    def rotated_translated(inner, position, rotation):
        return ShapeSpec("rotated_translated", inner=inner, vec=tuple(position), rot=tuple(rotation))

    @staticmethod
    def offset_com(inner, offset):
        return ShapeSpec("offset_com", inner=inner, vec=tuple(offset))

    @staticmethod
    def static_compound(children):
        return ShapeSpec("static_compound", children=list(children))

    @staticmethod
    def mutable_compound(children):
        return ShapeSpec("mutable_compound", children=list(children))


@dataclass
class ColliderPart:
    """One leaf of a compound collider as the device takes it (xpbd_collider_part): shape type,
    params (8), scale, placement steps (PART_OP), hull or tree."""

    type: int
    params: tuple
    scale: tuple = (1.0, 1.0, 1.0)
    ops: np.ndarray = field(default_factory=lambda: np.zeros(0, PART_OP))
    hull: ConvexHull | None = None
    tree: TriangleTree | None = None


@dataclass
class Collider:
    """A static collider body: shape, world transform (body position and rotation),
    material, world bounds.

    `params` are the shape's construction parameters (plane {nx, ny, nz, c, half extent},
    sphere {r}, box {hx, hy, hz}, capsule / cylinder {half height, r}, tapered capsule
    {half height of the tapered cylinder, top radius, bottom radius}, convex hull: its
    construction points, flattened xyz).  `device_params`, `hull` and `com` are what the
    soft-body update consumes (shape parameters and hull data as the shape holds them,
    centre-of-mass position); when unset they are derived here, for the tapered capsule
    approximately (the reference derives them with its own sin / tan; the tests take
    them from the reference).  A convex hull needs `hull` (ConvexHull.from_polytope, or
    the reference's own)."""

    type: int
    params: tuple
    pos: tuple = (0.0, 0.0, 0.0)
    rot: tuple = (0.0, 0.0, 0.0, 1.0)
    friction: float = 0.2
    restitution: float = 0.0
    bounds: tuple | None = None  # world-space AABB (min xyz, max xyz)
    device_params: tuple | None = None
    com: tuple | None = None
    hull: ConvexHull | None = None
    # motion: static / kinematic colliders take the static contact branch; a dynamic
    # one needs its mass properties (MotionProperties) and world velocities
    motion_type: int = MOTION_STATIC
    inv_mass: float = 0.0
    inv_inertia_diagonal: tuple = (0.0, 0.0, 0.0)
    inertia_rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    linear_velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    is_sensor: bool = False     # Body::IsSensor(): detects vertices, never collides
    user_id: int | None = None  # caller's identifier (BodyID); None: the object's identity
    scale: tuple = (1.0, 1.0, 1.0)  # ScaledShape around the shape
    # SHAPE_MESH construction: a MeshShape from (vertices (N, 3), indices (M, 3), max triangles
    # per leaf or 0), a TriangleShape from `triangle` (9 floats) or a HeightFieldShape from
    # `height_field` (samples (n, n), block size, bits per sample, offset (3), scale (3)); `tree`
    # is what the device consumes (taken from the reference's shape)
    mesh: tuple | None = None
    triangle: tuple | None = None
    height_field: tuple | None = None
    tree: TriangleTree | None = None
    # SHAPE_COMPOUND: the shape to build (ShapeSpec) and its leaves as the device takes them
    # (ColliderPart list, taken from the reference's shape)
    shape: ShapeSpec | None = None
    parts: list | None = None

    def param_array(self):
        if self.type == SHAPE_CONVEX_HULL:
            return np.zeros(8, dtype=np.float32)
        p = np.zeros(8, dtype=np.float32)
        p[:len(self.params)] = self.params
        return p

    def device_param_array(self):
        if self.device_params is not None:
            p = np.zeros(8, dtype=np.float32)
            p[:len(self.device_params)] = self.device_params
            return p
        if self.type == SHAPE_TAPERED_CAPSULE:
            return tapered_capsule_device_params(*self.params[:3])
        return self.param_array()

    def com_position(self):
        if self.com is not None:
            return tuple(self.com)
        if self.type == SHAPE_TAPERED_CAPSULE:
            p = tapered_capsule_device_params(*self.params[:3])
            off = np.array([0.0, self.params[0] - p[0], 0.0])      # centre of mass in the shape's frame
            return tuple(np.array(self.pos, np.float64) + _rotate(self.rot, off))
        if self.type == SHAPE_CONVEX_HULL:
            if self.hull is None:
                raise ValueError("convex hull collider without hull data")
            return tuple(np.array(self.pos, np.float64) + _rotate(self.rot, self.hull.center_of_mass))
        return tuple(self.pos)

    def hull_points(self):
        """Construction points of a convex hull collider, (P, 3) float32."""
        return np.asarray(self.params, np.float32).reshape(-1, 3)


def tapered_capsule_device_params(half_height, top_radius, bottom_radius):
    """TaperedCapsuleShape's derived parameters {top centre, bottom centre, top radius,
    bottom radius, sin alpha, tan alpha} (TaperedCapsuleShape.cpp:117-131), in float32
    with numpy's tan / arcsin (the reference uses its own approximations)."""
    f = np.float32
    h, t, b = f(half_height), f(top_radius), f(bottom_radius)
    tc = f(h + f(f(0.5) * f(b - t)))
    bc = f(f(-h) + f(f(0.5) * f(b - t)))
    sa = f(f(b - t) / f(tc - bc))
    ta = f(np.tan(np.arcsin(np.float64(sa))))
    p = np.zeros(8, dtype=np.float32)
    p[:6] = (tc, bc, t, b, sa, ta)
    return p


def _rotate(q, v):
    x, y, z, w = (float(c) for c in q)
    u = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    return 2.0 * np.dot(u, v) * u + (w * w - np.dot(u, u)) * v + 2.0 * w * np.cross(u, v)


# BodyCreationSettings defaults (BodyCreationSettings.h:111-112)
MAX_LINEAR_VELOCITY = np.float32(500.0)
MAX_ANGULAR_VELOCITY = np.float32(np.float32(np.float32(0.25) * np.float32(3.14159265358979323846)) * np.float32(60.0))


def add_velocity_clamped(velocity, delta, max_velocity):
    """BodyInterface::AddLinearAndAngularVelocity for one of the two velocities
    (BodyInterface.cpp:640-655): v + delta, clamped to max_velocity in length as
    MotionProperties::ClampLinearVelocity / ClampAngularVelocity do
    (MotionProperties.inl:23-41), in float32."""
    f = np.float32
    v = np.asarray(velocity, f) + np.asarray(delta, f)
    len_sq = f(f(v[0] * v[0]) + f(v[1] * v[1])) + f(v[2] * v[2])
    m = f(max_velocity)
    if len_sq > f(m * m):
        v = v * f(m / np.sqrt(len_sq))
    return tuple(float(x) for x in v)
