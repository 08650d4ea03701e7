"""Deterministic synthetic soft-body scenes (meshes, attributes, colliders).

A scene generator returns *mesh specs*: vertex records plus faces / explicit
constraints and the per-vertex attributes that the settings builder
(CreateConstraints / Calculate* / Optimize, SoftBodySharedSettings.cpp:189-1182)
turns into the constraint arrays.  The same spec is fed to the product's host
builder and, in the tests, to the reference's own builder, so both sides run
on identical inputs.

Random numbers come from numpy's PCG64 with fixed seeds and are drawn as
float64 then rounded once to float32, so the scenes are reproducible across
machines.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

from .records import (BEND_DIHEDRAL, BEND_NONE, EDGE, LRA, LRA_EUCLIDEAN, LRA_NONE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_COMPOUND,
                      SHAPE_CYLINDER, SHAPE_CONVEX_HULL, SHAPE_MESH, SHAPE_PLANE, SHAPE_SPHERE, SHAPE_TAPERED_CAPSULE,
                      SHAPE_TAPERED_CYLINDER, VERTEX, VOLUME, BodyParams, Collider, tapered_capsule_device_params)

# Build flags (same bit meaning on the product and oracle builders)
CREATE_CONSTRAINTS = 1
CALC_EDGE_LENGTHS = 2
CALC_VOLUMES = 4
CALC_LRA_LENGTHS = 8
OPTIMIZE = 16
CALC_BEND_CONSTANTS = 32


@dataclass
class MeshSpec:
    vertices: np.ndarray                       # VERTEX records
    faces: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.uint32))
    edges: np.ndarray = field(default_factory=lambda: np.zeros(0, EDGE))
    volumes: np.ndarray = field(default_factory=lambda: np.zeros(0, VOLUME))
    lras: np.ndarray = field(default_factory=lambda: np.zeros(0, LRA))
    attrs: np.ndarray = field(default_factory=lambda: np.zeros((0, 5), np.float32))
    bend_type: int = BEND_DIHEDRAL
    angle_tolerance: float = float(np.float32(np.deg2rad(8.0)))
    lra_multiplier: float = 1.0
    flags: int = CREATE_CONSTRAINTS | OPTIMIZE
    skinned: np.ndarray | None = None           # SKINNED records (normals are calculated)
    inv_binds: np.ndarray | None = None         # INV_BIND records
    # Cosserat rods: (stretch-shear vertex pairs (n, 2), their compliances (n,), bend-twist
    # rod pairs (m, 2), their compliances (m,)); CalculateRodProperties fills in the rest
    rods: tuple | None = None


@dataclass
class BodySpec:
    mesh: MeshSpec
    position: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    params: BodyParams = field(default_factory=BodyParams)
    # optional per-constraint compliance override applied after the build
    # (callable(settings_arrays) -> None)
    post: object = None


@dataclass
class Scene:
    name: str
    bodies: list
    colliders: list
    gravity: tuple = (0.0, 0.0, -9.81)
    dt: float = 1.0 / 60.0
    collision_steps: int = 1
    frames: int = 1


def attrs(compliance, shear, bend, lra_type=LRA_NONE, lra_mult=1.0, n=1):
    a = np.zeros((n, 5), np.float32)
    a[:, 0], a[:, 1], a[:, 2], a[:, 3], a[:, 4] = compliance, shear, bend, lra_type, lra_mult
    return a


# ------------------------------------------------------------------ meshes
def cloth(n, spacing, inv_mass=100.0, pinned=(), jitter=0.0, seed=0, nz=None):
    """n x nz vertex cloth in the x-y plane, centred on the origin.

    Vertex (x, y) has index x + y * n.  Each quad is split along its (x,y)-(x+1,y+1)
    diagonal into two counter-clockwise (seen from +z) triangles.
    """
    nz = n if nz is None else nz
    xs = (np.arange(n, dtype=np.float64) - 0.5 * (n - 1)) * spacing
    ys = (np.arange(nz, dtype=np.float64) - 0.5 * (nz - 1)) * spacing
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel(), np.zeros(n * nz)], axis=1)
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    v = np.zeros(n * nz, VERTEX)
    v["pos"] = pos.astype(np.float32)
    v["inv_mass"] = np.float32(inv_mass)
    for p in pinned:
        v["inv_mass"][p] = 0.0
    idx = np.arange(n * nz, dtype=np.uint32).reshape(nz, n)
    a, b = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel()
    c, d = idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    faces = np.empty((2 * a.size, 3), np.uint32)
    faces[0::2] = np.stack([a, b, d], 1)
    faces[1::2] = np.stack([a, d, c], 1)
    return v, faces


_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                     [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], dtype=np.int64)
# Five-tetrahedron split of a cell; the two mirror variants alternate with the
# parity of x+y+z so the face diagonals of neighbouring cells agree.
_TETS_EVEN = np.array([[1, 2, 4, 7], [0, 1, 2, 4], [3, 1, 2, 7], [5, 1, 4, 7], [6, 2, 4, 7]])
_TETS_ODD = np.array([[0, 3, 5, 6], [1, 0, 3, 5], [2, 0, 3, 6], [4, 0, 5, 6], [7, 3, 5, 6]])


def tet_cube(cells, spacing, inv_mass=1.0, compliance=1e-6, seed=None, jitter=0.0):
    """(cells+1)^3-vertex block of 5 tets per cell, centred on the origin.

    Returns vertex records, edge records (unique tet edges sorted by vertex
    pair, rest length to be computed) and volume records (rest volume to be
    computed).
    """
    g = cells + 1
    r = (np.arange(g, dtype=np.float64) - 0.5 * cells) * spacing
    Z, Y, X = np.meshgrid(r, r, r, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    v = np.zeros(g ** 3, VERTEX)
    v["pos"] = pos.astype(np.float32)
    v["inv_mass"] = np.float32(inv_mass)

    cz, cy, cx = np.meshgrid(np.arange(cells), np.arange(cells), np.arange(cells), indexing="ij")
    cx, cy, cz = cx.ravel(), cy.ravel(), cz.ravel()
    corner_idx = ((cx[:, None] + _CORNERS[None, :, 0]) + (cy[:, None] + _CORNERS[None, :, 1]) * g
                  + (cz[:, None] + _CORNERS[None, :, 2]) * g * g)          # [cells^3, 8]
    odd = ((cx + cy + cz) & 1).astype(bool)
    tets = np.where(odd[:, None, None], corner_idx[:, _TETS_ODD], corner_idx[:, _TETS_EVEN])
    tets = tets.reshape(-1, 4).astype(np.uint32)

    vol = np.zeros(len(tets), VOLUME)
    vol["v"] = tets
    vol["compliance"] = np.float32(compliance)

    pairs = np.concatenate([tets[:, [i, j]] for i in range(4) for j in range(i + 1, 4)])
    pairs = np.sort(pairs, axis=1)
    pairs = np.unique(pairs, axis=0)
    e = np.zeros(len(pairs), EDGE)
    e["v"] = pairs
    e["compliance"] = np.float32(compliance)
    return v, e, vol


def icosphere(subdivisions, radius=1.0, inv_mass=1.0):
    """Icosahedron subdivided `subdivisions` times and projected to a sphere
    (outward counter-clockwise faces)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in verts]
    for _ in range(subdivisions):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.zeros(len(verts), VERTEX)
    v["pos"] = (np.array(verts) * radius).astype(np.float32)
    v["inv_mass"] = np.float32(inv_mass)
    return v, np.array(faces, np.uint32)


# ------------------------------------------------------------------ colliders
def _rot_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def spec_radius(spec):
    """Radius of a ball around the shape's origin that holds a ShapeSpec (conservative)."""
    if spec.kind == "basic":
        t, p = spec.type, tuple(float(x) for x in spec.params)
        if t == SHAPE_SPHERE:
            return p[0]
        if t == SHAPE_BOX:
            return float(np.linalg.norm(p[:3]))
        if t == SHAPE_CAPSULE:
            return p[0] + p[1]
        if t == SHAPE_CYLINDER:
            return float(np.hypot(p[0], p[1]))
        if t == SHAPE_TAPERED_CAPSULE:
            return p[0] + 2.0 * max(p[1], p[2])
        if t == SHAPE_TAPERED_CYLINDER:
            return float(np.hypot(p[0], max(p[1], p[2])))
        raise ValueError(f"shape type {t} in a compound")
    if spec.kind in ("hull", "triangle"):
        return float(np.linalg.norm(np.asarray(spec.params, np.float64).reshape(-1, 3), axis=1).max())
    if spec.kind == "scaled":
        return float(np.abs(spec.vec).max()) * spec_radius(spec.inner)
    if spec.kind == "rotated_translated":
        return float(np.linalg.norm(spec.vec)) + spec_radius(spec.inner)
    if spec.kind == "offset_com":
        return spec_radius(spec.inner)
    return max(float(np.linalg.norm(pos)) + spec_radius(child) for child, pos, _ in spec.children)


def collider_world_bounds(c: Collider):
    """World-space AABB of a static collider body (conservative for planes, scaled shapes,
    tapered cylinders and compounds)."""
    p = np.array(c.pos, np.float64)
    R = _rot_matrix(c.rot)
    scale = np.abs(np.array(c.scale, np.float64))
    if c.type == SHAPE_COMPOUND:
        r = spec_radius(c.shape)
        return tuple(np.concatenate([p - r, p + r]).astype(np.float32))
    if c.type == SHAPE_MESH and c.height_field is not None:
        samples, _, _, off, sc = c.height_field
        n = np.asarray(samples).shape[0]
        h = np.asarray(samples, np.float64)
        h = h[h < 1e30]
        pts = np.array([[0.0, h.min(), 0.0], [n - 1.0, h.max(), n - 1.0]]) * np.array(sc, np.float64) + np.array(off, np.float64)
        corners = np.array([[x, y, z] for x in pts[:, 0] for y in pts[:, 1] for z in pts[:, 2]])
        w = (corners * np.array(c.scale, np.float64)) @ R.T + p
        return tuple(np.concatenate([w.min(axis=0), w.max(axis=0)]).astype(np.float32))
    if c.type == SHAPE_MESH:
        pts = np.asarray(c.triangle if c.triangle is not None else c.mesh[0], np.float64).reshape(-1, 3)
        w = (pts * np.array(c.scale, np.float64)) @ R.T + p
        return tuple(np.concatenate([w.min(axis=0), w.max(axis=0)]).astype(np.float32))
    if c.type == SHAPE_TAPERED_CYLINDER:
        hh, r = c.params[0], max(c.params[1], c.params[2])
        he = np.abs(R) @ (scale * np.array([r, 2.0 * hh, r]))
        return tuple(np.concatenate([p - he, p + he]).astype(np.float32))
    if (scale != 1.0).any() and c.type != SHAPE_PLANE:
        # the unscaled bounds' extent around the body position, grown by the scale
        unscaled = Collider(c.type, c.params, pos=(0.0, 0.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0))
        b = np.array(collider_world_bounds(unscaled), np.float64)
        he = np.abs(R) @ (scale.max() * np.maximum(np.abs(b[:3]), np.abs(b[3:])))
        return tuple(np.concatenate([p - he, p + he]).astype(np.float32))
    if c.type == SHAPE_SPHERE:
        he = np.full(3, c.params[0])
        return tuple(np.concatenate([p - he, p + he]).astype(np.float32))
    if c.type == SHAPE_BOX:
        he = np.abs(R) @ np.array(c.params[:3], np.float64)
    elif c.type == SHAPE_CAPSULE:
        hh, r = c.params[0], c.params[1]
        he = np.abs(R) @ np.array([r, hh + r, r]) + 0.0
    elif c.type == SHAPE_CYLINDER:
        hh, r = c.params[0], c.params[1]
        he = np.abs(R) @ np.array([r, hh, r])
    elif c.type == SHAPE_TAPERED_CAPSULE:
        # conservative: a box around both end spheres, centred on the centre of mass
        d = tapered_capsule_device_params(*c.params[:3])
        r = max(float(d[2]), float(d[3]))
        ext = max(abs(float(d[0])) + float(d[2]), abs(float(d[1])) + float(d[3]))
        he = np.abs(R) @ np.array([r, ext, r])
        p = np.array(c.com_position(), np.float64)
    elif c.type == SHAPE_CONVEX_HULL:
        w = c.hull_points().astype(np.float64) @ R.T + p
        return tuple(np.concatenate([w.min(axis=0), w.max(axis=0)]).astype(np.float32))
    else:
        return plane_world_bounds(c)
    return tuple(np.concatenate([p - he, p + he]).astype(np.float32))


def plane_world_bounds(c: Collider):
    """Body::GetWorldSpaceBounds of a plane collider, in float32 as the reference computes it:
    PlaneShape::CalculateLocalBounds (the four corners of the half-extent square around the
    plane point and the same corners half extent behind it; PlaneShape.cpp:45-85), then
    AABox::Transformed by the body transform (AABox.h:193-213)."""
    f = np.float32
    n = np.array(c.params[:3], f)
    n = n / f(np.sqrt(f(f(n[0] * n[0] + n[1] * n[1]) + n[2] * n[2])))
    const, half = f(c.params[3]), f(c.params[4])

    def cross(a, b):
        return np.array([f(a[1] * b[2]) - f(a[2] * b[1]), f(a[2] * b[0]) - f(a[0] * b[2]),
                         f(a[0] * b[1]) - f(a[1] * b[0])], f)

    def length(a):
        return f(np.sqrt(f(f(a[0] * a[0] + a[1] * a[1]) + a[2] * a[2])))

    # sPlaneGetOrthogonalBasis: NormalizedOr(AxisX) of n x Y, then the two crosses
    p1 = cross(n, np.array([0, 1, 0], f))
    l2 = f(f(p1[0] * p1[0] + p1[1] * p1[1]) + p1[2] * p1[2])
    p1 = np.array([1, 0, 0], f) if l2 <= np.finfo(np.float32).tiny else p1 / f(np.sqrt(l2))
    p2 = cross(p1, n)
    p2 = p2 / length(p2)
    p1 = cross(n, p2)
    p1, p2 = p1 * half, p2 * half
    point = -n * const
    corners = [point + p1 + p2, point + p1 - p2, point - p1 - p2, point - p1 + p2]
    pts = np.array(corners + [v - half * n for v in corners], f)
    lmin, lmax = pts.min(axis=0), pts.max(axis=0)
    R = _rot_matrix(c.rot).astype(f)
    t = np.array(c.pos, f)
    wmin, wmax = t.copy(), t.copy()
    for k in range(3):
        a, b = R[:, k] * lmin[k], R[:, k] * lmax[k]
        wmin = wmin + np.minimum(a, b)
        wmax = wmax + np.maximum(a, b)
    return tuple(np.concatenate([wmin, wmax]).astype(np.float32))


# ------------------------------------------------------------------ configs
def config1_cloth_on_box(frames=300):
    """32x32 cloth (spacing 0.05 m, inv-mass 100, two top corners pinned), structural +
    shear edges (compliance 1e-5) and dihedral bends (1e-3), jitter U(-1e-4, 1e-4) seed 1,
    gravity -9.81 z, dt 1/60, 1 collision step x 5 iterations, over a static
    1 x 1 x 0.2 m box whose top is 0.5 m below the cloth; sleeping off."""
    n = 32
    v, f = cloth(n, 0.05, 100.0, pinned=((n - 1) * n, n * n - 1), jitter=1e-4, seed=1)
    mesh = MeshSpec(vertices=v, faces=f, attrs=attrs(1e-5, 1e-5, 1e-3), bend_type=BEND_DIHEDRAL)
    body = BodySpec(mesh=mesh, params=BodyParams(num_iterations=5, allow_sleeping=False))
    box = Collider(SHAPE_BOX, (0.5, 0.5, 0.1), pos=(0.0, 0.0, -0.6))
    box.bounds = collider_world_bounds(box)
    return Scene("cloth32_on_box", [body], [box], gravity=(0.0, 0.0, -9.81), dt=1.0 / 60.0,
                 collision_steps=1, frames=frames)


def _log_uniform(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=size))


def config2_cloth_and_cube(cloth_n=512, cube_cells=48, num_spheres=64, seed=2, frames=1):
    """512^2 cloth (spacing 0.02 m, per-constraint compliance log-U[1e-6, 1e-3]) plus a
    48^3-cell x 5-tet cube (edge + volume compliance 1e-6), ground plane and 64 static
    spheres r ~ U(0.1, 0.5); 4 collision steps x 8 iterations."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v, f = cloth(cloth_n, 0.02, 100.0, jitter=1e-4, seed=seed)
    mesh = MeshSpec(vertices=v, faces=f, attrs=attrs(1e-4, 1e-4, 1e-4), bend_type=BEND_DIHEDRAL)

    def cloth_compliance(s):
        r = np.random.Generator(np.random.PCG64(seed + 100))
        s.edges["compliance"] = _log_uniform(r, 1e-6, 1e-3, len(s.edges)).astype(np.float32)
        s.bends["compliance"] = _log_uniform(r, 1e-6, 1e-3, len(s.bends)).astype(np.float32)

    half = 0.5 * 0.02 * (cloth_n - 1)
    cloth_body = BodySpec(mesh=mesh, position=(0.0, 0.0, 3.0),
                          params=BodyParams(num_iterations=8, allow_sleeping=False), post=cloth_compliance)
    tv, te, tt = tet_cube(cube_cells, 0.05, inv_mass=1.0, compliance=1e-6)
    cube_mesh = MeshSpec(vertices=tv, edges=te, volumes=tt, flags=CALC_EDGE_LENGTHS | CALC_VOLUMES | OPTIMIZE)
    cube_body = BodySpec(mesh=cube_mesh, position=(half + 3.0, 0.0, 2.0),
                         params=BodyParams(num_iterations=8, allow_sleeping=False))
    colliders = [Collider(SHAPE_PLANE, (0.0, 0.0, 1.0, 0.0, 1000.0))]
    for _ in range(num_spheres):
        r = float(rng.uniform(0.1, 0.5))
        p = (float(rng.uniform(-half, half)), float(rng.uniform(-half, half)), float(rng.uniform(0.5, 2.0)))
        colliders.append(Collider(SHAPE_SPHERE, (r,), pos=p))
    for c in colliders:
        c.bounds = collider_world_bounds(c)
    return Scene("cloth512_cube48", [cloth_body, cube_body], colliders, gravity=(0.0, 0.0, -9.81),
                 dt=1.0 / 60.0, collision_steps=4, frames=frames)


def config5_block(cells=64, spacing=0.05, num_boxes=512, seed=5, frames=1, region=10.0, touching=True):
    """Tetrahedralised cells^3 block (5 tets per cell, spacing 0.05 m, inv-mass 1) with edge,
    volume and long-range-attachment constraints: the top face's first cells x cells vertices
    are pinned (4096 for 64 cells) and every free vertex gets a Euclidean LRA to its nearest
    pinned vertex (max distance = rest distance).  Compliances log-uniform [1e-7, 1e-4] per
    constraint; 512 static boxes (half extents U(0.2, 2) m, random rotation) in a 20 x 20 x 10 m
    region around the block (`region` = 10: its half width); 8 collision steps x 10 iterations.
    `touching`: a box whose world bounds overlap the block's is moved out along its shallowest
    axis until the two bounds overlap by 1 cm, so the boxes touch the block instead of starting
    deep inside it (a block embedded in boxes blows up within the first frame, in the
    reference as here)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v, e, t = tet_cube(cells, spacing, inv_mass=1.0, compliance=1e-6)
    g = cells + 1
    top = cells * g * g
    pinned = np.array([top + ix + iy * g for iy in range(cells) for ix in range(cells)], np.int64)
    v["inv_mass"][pinned] = 0.0
    idx = np.arange(g ** 3, dtype=np.int64)
    free = idx[v["inv_mass"] > 0.0]
    ix, iy = free % g, (free // g) % g
    anchor = top + np.minimum(ix, cells - 1) + np.minimum(iy, cells - 1) * g     # nearest pinned vertex
    lra = np.zeros(len(free), LRA)
    lra["v"][:, 0] = anchor
    lra["v"][:, 1] = free
    mesh = MeshSpec(vertices=v, edges=e, volumes=t, lras=lra, lra_multiplier=1.0,
                    flags=CALC_EDGE_LENGTHS | CALC_VOLUMES | CALC_LRA_LENGTHS | OPTIMIZE)

    def compliance(s):
        r = np.random.Generator(np.random.PCG64(seed + 100))
        s.edges["compliance"] = _log_uniform(r, 1e-7, 1e-4, len(s.edges)).astype(np.float32)
        s.volumes["compliance"] = _log_uniform(r, 1e-7, 1e-4, len(s.volumes)).astype(np.float32)

    body = BodySpec(mesh=mesh, position=(0.0, 0.0, 2.0),
                    params=BodyParams(num_iterations=10, allow_sleeping=False), post=compliance)
    colliders = []
    for _ in range(num_boxes):
        he = tuple(float(x) for x in rng.uniform(0.2, 2.0, 3))
        p = (float(rng.uniform(-region, region)), float(rng.uniform(-region, region)),
             float(rng.uniform(-0.5 * region, 0.5 * region)))
        colliders.append(Collider(SHAPE_BOX, he, pos=p, rot=random_quaternion(rng)))
    for c in colliders:
        c.bounds = collider_world_bounds(c)
    if touching:
        lo = v["pos"].min(axis=0).astype(np.float64) + np.array(body.position)
        hi = v["pos"].max(axis=0).astype(np.float64) + np.array(body.position)
        for c in colliders:
            b = np.array(c.bounds, np.float64)
            if np.any(lo > b[3:]) or np.any(hi < b[:3]):
                continue
            # the shift that clears the block's bounds, per axis and side; the shallowest wins
            shift, axis, side = min(((b[3 + k] - lo[k]) if sd < 0 else (hi[k] - b[k]), k, sd)
                                    for k in range(3) for sd in (-1, 1))
            p = list(c.pos)
            p[axis] += side * (shift - 0.01)
            c.pos = tuple(float(x) for x in p)
            c.bounds = collider_world_bounds(c)
    return Scene(f"block{cells}_lra_boxes{num_boxes}", [body], colliders, gravity=(0.0, 0.0, -9.81),
                 dt=1.0 / 60.0, collision_steps=8, frames=frames)


def random_quaternion(rng):
    """Uniform random rotation (Shoemake), float32 components, normalised in float64."""
    u1, u2, u3 = rng.uniform(size=3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    return tuple(float(x) for x in np.float32(q / np.linalg.norm(q)))


def config3_batched(num_bodies=2048, seed=3, frames=600):
    """Batched uneven bodies: 60 % cloths (side log-uniform 8-128 vertices, spacing 0.05 m,
    edges + shear + dihedral bends), 30 % 5-tet cubes (4^3-16^3 cells, spacing 0.1 m, edges +
    volumes), 10 % icospheres (subdivision 2-4, radius 1 m, pressure U(500, 2000)); one
    compliance per body, log-uniform in [1e-6, 1e-3]; random position in a 100 m box above a
    ground plane and random rotation; dt 1/60, 2 collision steps x 5 iterations."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_cloth = int(round(0.6 * num_bodies))
    n_cube = int(round(0.3 * num_bodies))
    n_ico = num_bodies - n_cloth - n_cube
    kinds = ["cloth"] * n_cloth + ["cube"] * n_cube + ["ico"] * n_ico
    rng.shuffle(kinds)
    bodies = []
    for kind in kinds:
        compliance = float(np.float32(np.exp(rng.uniform(np.log(1e-6), np.log(1e-3)))))
        pos = (float(rng.uniform(-50, 50)), float(rng.uniform(-50, 50)), float(rng.uniform(5, 105)))
        rot = random_quaternion(rng)
        if kind == "cloth":
            n = int(round(np.exp(rng.uniform(np.log(8), np.log(128)))))
            key = ("cloth", n)
            params = BodyParams(num_iterations=5, allow_sleeping=False)
        elif kind == "cube":
            n = int(rng.integers(4, 17))
            key = ("cube", n)
            params = BodyParams(num_iterations=5, allow_sleeping=False)
        else:
            n = int(rng.integers(2, 5))
            key = ("ico", n)
            params = BodyParams(num_iterations=5, allow_sleeping=False, pressure=float(rng.uniform(500, 2000)))
        bodies.append(BodySpec(mesh=key, position=pos, rotation=rot, params=params,
                               post=_uniform_compliance(compliance)))
    plane = Collider(SHAPE_PLANE, (0.0, 0.0, 1.0, 0.0, 1000.0))
    plane.bounds = collider_world_bounds(plane)
    return Scene("batched2048", bodies, [plane], gravity=(0.0, 0.0, -9.81), dt=1.0 / 60.0,
                 collision_steps=2, frames=frames)


def batched_check_sample(scene, settings):
    """64 bodies of the batched scene covering every size: the 12 largest cloths (streamed,
    above the LDS limit), 24 cloths and 18 cubes evenly spread over their sorted sizes
    (single-body workgroups and bundles of both launch classes) and 10 icospheres of every
    subdivision (pressure)."""
    kinds = {"cloth": [], "cube": [], "ico": []}
    for i, b in enumerate(scene.bodies):
        kinds[b.mesh[0]].append(i)

    def spread(ids, n, skip=0):
        ids = sorted(ids, key=lambda i: (-len(settings[i].vertices), i))[skip:]
        return [ids[int(k * (len(ids) - 1) / max(n - 1, 1))] for k in range(n)]

    cloths = sorted(kinds["cloth"], key=lambda i: (-len(settings[i].vertices), i))
    pick = cloths[:12] + spread(kinds["cloth"], 24, skip=12) + spread(kinds["cube"], 18) + spread(kinds["ico"], 10)
    return sorted(set(pick))


def state_digest(r):
    """SHA-256 of a body's float32 positions, velocities and body position (bit patterns):
    the form in which reference trajectories of the batched scene are kept."""
    h = hashlib.sha256()
    for k in ("pos", "vel", "body_pos"):
        h.update(np.ascontiguousarray(r[k], np.float32).tobytes())
    return h.hexdigest()


def _uniform_compliance(c):
    def apply(s):
        s.edges["compliance"] = np.float32(c)
        s.bends["compliance"] = np.float32(c)
        s.volumes["compliance"] = np.float32(c)
    return apply


@dataclass
class RolloutCase:
    """Differentiation workload: `num_bodies` copies of one edge-only cloth, per-edge
    compliance and initial velocities as parameters, and a target produced by a forward
    run with the compliance scaled per edge."""

    name: str
    mesh: MeshSpec
    num_bodies: int
    frames: int
    iterations: int
    dt: float
    gravity: tuple
    compliance: np.ndarray          # (num_bodies * E,) float32: the parameter values
    velocity: np.ndarray            # (num_bodies * V, 3) float32
    target_scale: np.ndarray        # (num_bodies * E,) float32: compliance multiplier of the target run
    checkpoint_interval: int = 10
    pressure: float = 0.0           # mPressure of every body (ApplyPressure on its faces)


def config4_rollout(num_bodies=64, n=64, frames=60, seed=4, compliance=1e-3, lra=False, bends=False):
    """64 cloths of 64 x 64 vertices (spacing 0.05 m, inv-mass 100, two top corners
    pinned), structural + shear edges only (compliance 1e-3), gravity -9.81 z, dt 1/60,
    one collision step x 5 iterations, 60 frames, no colliders.  Initial velocities
    U(-0.1, 0.1) m/s and target compliance multipliers U(0.5, 2.0) per edge, from `seed`.
    lra: also Euclidean LRA constraints from the pinned corners (max distance 1.0 x the rest
    distance), which act once the sheet stretches.  bends: also dihedral bends (compliance
    1e-3) on a sheet jittered by up to 1 cm, so that no hinge starts flat."""
    v, f = cloth(n, 0.05, 100.0, pinned=((n - 1) * n, n * n - 1), jitter=0.01 if bends else 0.0, seed=seed)
    a = attrs(compliance, compliance, 1e-3 if bends else 0.0, lra_type=LRA_EUCLIDEAN if lra else LRA_NONE, lra_mult=1.0)
    mesh = MeshSpec(vertices=v, faces=f, attrs=a, bend_type=BEND_DIHEDRAL if bends else BEND_NONE)
    rng = np.random.Generator(np.random.PCG64(seed))
    num_edges = 2 * n * (n - 1) + 2 * (n - 1) * (n - 1)     # structural + shear
    if bends:
        # with dihedral bends CreateConstraints makes another edge set: count it
        from .xpbd import build_settings
        num_edges = len(build_settings(mesh).edges)
    comp = np.full(num_bodies * num_edges, compliance, np.float32)
    vel = rng.uniform(-0.1, 0.1, size=(num_bodies * n * n, 3)).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, size=num_bodies * num_edges).astype(np.float32)
    return RolloutCase(f"cloth{n}x{n}x{num_bodies}_rollout", mesh, num_bodies, frames, 5, 1.0 / 60.0,
                       (0.0, 0.0, -9.81), comp, vel, scale)


def rollout_tet_block(num_bodies=2, cells=3, frames=8, seed=6, compliance=1e-4, volume_compliance=1e-5):
    """Differentiation workload with volume constraints: `num_bodies` copies of a (cells+1)^3
    5-tet block (spacing 0.1 m) hanging from its top layer (inverse mass 0), edges of
    compliance `compliance` (the parameters) and volumes of compliance `volume_compliance`,
    gravity -9.81 z, dt 1/60, 5 iterations; initial velocities U(-0.1, 0.1) m/s and target
    compliance multipliers U(0.5, 2.0) per edge, from `seed`."""
    v, e, t = tet_cube(cells, 0.1, inv_mass=1.0, compliance=compliance)
    t["compliance"] = volume_compliance
    top = v["pos"][:, 2] > v["pos"][:, 2].max() - 1e-6
    v["inv_mass"][top] = 0.0
    mesh = MeshSpec(vertices=v, edges=e, volumes=t, flags=CALC_EDGE_LENGTHS | CALC_VOLUMES | OPTIMIZE)
    rng = np.random.Generator(np.random.PCG64(seed))
    comp = np.full(num_bodies * len(e), compliance, np.float32)
    vel = rng.uniform(-0.1, 0.1, size=(num_bodies * len(v), 3)).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, size=num_bodies * len(e)).astype(np.float32)
    return RolloutCase(f"tet{cells}x{num_bodies}_rollout", mesh, num_bodies, frames, 5, 1.0 / 60.0, (0.0, 0.0, -9.81),
                       comp, vel, scale)


def rollout_balloon(num_bodies=2, subdivisions=2, frames=8, seed=8, compliance=1e-4, pressure=1000.0):
    """Differentiation workload with pressure: `num_bodies` copies of an icosphere of
    radius 0.5 m (`subdivisions` subdivisions, inverse mass 1, its highest vertex pinned),
    edges of compliance `compliance` (the parameters), no bends, pressure `pressure` on its
    faces, gravity -9.81 z, dt 1/60, 5 iterations; initial velocities U(-0.1, 0.1) m/s and
    target compliance multipliers U(0.5, 2.0) per edge, from `seed`."""
    v, f = icosphere(subdivisions, radius=0.5, inv_mass=1.0)
    v["inv_mass"][int(np.argmax(v["pos"][:, 2]))] = 0.0
    mesh = MeshSpec(vertices=v, faces=f, attrs=attrs(compliance, compliance, 0.0), bend_type=BEND_NONE)
    from .xpbd import build_settings
    num_edges = len(build_settings(mesh).edges)
    rng = np.random.Generator(np.random.PCG64(seed))
    comp = np.full(num_bodies * num_edges, compliance, np.float32)
    vel = rng.uniform(-0.1, 0.1, size=(num_bodies * len(v), 3)).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, size=num_bodies * num_edges).astype(np.float32)
    return RolloutCase(f"balloon{subdivisions}x{num_bodies}_rollout", mesh, num_bodies, frames, 5, 1.0 / 60.0,
                       (0.0, 0.0, -9.81), comp, vel, scale, pressure=float(pressure))


def mesh_for_key(key):
    """MeshSpec of a shape key used by config3_batched (meshes shared between bodies)."""
    kind, n = key
    if kind == "cloth":
        v, f = cloth(n, 0.05, 100.0)
        return MeshSpec(vertices=v, faces=f, attrs=attrs(1e-4, 1e-4, 1e-4), bend_type=BEND_DIHEDRAL)
    if kind == "cube":
        v, e, t = tet_cube(n, 0.1, inv_mass=1.0, compliance=1e-4)
        return MeshSpec(vertices=v, edges=e, volumes=t, flags=CALC_EDGE_LENGTHS | CALC_VOLUMES | OPTIMIZE)
    v, f = icosphere(n, radius=1.0, inv_mass=1.0)
    return MeshSpec(vertices=v, faces=f, attrs=attrs(1e-4, 1e-4, 1e-4), bend_type=BEND_DIHEDRAL)


def resolve_meshes(scene, build):
    """Replaces shape keys by MeshSpecs and returns per-body settings built with `build`
    (one build per distinct shape; the per-body compliance is applied on a copy)."""
    cache = {}
    settings = []
    for b in scene.bodies:
        key = b.mesh if isinstance(b.mesh, tuple) else None
        if key is not None:
            if key not in cache:
                cache[key] = build(mesh_for_key(key))
            s = cache[key].copy()
        else:
            s = build(b.mesh)
        if b.post is not None:
            b.post(s)
        settings.append(s)
    return settings


def solves_per_frame(scene, settings):
    """Constraint solves per frame: sum over bodies of steps x iterations x constraints."""
    return sum(scene.collision_steps * b.params.num_iterations * s.num_constraints()
               for b, s in zip(scene.bodies, settings))
