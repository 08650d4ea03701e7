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

assert VERTEX.itemsize == 28 and FACE.itemsize == 16 and EDGE.itemsize == 16
assert BEND.itemsize == 24 and VOLUME.itemsize == 24 and LRA.itemsize == 12
assert GROUP.itemsize == 28

# EBendType / ELRAType (SoftBodySharedSettings.h:21-34)
BEND_NONE, BEND_DISTANCE, BEND_DIHEDRAL = 0, 1, 2
LRA_NONE, LRA_EUCLIDEAN, LRA_GEODESIC = 0, 1, 2

# Collider shape types of the C ABI
SHAPE_PLANE, SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE = 0, 1, 2, 3


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

    def counts(self):
        return dict(V=len(self.vertices), F=len(self.faces), E=len(self.edges), B=len(self.bends),
                    T=len(self.volumes), L=len(self.lras), G=len(self.groups))

    def num_constraints(self):
        return len(self.edges) + len(self.bends) + len(self.volumes) + len(self.lras)

    def copy(self):
        return SettingsArrays(*(np.array(a, copy=True) for a in (
            self.vertices, self.faces, self.edges, self.bends, self.volumes, self.lras, self.groups)))


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
class Collider:
    """A static collider body: shape, world transform, material, world bounds."""

    type: int
    params: tuple
    pos: tuple = (0.0, 0.0, 0.0)
    rot: tuple = (0.0, 0.0, 0.0, 1.0)
    friction: float = 0.2
    restitution: float = 0.0
    bounds: tuple | None = None  # world-space AABB (min xyz, max xyz)

    def param_array(self):
        p = np.zeros(5, dtype=np.float32)
        p[:len(self.params)] = self.params
        return p
