"""ctypes front end of the two CPU oracles.  TEST INFRASTRUCTURE ONLY.

* ``JoltReference`` drives the reference library compiled from the reference
  sources (oracle/_ref/libjolt_oracle_{nofma,fma}.so, built by oracle/Makefile):
  the reference's own CreateConstraints / Optimize / PhysicsSystem::Update.
* ``Restatement`` drives the own C++ restatement (oracle/libxpbd_oracle.so) in
  FP32 (bit-exact target) or FP64 (the error yardstick).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from joltphysics_amd.records import (BEND, EDGE, FACE, GROUP, INV_BIND, LRA, MOTION_DYNAMIC, PART_OP, ROD_BT, ROD_SS,
                                     SHAPE_COMPOUND, SHAPE_CONVEX_HULL, SHAPE_MESH, SKINNED, TREE_NODE,
                                     VERTEX, VOLUME, ColliderPart, ConvexHull, SettingsArrays, TriangleTree)

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_vp = C.c_void_p
_u32 = C.c_uint32
_f = C.c_float
_i = C.c_int


def _ptr(a):
    if a is None or (hasattr(a, "size") and a.size == 0):
        return None
    return a.ctypes.data_as(_vp)


def ref_library_path(variant="nofma"):
    return os.path.join(REF_DIR, f"libjolt_oracle_{variant}.so")


def restatement_path():
    return os.path.join(HERE, "libxpbd_oracle.so")


def have_reference(variant="nofma"):
    return os.path.exists(ref_library_path(variant))


def have_restatement():
    return os.path.exists(restatement_path())


def _faces4(faces):
    """Face records from either (F,3) index arrays or FACE records."""
    if faces.dtype == FACE:
        return np.ascontiguousarray(faces)
    f = np.zeros(len(faces), FACE)
    f["v"] = faces
    return f


class JoltReference:
    """The reference library (compiled from its sources)."""

    _libs = {}

    def __init__(self, variant="nofma"):
        if variant not in self._libs:
            lib = C.CDLL(ref_library_path(variant))
            lib.jo_settings_build.restype = _vp
            lib.jo_settings_build.argtypes = [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32,
                                              _vp, _u32, _i, _f, _f, _u32]
            lib.jo_settings_from_arrays.restype = _vp
            lib.jo_settings_from_arrays.argtypes = [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32,
                                                    _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32,
                                                    _vp, _u32, _vp, _u32]
            lib.jo_settings_add_skin.argtypes = [_vp, _vp, _u32, _vp, _u32]
            lib.jo_settings_add_rods.argtypes = [_vp, _vp, _vp, _u32, _vp, _vp, _u32]
            lib.jo_settings_rod_counts.argtypes = [_vp, _vp]
            lib.jo_settings_rod_read.argtypes = [_vp, _vp, _vp]
            lib.jo_settings_optimize.argtypes = [_vp]
            lib.jo_settings_skin_counts.argtypes = [_vp, _vp]
            lib.jo_settings_skin_read.argtypes = [_vp, _vp, _vp, _vp]
            lib.jo_settings_counts.argtypes = [_vp, _vp]
            lib.jo_settings_read.argtypes = [_vp] * 8
            lib.jo_settings_free.argtypes = [_vp]
            lib.jo_acos_approximate.restype = _f
            lib.jo_acos_approximate.argtypes = [_f]
            lib.jo_world_create.restype = _vp
            lib.jo_world_create.argtypes = [_u32, _i]
            lib.jo_world_free.argtypes = [_vp]
            lib.jo_world_set_gravity.argtypes = [_vp, _f, _f, _f]
            lib.jo_world_add_collider.restype = _i
            lib.jo_world_add_collider.argtypes = [_vp, _i, _vp, _vp, _vp, _f, _f]
            lib.jo_world_collider_bounds.restype = _i
            lib.jo_world_collider_bounds.argtypes = [_vp, _i, _vp]
            lib.jo_world_collider_device.restype = _i
            lib.jo_world_collider_device.argtypes = [_vp, _i, _vp, _vp]
            lib.jo_world_add_hull.restype = _i
            lib.jo_world_add_hull.argtypes = [_vp, _vp, _i, _vp, _vp, _f, _f]
            lib.jo_world_set_next_scale.restype = None
            lib.jo_world_set_next_scale.argtypes = [_vp, _vp]
            lib.jo_world_add_triangle.restype = _i
            lib.jo_world_add_triangle.argtypes = [_vp, _vp, _vp, _vp, _f, _f]
            lib.jo_world_add_mesh.restype = _i
            lib.jo_world_add_mesh.argtypes = [_vp, _vp, _i, _vp, _i, _i, _vp, _vp, _f, _f]
            lib.jo_world_collider_describe.restype = _i
            lib.jo_world_collider_describe.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
            lib.jo_world_add_height_field.restype = _i
            lib.jo_world_add_height_field.argtypes = [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _f, _f]
            lib.jo_world_hull_data.restype = _i
            for name, args in (("jo_shape_basic", [_vp, _i, _vp]), ("jo_shape_hull", [_vp, _vp, _i]),
                               ("jo_shape_scaled", [_vp, _i, _vp]), ("jo_shape_triangle", [_vp, _vp]), ("jo_shape_rotated_translated", [_vp, _i, _vp, _vp]),
                               ("jo_shape_offset_com", [_vp, _i, _vp]), ("jo_shape_compound", [_vp, _i, _i, _vp, _vp, _vp]),
                               ("jo_world_add_shape", [_vp, _i, _vp, _vp, _f, _f, _i, _vp, _vp]),
                               ("jo_world_collider_part", [_vp, _i, _i] + [_vp] * 14),
                               ("jo_collect_leaves", [_vp, _i, _vp, _vp, _vp, _vp, _i])):
                getattr(lib, name).restype = _i
                getattr(lib, name).argtypes = args
            lib.jo_world_add_dynamic.restype = _i
            lib.jo_world_add_dynamic.argtypes = [_vp, _i, _vp, _vp, _vp, _f, _f, _vp, _vp]
            lib.jo_world_collider_motion.restype = _i
            lib.jo_world_record_contacts.argtypes = [_vp]
            lib.jo_world_set_sensor.argtypes = [_vp, _i]
            lib.jo_world_sensor_contacts.restype = _i
            lib.jo_world_sensor_contacts.argtypes = [_vp, _i, _vp, _i]
            lib.jo_world_contacts.restype = _i
            lib.jo_body_skin_vertices.restype = _i
            lib.jo_body_skin_vertices.argtypes = [_vp, _i, _vp, _u32, _i, _i, _f]
            lib.jo_world_set_contact_settings.argtypes = [_vp, _i, _vp, _i]
            lib.jo_body_rod_states.restype = _i
            lib.jo_body_rod_states.argtypes = [_vp, _i, _vp, _vp, _u32]
            lib.jo_body_save_state.restype = C.c_long
            lib.jo_body_save_state.argtypes = [_vp, _i, _vp, C.c_long]
            lib.jo_body_restore_state.restype = _i
            lib.jo_body_restore_state.argtypes = [_vp, _i, _vp, C.c_long]
            lib.jo_world_contacts.argtypes = [_vp, _i, _vp, _vp]
            lib.jo_world_collider_motion.argtypes = [_vp, _i, _vp]
            lib.jo_world_hull_data.argtypes = [_vp, _i] + [_vp] * 6
            lib.jo_world_add_soft_body.restype = _i
            lib.jo_world_add_soft_body.argtypes = [_vp, _vp, _vp, _vp, _vp]
            lib.jo_world_optimize_broadphase.argtypes = [_vp]
            lib.jo_world_update.restype = C.c_double
            lib.jo_world_update.argtypes = [_vp, _f, _i, _i]
            lib.jo_body_custom_update.restype = _i
            lib.jo_body_custom_update.argtypes = [_vp, _i, _f]
            lib.jo_body_read.restype = _i
            lib.jo_body_read.argtypes = [_vp, _i] + [_vp] * 10
            lib.jo_body_write.restype = _i
            lib.jo_body_write.argtypes = [_vp, _i, _vp, _vp]
            lib.jo_hardware_concurrency.restype = _i
            self._libs[variant] = lib
        self.lib = self._libs[variant]
        self.variant = variant

    # ---------------------------------------------------------------- settings
    def build_settings(self, mesh) -> SettingsArrays:
        """Runs the reference's CreateConstraints / Calculate* / Optimize on a MeshSpec."""
        v = np.ascontiguousarray(mesh.vertices, VERTEX)
        f = np.ascontiguousarray(mesh.faces, np.uint32).reshape(-1, 3)
        e = np.ascontiguousarray(mesh.edges, EDGE)
        t = np.ascontiguousarray(mesh.volumes, VOLUME)
        lr = np.ascontiguousarray(mesh.lras, LRA)
        a = np.ascontiguousarray(mesh.attrs, np.float32).reshape(-1, 5)
        skinned = getattr(mesh, "skinned", None)
        rods = getattr(mesh, "rods", None)
        flags = int(mesh.flags)
        optimize = 16
        if skinned is not None or rods is not None:
            flags &= ~optimize          # skinned constraints and rods join before Optimize
        h = self.lib.jo_settings_build(_ptr(v), len(v), _ptr(f), len(f), _ptr(e), len(e), None, 0,
                                       _ptr(t), len(t), _ptr(lr), len(lr), _ptr(a), len(a), int(mesh.bend_type),
                                       float(mesh.angle_tolerance), float(mesh.lra_multiplier), flags)
        try:
            if skinned is not None:
                sk = np.ascontiguousarray(skinned, SKINNED)
                ib = np.ascontiguousarray(mesh.inv_binds, INV_BIND)
                self.lib.jo_settings_add_skin(h, _ptr(sk), len(sk), _ptr(ib), len(ib))
            if rods is not None:
                sv = np.ascontiguousarray(rods[0], np.uint32).reshape(-1, 2)
                sc = np.ascontiguousarray(rods[1], np.float32).reshape(-1)
                bv = np.ascontiguousarray(rods[2], np.uint32).reshape(-1, 2)
                bc = np.ascontiguousarray(rods[3], np.float32).reshape(-1)
                self.lib.jo_settings_add_rods(h, _ptr(sv), _ptr(sc), len(sv), _ptr(bv), _ptr(bc), len(bv))
            if (skinned is not None or rods is not None) and int(mesh.flags) & optimize:
                self.lib.jo_settings_optimize(h)
            return self._read_settings(h)
        finally:
            self.lib.jo_settings_free(h)

    def _read_settings(self, h) -> SettingsArrays:
        counts = np.zeros(7, np.uint32)
        self.lib.jo_settings_counts(h, _ptr(counts))
        nv, nf, ne, nb, nt, nl, ng = (int(c) for c in counts)
        s = SettingsArrays(np.zeros(nv, VERTEX), np.zeros(nf, FACE), np.zeros(ne, EDGE), np.zeros(nb, BEND),
                           np.zeros(nt, VOLUME), np.zeros(nl, LRA), np.zeros(ng, GROUP))
        self.lib.jo_settings_read(h, _ptr(s.vertices), _ptr(s.faces), _ptr(s.edges), _ptr(s.bends),
                                  _ptr(s.volumes), _ptr(s.lras), _ptr(s.groups))
        sc = np.zeros(3, np.uint32)
        self.lib.jo_settings_skin_counts(h, _ptr(sc))
        if sc[0] or sc[2]:
            s.skinned = np.zeros(int(sc[0]), SKINNED)
            s.skinned_normals = np.zeros(int(sc[1]), np.uint32)
            s.inv_binds = np.zeros(int(sc[2]), INV_BIND)
            self.lib.jo_settings_skin_read(h, _ptr(s.skinned), _ptr(s.skinned_normals), _ptr(s.inv_binds))
        rc = np.zeros(2, np.uint32)
        self.lib.jo_settings_rod_counts(h, _ptr(rc))
        if rc[0] or rc[1]:
            s.rod_stretch_shear = np.zeros(int(rc[0]), ROD_SS)
            s.rod_bend_twist = np.zeros(int(rc[1]), ROD_BT)
            self.lib.jo_settings_rod_read(h, _ptr(s.rod_stretch_shear), _ptr(s.rod_bend_twist))
        return s

    def settings_handle(self, s: SettingsArrays):
        f4 = _faces4(s.faces)
        h = self.lib.jo_settings_from_arrays(_ptr(s.vertices), len(s.vertices), _ptr(f4), len(f4),
                                             _ptr(s.edges), len(s.edges), _ptr(s.bends), len(s.bends),
                                             _ptr(s.volumes), len(s.volumes), _ptr(s.lras), len(s.lras),
                                             _ptr(s.groups), len(s.groups), _ptr(s.skinned), len(s.skinned),
                                             _ptr(s.skinned_normals), len(s.skinned_normals), _ptr(s.inv_binds),
                                             len(s.inv_binds), _ptr(s.rod_stretch_shear), len(s.rod_stretch_shear),
                                             _ptr(s.rod_bend_twist), len(s.rod_bend_twist))
        if not h:
            raise RuntimeError("reference rejected the settings arrays")
        return h

    def acos_approximate(self, x):
        return self.lib.jo_acos_approximate(float(x))

    def hardware_concurrency(self):
        return self.lib.jo_hardware_concurrency()


class JoltWorld:
    """PhysicsSystem with static colliders and soft bodies (reference)."""

    def __init__(self, ref: JoltReference, gravity, max_bodies=4096, num_threads=0):
        self.ref = ref
        self.lib = ref.lib
        self.h = self.lib.jo_world_create(max_bodies, num_threads)
        self.lib.jo_world_set_gravity(self.h, *map(float, gravity))
        self._settings = []
        self.sizes = []
        self.rods = []

    def close(self):
        if self.h:
            for s in self._settings:
                self.lib.jo_settings_free(s)
            self.lib.jo_world_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _build_shape(self, spec):
        """Builds a ShapeSpec with the reference's shape classes; returns its handle."""
        f3 = lambda v: np.array(v, np.float32)
        if spec.kind == "basic":
            p = np.zeros(8, np.float32)
            p[:len(spec.params)] = spec.params
            h = self.lib.jo_shape_basic(self.h, int(spec.type), _ptr(p))
        elif spec.kind == "hull":
            pts = np.ascontiguousarray(np.array(spec.params, np.float32).reshape(-1, 3))
            h = self.lib.jo_shape_hull(self.h, _ptr(pts), len(pts))
        elif spec.kind == "triangle":
            h = self.lib.jo_shape_triangle(self.h, _ptr(np.array(spec.params, np.float32)))
        elif spec.kind == "scaled":
            h = self.lib.jo_shape_scaled(self.h, self._build_shape(spec.inner), _ptr(f3(spec.vec)))
        elif spec.kind == "rotated_translated":
            h = self.lib.jo_shape_rotated_translated(self.h, self._build_shape(spec.inner), _ptr(f3(spec.vec)), _ptr(f3(spec.rot)))
        elif spec.kind == "offset_com":
            h = self.lib.jo_shape_offset_com(self.h, self._build_shape(spec.inner), _ptr(f3(spec.vec)))
        elif spec.kind in ("static_compound", "mutable_compound"):
            kids = np.array([self._build_shape(c[0]) for c in spec.children], np.int32)
            pos = np.ascontiguousarray(np.array([c[1] for c in spec.children], np.float32).reshape(-1, 3))
            rot = np.ascontiguousarray(np.array([c[2] for c in spec.children], np.float32).reshape(-1, 4))
            h = self.lib.jo_shape_compound(self.h, int(spec.kind == "mutable_compound"), len(kids), _ptr(kids), _ptr(pos), _ptr(rot))
        else:
            raise ValueError(f"unknown shape kind {spec.kind}")
        if h < 0:
            raise RuntimeError(f"building a {spec.kind} shape failed")
        return h

    def collect_leaves(self, i, soft_pos, soft_rot, box):
        """The reference's CollectTransformedShapes of collider i for a soft body at (soft_pos,
        soft_rot) with local bounds box (min xyz, max xyz): [(4x4 leaf transform, scale)]."""
        f = lambda v: np.array(v, np.float32)
        out = np.zeros((256, 19), np.float32)
        n = self.lib.jo_collect_leaves(self.h, i, _ptr(f(soft_pos)), _ptr(f(soft_rot)), _ptr(f(box)), _ptr(out), 256)
        if n < 0 or n > 256:
            raise RuntimeError("collect_leaves failed")
        return [(out[k, :16].reshape(4, 4).T.copy(), tuple(out[k, 16:])) for k in range(n)]

    def collider_parts(self, i):
        """The leaves of compound collider i as the device takes them (ColliderPart list)."""
        n = self.lib.jo_world_collider_part(self.h, i, -1, *([None] * 14))
        if n < 0:
            raise RuntimeError("collider_part failed")
        parts = []
        for k in range(n):
            sizes = np.zeros(6, np.int32)
            self.lib.jo_world_collider_part(self.h, i, k, _ptr(sizes), *([None] * 13))
            nops, nn, nt, npt, nf, ni = (int(x) for x in sizes)
            t = C.c_int()
            params, scale = np.zeros(8, np.float32), np.zeros(3, np.float32)
            ops = np.zeros(max(nops, 1), PART_OP)
            root, flags = C.c_uint32(), C.c_uint32()
            nodes = np.zeros(max(nn, 1), TREE_NODE)
            tris = np.zeros((max(nt, 1), 9), np.float32)
            pts, planes = np.zeros((max(npt, 1), 3), np.float32), np.zeros((max(nf, 1), 4), np.float32)
            first, count, index = np.zeros(max(nf, 1), np.uint32), np.zeros(max(nf, 1), np.uint32), np.zeros(max(ni, 1), np.uint32)
            if self.lib.jo_world_collider_part(self.h, i, k, _ptr(sizes), C.byref(t), _ptr(params), _ptr(scale), ops.ctypes.data,
                                               C.byref(root), C.byref(flags), nodes.ctypes.data, _ptr(tris), _ptr(pts),
                                               _ptr(planes), _ptr(first), _ptr(count), _ptr(index)) != n:
                raise RuntimeError("collider_part failed")
            hull = tree = None
            if t.value == SHAPE_CONVEX_HULL:
                hull = ConvexHull(pts[:npt], planes[:nf], first[:nf], count[:nf], index[:ni])
            if t.value == SHAPE_MESH:
                tree = TriangleTree(int(root.value), nodes[:nn], tris[:nt], int(flags.value))
            parts.append(ColliderPart(int(t.value), tuple(float(x) for x in params), tuple(float(x) for x in scale),
                                      ops[:nops].copy(), hull, tree))
        return parts

    def add_collider(self, c):
        pos = np.array(c.pos, np.float32)
        rot = np.array(c.rot, np.float32)
        if c.type == SHAPE_COMPOUND:
            h = self._build_shape(c.shape)
            dyn = c.motion_type == MOTION_DYNAMIC
            lin = np.array(c.linear_velocity, np.float32)
            ang = np.array(c.angular_velocity, np.float32)
            i = self.lib.jo_world_add_shape(self.h, h, _ptr(pos), _ptr(rot), float(c.friction), float(c.restitution), int(dyn),
                                            _ptr(lin), _ptr(ang))
            if i < 0:
                raise RuntimeError("add_collider failed")
            if c.is_sensor:
                self.lib.jo_world_set_sensor(self.h, i)
            self._num_colliders = i + 1
            return i
        if tuple(float(x) for x in c.scale) != (1.0, 1.0, 1.0):
            self.lib.jo_world_set_next_scale(self.h, _ptr(np.array(c.scale, np.float32)))
        if c.type == SHAPE_MESH and c.motion_type != MOTION_DYNAMIC:
            if c.triangle is not None:
                v = np.array(c.triangle, np.float32)
                i = self.lib.jo_world_add_triangle(self.h, _ptr(v), _ptr(pos), _ptr(rot), float(c.friction), float(c.restitution))
            elif c.height_field is not None:
                samples, block, bits, off, sc = c.height_field
                smp = np.ascontiguousarray(samples, np.float32)
                i = self.lib.jo_world_add_height_field(self.h, _ptr(smp), int(smp.shape[0]), int(block), int(bits),
                                                       _ptr(np.array(off, np.float32)), _ptr(np.array(sc, np.float32)),
                                                       _ptr(pos), _ptr(rot), float(c.friction), float(c.restitution))
            else:
                verts = np.ascontiguousarray(c.mesh[0], np.float32).reshape(-1, 3)
                idx = np.ascontiguousarray(c.mesh[1], np.uint32).reshape(-1, 3)
                per_leaf = int(c.mesh[2]) if len(c.mesh) > 2 else 0
                i = self.lib.jo_world_add_mesh(self.h, _ptr(verts), len(verts), _ptr(idx), len(idx), per_leaf, _ptr(pos), _ptr(rot),
                                               float(c.friction), float(c.restitution))
            if i < 0:
                raise RuntimeError("add_collider failed")
            if c.is_sensor:
                self.lib.jo_world_set_sensor(self.h, i)
            self._num_colliders = i + 1
            return i
        if c.motion_type == MOTION_DYNAMIC:
            p = c.param_array()
            lin = np.array(c.linear_velocity, np.float32)
            ang = np.array(c.angular_velocity, np.float32)
            i = self.lib.jo_world_add_dynamic(self.h, int(c.type), _ptr(p), _ptr(pos), _ptr(rot), float(c.friction),
                                              float(c.restitution), _ptr(lin), _ptr(ang))
        elif c.type == SHAPE_CONVEX_HULL:
            pts = np.ascontiguousarray(c.hull_points(), np.float32)
            i = self.lib.jo_world_add_hull(self.h, _ptr(pts), len(pts), _ptr(pos), _ptr(rot),
                                           float(c.friction), float(c.restitution))
        else:
            p = c.param_array()
            i = self.lib.jo_world_add_collider(self.h, int(c.type), _ptr(p), _ptr(pos), _ptr(rot),
                                               float(c.friction), float(c.restitution))
        if i < 0:
            raise RuntimeError("add_collider failed")
        if c.is_sensor:
            self.lib.jo_world_set_sensor(self.h, i)
        self._num_colliders = i + 1
        return i

    def collider_bounds(self, i):
        out = np.zeros(6, np.float32)
        self.lib.jo_world_collider_bounds(self.h, i, _ptr(out))
        return tuple(float(x) for x in out)

    def collider_tree(self, i):
        """The triangle tree of mesh / triangle collider i as the device takes it."""
        t, params, scale, com = C.c_int(), np.zeros(8, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32)
        sizes = np.zeros(2, np.int32)
        if self.lib.jo_world_collider_describe(self.h, i, C.byref(t), _ptr(params), _ptr(scale), _ptr(com), _ptr(sizes),
                                               None, None, None, None) != 0:
            raise RuntimeError("collider_describe failed")
        root, flags = C.c_uint32(), C.c_uint32()
        nodes = np.zeros(int(sizes[0]), TREE_NODE)
        tris = np.zeros((int(sizes[1]), 9), np.float32)
        self.lib.jo_world_collider_describe(self.h, i, C.byref(t), _ptr(params), _ptr(scale), _ptr(com), _ptr(sizes),
                                            C.byref(root), nodes.ctypes.data if len(nodes) else _ptr(np.zeros(28, np.uint32)),
                                            _ptr(tris) if len(tris) else _ptr(np.zeros(9, np.float32)), C.byref(flags))
        return TriangleTree(int(root.value), nodes, tris, int(flags.value))

    def collider_device(self, i):
        """(shape parameters as the shape holds them [8], centre-of-mass position [3])."""
        params = np.zeros(8, np.float32)
        com = np.zeros(3, np.float32)
        self.lib.jo_world_collider_device(self.h, i, _ptr(params), _ptr(com))
        return tuple(float(x) for x in params), tuple(float(x) for x in com)

    def skin_vertices(self, body, joints, hard_skin_all=False, enable=None, multiplier=1.0):
        """SoftBodyMotionProperties::SkinVertices with joint matrices (n, 16) column major;
        enable / multiplier set SetEnableSkinConstraints / SetSkinnedMaxDistanceMultiplier."""
        j = np.ascontiguousarray(joints, np.float32).reshape(-1, 16)
        e = -1 if enable is None else int(bool(enable))
        if self.lib.jo_body_skin_vertices(self.h, body, _ptr(j), len(j), int(bool(hard_skin_all)), e, float(multiplier)) != 0:
            raise RuntimeError("SkinVertices failed")

    def set_contact_settings(self, body, settings):
        """OnSoftBodyContactValidate answers for soft body `body`: per collider {scale1, scale2,
        inertia scale2, accept}; needs the recording listener (record_contacts)."""
        t = np.ascontiguousarray(settings, np.float32).reshape(-1, 4)
        self.lib.jo_world_set_contact_settings(self.h, body, _ptr(t), len(t))

    def rod_states(self, body):
        """(rotations (n, 4), angular velocities (n, 3)) of the rods of soft body `body`."""
        n = self.rods[body]
        rot = np.zeros((max(n, 1), 4), np.float32)
        ang = np.zeros((max(n, 1), 3), np.float32)
        if self.lib.jo_body_rod_states(self.h, body, _ptr(rot), _ptr(ang), n) != 0:
            raise RuntimeError("rod states failed")
        return rot[:n], ang[:n]

    def save_state(self, body):
        """SoftBodyMotionProperties::SaveState bytes of soft body `body`."""
        n = self.lib.jo_body_save_state(self.h, body, None, 0)
        buf = C.create_string_buffer(n)
        self.lib.jo_body_save_state(self.h, body, buf, n)
        return buf.raw[:n]

    def restore_state(self, body, data):
        if self.lib.jo_body_restore_state(self.h, body, data, len(data)) != 0:
            raise RuntimeError("RestoreState failed")

    def record_contacts(self):
        """Installs the contact listener that records each soft body's SoftBodyManifold of
        the last collision step (call after adding the colliders)."""
        self.lib.jo_world_record_contacts(self.h)

    def sensor_contacts(self, body):
        """Per collider: 1 where the manifold of the last step listed it as a sensor contact."""
        n = getattr(self, "_num_colliders", 0)
        hits = np.zeros(max(n, 1), np.int32)
        if self.lib.jo_world_sensor_contacts(self.h, body, _ptr(hits), n) != 0:
            raise RuntimeError("contact recording is not enabled")
        return hits[:n]

    def contacts(self, body):
        """(collider index per vertex, -1 without contact; local contact points (V, 3))."""
        nv = self.sizes[body]
        col = np.zeros(nv, np.int32)
        pts = np.zeros((nv, 3), np.float32)
        if self.lib.jo_world_contacts(self.h, body, _ptr(col), _ptr(pts)) != 0:
            raise RuntimeError("contact recording is not enabled")
        return col, pts

    def collider_motion(self, i):
        """(EMotionType, {inv_mass, inv_inertia_diagonal, inertia_rotation, linear_velocity,
        angular_velocity}) of collider i, velocities current and in world space."""
        out = np.zeros(16, np.float32)
        m = self.lib.jo_world_collider_motion(self.h, i, _ptr(out))
        return m, dict(inv_mass=float(out[0]), inv_inertia_diagonal=tuple(float(x) for x in out[1:4]),
                       inertia_rotation=tuple(float(x) for x in out[4:8]),
                       linear_velocity=tuple(float(x) for x in out[8:11]),
                       angular_velocity=tuple(float(x) for x in out[11:14]))

    def collider_hull(self, i):
        """The hull of convex hull collider i as the shape holds it (ConvexHull); the
        centre of mass in the shape's frame is left at zero (collider_device gives
        the world position)."""
        sizes = np.zeros(3, np.int32)
        if self.lib.jo_world_hull_data(self.h, i, _ptr(sizes), None, None, None, None, None) != 0:
            raise RuntimeError("collider is not a convex hull")
        np_, nf, ni = (int(x) for x in sizes)
        pts = np.zeros((np_, 3), np.float32)
        planes = np.zeros((nf, 4), np.float32)
        first = np.zeros(nf, np.uint32)
        count = np.zeros(nf, np.uint32)
        index = np.zeros(ni, np.uint32)
        self.lib.jo_world_hull_data(self.h, i, _ptr(sizes), _ptr(pts), _ptr(planes), _ptr(first), _ptr(count),
                                    _ptr(index))
        return ConvexHull(pts, planes, first, count, index)

    def add_soft_body(self, settings: SettingsArrays, position, rotation, params):
        sh = self.ref.settings_handle(settings)
        self._settings.append(sh)
        pos = np.array(position, np.float32)
        rot = np.array(rotation, np.float32)
        p = params.as_floats()
        i = self.lib.jo_world_add_soft_body(self.h, sh, _ptr(pos), _ptr(rot), _ptr(p))
        if i < 0:
            raise RuntimeError("add_soft_body failed")
        self.sizes.append(len(settings.vertices))
        self.rods.append(len(settings.rod_stretch_shear))
        return i

    def optimize_broadphase(self):
        self.lib.jo_world_optimize_broadphase(self.h)

    def update(self, dt, collision_steps=1, frames=1):
        t = self.lib.jo_world_update(self.h, float(dt), int(collision_steps), int(frames))
        if t < 0:
            raise RuntimeError("PhysicsSystem::Update reported an error")
        return t

    def custom_update(self, body, dt):
        if self.lib.jo_body_custom_update(self.h, body, float(dt)) != 0:
            raise RuntimeError("CustomUpdate failed")

    def read(self, body):
        n = self.sizes[body]
        out = dict(pos=np.zeros((n, 3), np.float32), prev=np.zeros((n, 3), np.float32),
                   vel=np.zeros((n, 3), np.float32), plane=np.zeros((n, 4), np.float32),
                   shape_idx=np.zeros(n, np.int32), penetration=np.zeros(n, np.float32),
                   contact=np.zeros(n, np.int32), inv_mass=np.zeros(n, np.float32),
                   body_pos=np.zeros(3, np.float32), body_rot=np.zeros(4, np.float32))
        self.lib.jo_body_read(self.h, body, *(_ptr(out[k]) for k in (
            "pos", "prev", "vel", "plane", "shape_idx", "penetration", "contact", "inv_mass",
            "body_pos", "body_rot")))
        return out

    def write(self, body, pos=None, vel=None):
        pos = None if pos is None else np.ascontiguousarray(pos, np.float32)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32)
        self.lib.jo_body_write(self.h, body, _ptr(pos), _ptr(vel))


class Restatement:
    """Own CPU restatement of the update, FP32 (precision=32) or FP64 (64)."""

    _lib = None

    def __init__(self, gravity, precision=32):
        if Restatement._lib is None:
            lib = C.CDLL(restatement_path())
            lib.xr_world_create.restype = _vp
            lib.xr_world_create.argtypes = [_i]
            lib.xr_world_free.argtypes = [_vp]
            lib.xr_world_set_gravity.argtypes = [_vp, _f, _f, _f]
            lib.xr_world_add_collider.restype = _i
            lib.xr_world_add_collider.argtypes = [_vp, _i, _vp, _vp, _vp, _f, _f, _vp]
            lib.xr_world_set_hull.restype = _i
            lib.xr_world_set_sensor.restype = _i
            lib.xr_world_set_sensor.argtypes = [_vp, _i]
            lib.xr_world_set_motion.restype = _i
            lib.xr_world_set_motion.argtypes = [_vp, _i, _i, _f, _vp, _vp, _vp, _vp]
            lib.xr_body_velocity_deltas.restype = _i
            lib.xr_body_velocity_deltas.argtypes = [_vp, _i, _i, _vp, _vp, _vp]
            lib.xr_world_set_hull.argtypes = [_vp, _i, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _u32]
            lib.xr_world_set_collider_scale.restype = _i
            lib.xr_world_set_collider_scale.argtypes = [_vp, _i, _vp]
            lib.xr_world_set_tree.restype = _i
            lib.xr_world_set_tree.argtypes = [_vp, _i, _u32, _vp, _vp, _u32, _vp, _u32, _u32]
            lib.xr_world_set_part_tree.restype = _i
            lib.xr_world_set_part_tree.argtypes = [_vp, _i, _i, _u32, _vp, _vp, _u32, _vp, _u32, _u32]
            lib.xr_world_set_part_hull.restype = _i
            lib.xr_world_set_part_hull.argtypes = [_vp, _i, _i, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _u32]
            lib.xr_collect_leaves.restype = _i
            lib.xr_collect_leaves.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _i]
            lib.xr_world_add_part.restype = _i
            lib.xr_world_add_part.argtypes = [_vp, _i, _i, _vp, _vp, _u32, _vp]
            lib.xr_world_add_soft_body.restype = _i
            lib.xr_world_add_soft_body.argtypes = [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32,
                                                   _vp, _u32, _vp, _u32, _vp, _vp, _vp]
            lib.xr_world_update.restype = C.c_double
            lib.xr_world_update.argtypes = [_vp, _f, _i, _i]
            lib.xr_body_read.restype = _i
            lib.xr_body_read.argtypes = [_vp, _i] + [_vp] * 8
            lib.xr_body_write.restype = _i
            lib.xr_body_write.argtypes = [_vp, _i, _vp, _vp]
            lib.xr_acos_approximate.restype = _f
            lib.xr_acos_approximate.argtypes = [_f]
            Restatement._lib = lib
        self.lib = Restatement._lib
        self.h = self.lib.xr_world_create(precision)
        self.lib.xr_world_set_gravity(self.h, *map(float, gravity))
        self.sizes = []

    def close(self):
        if self.h:
            self.lib.xr_world_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _set_hull(self, i, part, h):
        a = (np.ascontiguousarray(h.points, np.float32), np.ascontiguousarray(h.planes, np.float32),
             np.ascontiguousarray(h.face_first, np.uint32), np.ascontiguousarray(h.face_count, np.uint32),
             np.ascontiguousarray(h.vertex_index, np.uint32))
        if self.lib.xr_world_set_part_hull(self.h, i, part, _ptr(a[0]), len(a[0]), _ptr(a[1]), len(a[1]), _ptr(a[2]),
                                           _ptr(a[3]), _ptr(a[4]), len(a[4])) != 0:
            raise RuntimeError("invalid convex hull data")

    def _set_tree(self, i, part, tree):
        if tree is None:
            raise RuntimeError("mesh collider without tree data")
        nodes = np.ascontiguousarray(tree.nodes, TREE_NODE)
        bounds = np.ascontiguousarray(np.concatenate([nodes["min"].reshape(-1, 12), nodes["max"].reshape(-1, 12)], axis=1), np.float32)
        children = np.ascontiguousarray(nodes["child"], np.uint32)
        tris = np.ascontiguousarray(tree.triangles, np.float32)
        if self.lib.xr_world_set_part_tree(self.h, i, part, int(tree.root), _ptr(bounds) if len(nodes) else None,
                                           _ptr(children) if len(nodes) else None, len(nodes), _ptr(tris) if len(tris) else None,
                                           len(tris), int(tree.flags)) != 0:
            raise RuntimeError("invalid triangle tree")

    def add_collider(self, c, bounds=None):
        compound = c.type == SHAPE_COMPOUND
        p = np.zeros(8, np.float32) if compound else c.device_param_array()
        pos = np.array(c.com_position(), np.float32)
        rot = np.array(c.rot, np.float32)
        b = np.array(bounds if bounds is not None else c.bounds, np.float32)
        # a compound body holds no shape of its own: its leaves follow as parts
        i = self.lib.xr_world_add_collider(self.h, 1 if compound else int(c.type), _ptr(p), _ptr(pos), _ptr(rot),
                                           float(c.friction), float(c.restitution), _ptr(b))
        if i >= 0 and compound:
            if not c.parts:
                raise RuntimeError("compound collider without parts")
            for pt in c.parts:
                prm = np.zeros(8, np.float32)
                prm[:len(pt.params)] = pt.params
                ops = np.ascontiguousarray(pt.ops, PART_OP)
                k = self.lib.xr_world_add_part(self.h, i, int(pt.type), _ptr(prm), _ptr(np.array(pt.scale, np.float32)), len(ops),
                                               ops.ctypes.data if len(ops) else None)
                if k < 0:
                    raise RuntimeError("invalid compound part")
                if pt.type == SHAPE_CONVEX_HULL:
                    self._set_hull(i, k, pt.hull)
                if pt.type == SHAPE_MESH:
                    self._set_tree(i, k, pt.tree)
        if i >= 0 and not compound and c.type == SHAPE_CONVEX_HULL:
            self._set_hull(i, -1, c.hull)
        if i >= 0 and not compound and c.type == SHAPE_MESH:
            self._set_tree(i, -1, c.tree)
        if i >= 0 and not compound and tuple(float(x) for x in c.scale) != (1.0, 1.0, 1.0):
            self.lib.xr_world_set_collider_scale(self.h, i, _ptr(np.array(c.scale, np.float32)))
        if i >= 0 and c.is_sensor:
            self.lib.xr_world_set_sensor(self.h, i)
        if i >= 0 and c.motion_type != 0:
            f = lambda t: np.array(t, np.float32)
            if self.lib.xr_world_set_motion(self.h, i, int(c.motion_type), float(c.inv_mass),
                                            _ptr(f(c.inv_inertia_diagonal)), _ptr(f(c.inertia_rotation)),
                                            _ptr(f(c.linear_velocity)), _ptr(f(c.angular_velocity))) != 0:
                raise RuntimeError("invalid collider motion")
        self._num_colliders = max(getattr(self, "_num_colliders", 0), i + 1)
        return i

    def collect_leaves(self, i, soft_pos, soft_rot, box):
        """The restatement's leaf transforms (4x4 each) of collider i for a soft body at
        (soft_pos, soft_rot) with local bounds box (min xyz, max xyz)."""
        f = lambda v: np.array(v, np.float32)
        out = np.zeros((256, 16), np.float32)
        n = self.lib.xr_collect_leaves(self.h, i, _ptr(f(soft_pos)), _ptr(f(soft_rot)), _ptr(f(box)), _ptr(out), 256)
        if n < 0 or n > 256:
            raise RuntimeError("collect_leaves failed")
        return [out[k].reshape(4, 4).T.copy() for k in range(n)]

    def set_motion(self, i, c):
        """Updates collider i's motion (velocities) from Collider c between update calls."""
        f = lambda t: np.array(t, np.float32)
        if self.lib.xr_world_set_motion(self.h, i, int(c.motion_type), float(c.inv_mass),
                                        _ptr(f(c.inv_inertia_diagonal)), _ptr(f(c.inertia_rotation)),
                                        _ptr(f(c.linear_velocity)), _ptr(f(c.angular_velocity))) != 0:
            raise RuntimeError("invalid collider motion")

    def velocity_deltas(self, body):
        """(linear (n, 3), angular (n, 3), updated (n,)) of the last update call."""
        n = getattr(self, "_num_colliders", 0)
        lin = np.zeros((max(n, 1), 3), np.float64)
        ang = np.zeros((max(n, 1), 3), np.float64)
        upd = np.zeros(max(n, 1), np.int32)
        if self.lib.xr_body_velocity_deltas(self.h, int(body), n, _ptr(lin), _ptr(ang), _ptr(upd)) != 0:
            raise RuntimeError("velocity_deltas failed")
        return lin[:n], ang[:n], upd[:n]

    def add_soft_body(self, s: SettingsArrays, position, rotation, params):
        f4 = _faces4(s.faces)
        pos = np.array(position, np.float32)
        rot = np.array(rotation, np.float32)
        p = params.as_floats()
        i = self.lib.xr_world_add_soft_body(self.h, _ptr(s.vertices), len(s.vertices), _ptr(f4), len(f4),
                                            _ptr(s.edges), len(s.edges), _ptr(s.bends), len(s.bends),
                                            _ptr(s.volumes), len(s.volumes), _ptr(s.lras), len(s.lras),
                                            _ptr(s.groups), len(s.groups), _ptr(pos), _ptr(rot), _ptr(p))
        if i < 0:
            raise RuntimeError("add_soft_body failed")
        self.sizes.append(len(s.vertices))
        return i

    def update(self, dt, collision_steps=1, frames=1):
        return self.lib.xr_world_update(self.h, float(dt), int(collision_steps), int(frames))

    def read(self, body):
        n = self.sizes[body]
        out = dict(pos=np.zeros((n, 3)), prev=np.zeros((n, 3)), vel=np.zeros((n, 3)),
                   plane=np.zeros((n, 4)), shape_idx=np.zeros(n, np.int32), penetration=np.zeros(n),
                   contact=np.zeros(n, np.int32), body_pos=np.zeros(3))
        self.lib.xr_body_read(self.h, body, *(_ptr(out[k]) for k in (
            "pos", "prev", "vel", "plane", "shape_idx", "penetration", "contact", "body_pos")))
        return out

    def write(self, body, pos=None, vel=None):
        pos = None if pos is None else np.ascontiguousarray(pos, np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64)
        self.lib.xr_body_write(self.h, body, _ptr(pos), _ptr(vel))

    def acos_approximate(self, x):
        return self.lib.xr_acos_approximate(float(x))


class RolloutReference:
    """CPU restatement of the differentiable rollout of edge, bend, volume and LRA bodies, with pressure on their
    faces when `pressure` > 0 (oracle/xpbd_adjoint.cpp): FP32 or FP64 forward pass and the FP64 adjoint from a
    full tape."""

    _lib = None

    def __init__(self, settings: SettingsArrays, iterations=5, frames=60, dt=1.0 / 60.0, linear_damping=0.1,
                 max_linear_velocity=500.0, gravity=(0.0, 0.0, -9.81), gravity_factor=1.0, pressure=0.0):
        if RolloutReference._lib is None:
            lib = C.CDLL(restatement_path())
            common = [_u32, _vp, _vp, _u32, _vp, _vp, _u32, _u32, _f, _f, _f, _vp, _f]
            lib.xa_rollout_forward.restype = _i
            lib.xa_rollout_forward.argtypes = [_i] + common + [_vp, _vp, _vp]
            lib.xa_rollout_gradient.restype = _i
            lib.xa_rollout_gradient.argtypes = [_i] + common + [_vp] * 6
            common_lra = [_u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _u32, _vp,
                          _u32, _vp, _f, _u32, _u32, _f, _f, _f, _vp, _f]
            lib.xa_rollout_forward_lra.restype = _i
            lib.xa_rollout_forward_lra.argtypes = [_i] + common_lra + [_vp, _vp, _vp]
            lib.xa_rollout_gradient_lra.restype = _i
            lib.xa_rollout_gradient_lra.argtypes = [_i] + common_lra + [_vp] * 6
            RolloutReference._lib = lib
        self.lib = RolloutReference._lib
        self.nb = len(settings.bends)
        self.bv = np.ascontiguousarray(settings.bends["v"], np.uint32).reshape(-1) if self.nb else np.zeros(4, np.uint32)
        self.bend_comp = np.ascontiguousarray(settings.bends["compliance"], np.float32) if self.nb else np.zeros(1, np.float32)
        self.angle = np.ascontiguousarray(settings.bends["initial_angle"], np.float32) if self.nb else np.zeros(1, np.float32)
        self.nt = len(settings.volumes)
        self.tv = np.ascontiguousarray(settings.volumes["v"], np.uint32).reshape(-1) if self.nt else np.zeros(4, np.uint32)
        self.six_rest = np.ascontiguousarray(settings.volumes["six_rest_volume"], np.float32) if self.nt else np.zeros(1, np.float32)
        self.vol_comp = np.ascontiguousarray(settings.volumes["compliance"], np.float32) if self.nt else np.zeros(1, np.float32)
        self.nl = len(settings.lras)
        self.lv = np.ascontiguousarray(settings.lras["v"], np.uint32).reshape(-1) if self.nl else np.zeros(2, np.uint32)
        self.maxd = np.ascontiguousarray(settings.lras["max_distance"], np.float32) if self.nl else np.zeros(1, np.float32)
        g = settings.groups
        self.groups = np.ascontiguousarray(np.stack([g["volume_end"], g["bend_end"], g["edge_end"], g["lra_end"]], axis=1),
                                           np.uint32).reshape(-1)
        self.ng = len(g)
        self.nf = len(settings.faces) if pressure > 0.0 else 0
        self.fv = np.ascontiguousarray(settings.faces["v"], np.uint32).reshape(-1) if self.nf else np.zeros(3, np.uint32)
        self.pressure = float(pressure)
        v = settings.vertices
        self.nv, self.ne = len(v), len(settings.edges)
        self.pos = np.ascontiguousarray(v["pos"], np.float32).reshape(-1)
        self.inv_mass = np.ascontiguousarray(v["inv_mass"], np.float32)
        self.ev = np.ascontiguousarray(settings.edges["v"], np.uint32).reshape(-1)
        self.rest = np.ascontiguousarray(settings.edges["rest"], np.float32)
        self.gravity = np.asarray(gravity, np.float32)
        self.args = (int(iterations), int(frames), float(dt), float(linear_damping), float(max_linear_velocity))
        self.gravity_factor = float(gravity_factor)

    def _common(self):
        return [self.nv, _ptr(self.pos), _ptr(self.inv_mass), self.ne, _ptr(self.ev), _ptr(self.rest), self.nl, _ptr(self.lv),
                _ptr(self.maxd), self.nt, _ptr(self.tv), _ptr(self.six_rest), _ptr(self.vol_comp), self.nb, _ptr(self.bv),
                _ptr(self.bend_comp), _ptr(self.angle), self.ng, _ptr(self.groups), self.nf, _ptr(self.fv), self.pressure,
                *self.args, _ptr(self.gravity), self.gravity_factor]

    def forward(self, compliance, velocity, precision=64):
        c = np.ascontiguousarray(compliance, np.float64).reshape(-1)
        vel = np.ascontiguousarray(velocity, np.float64).reshape(-1)
        out = np.zeros(3 * self.nv, np.float64)
        rc = self.lib.xa_rollout_forward_lra(int(precision), *self._common(), _ptr(c), _ptr(vel), _ptr(out))
        if rc != 0:
            raise RuntimeError("xa_rollout_forward failed")
        return out.reshape(-1, 3)

    def gradient(self, compliance, velocity, grad_positions, forward_precision=64):
        """FP64 adjoint on the tape of the FP64 (64) or FP32 (32) forward pass."""
        c = np.ascontiguousarray(compliance, np.float64).reshape(-1)
        vel = np.ascontiguousarray(velocity, np.float64).reshape(-1)
        g = np.ascontiguousarray(grad_positions, np.float64).reshape(-1)
        gc = np.zeros(self.ne, np.float64)
        gv = np.zeros(3 * self.nv, np.float64)
        out = np.zeros(3 * self.nv, np.float64)
        rc = self.lib.xa_rollout_gradient_lra(int(forward_precision), *self._common(), _ptr(c), _ptr(vel), _ptr(g), _ptr(gc),
                                          _ptr(gv), _ptr(out))
        if rc != 0:
            raise RuntimeError("xa_rollout_gradient failed")
        return out.reshape(-1, 3), gc, gv.reshape(-1, 3)
