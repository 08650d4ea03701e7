"""ctypes binding of the C ABI in include/jolt_xpbd.h (libjolt_xpbd.so).

The binding is a thin mirror of the reference's soft-body interface:
``Context.create_settings`` takes the arrays of a SoftBodySharedSettings after
Optimize, ``Context.create_body`` the fields of SoftBodyCreationSettings,
``Context.set_colliders`` the static bodies and ``Context.update`` runs
PhysicsSystem::Update(dt, collision_steps) for every soft body on the GPU.

There is no CPU fallback: if the HIP library is missing or no GPU is present
the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

from .records import (BEND, EDGE, FACE, GROUP, INV_BIND, LRA, MOTION_DYNAMIC, PART_OP, ROD_BT, ROD_SS, SHAPE_COMPOUND,
                      SHAPE_CONVEX_HULL, SHAPE_MESH, SKINNED, TREE_NODE, VERTEX, VOLUME, BodyParams, SettingsArrays)

_HERE = os.path.dirname(os.path.abspath(__file__))
# JOLT_XPBD_LIBRARY selects another build of the library (the sanitized one in the CPU tests)
LIB_PATH = os.environ.get("JOLT_XPBD_LIBRARY") or os.path.join(_HERE, "_lib", "libjolt_xpbd.so")

XPBD_OK = 0


class XpbdError(RuntimeError):
    pass


class _SettingsDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint32),
                ("faces", C.c_void_p), ("num_faces", C.c_uint32),
                ("edges", C.c_void_p), ("num_edges", C.c_uint32),
                ("dihedral_bends", C.c_void_p), ("num_dihedral_bends", C.c_uint32),
                ("volumes", C.c_void_p), ("num_volumes", C.c_uint32),
                ("lras", C.c_void_p), ("num_lras", C.c_uint32),
                ("update_groups", C.c_void_p), ("num_update_groups", C.c_uint32),
                ("num_skinned", C.c_uint32), ("num_rod_stretch_shear", C.c_uint32),
                ("num_rod_bend_twist", C.c_uint32), ("skinned", C.c_void_p),
                ("skinned_normals", C.c_void_p), ("num_skinned_normals", C.c_uint32),
                ("inv_bind_matrices", C.c_void_p), ("num_inv_bind_matrices", C.c_uint32),
                ("rod_stretch_shear", C.c_void_p), ("rod_bend_twist", C.c_void_p)]


class _BodyDesc(C.Structure):
    _fields_ = [("settings", C.c_void_p), ("position", C.c_float * 3), ("rotation", C.c_float * 4),
                ("num_iterations", C.c_uint32), ("linear_damping", C.c_float),
                ("max_linear_velocity", C.c_float), ("restitution", C.c_float), ("friction", C.c_float),
                ("pressure", C.c_float), ("gravity_factor", C.c_float), ("vertex_radius", C.c_float),
                ("update_position", C.c_int32), ("make_rotation_identity", C.c_int32),
                ("allow_sleeping", C.c_int32)]


class _ConvexHull(C.Structure):
    _fields_ = [("num_points", C.c_uint32), ("points", C.c_void_p), ("num_faces", C.c_uint32),
                ("planes", C.c_void_p), ("face_first", C.c_void_p), ("face_count", C.c_void_p),
                ("num_vertex_index", C.c_uint32), ("vertex_index", C.c_void_p)]


class _TriangleTree(C.Structure):
    _fields_ = [("root", C.c_uint32), ("num_nodes", C.c_uint32), ("nodes", C.c_void_p),
                ("num_triangles", C.c_uint32), ("triangles", C.c_void_p), ("flags", C.c_uint32)]


class _ColliderPart(C.Structure):
    _fields_ = [("type", C.c_int32), ("params", C.c_float * 8), ("scale", C.c_float * 3),
                ("hull", C.POINTER(_ConvexHull)), ("tree", C.POINTER(_TriangleTree)),
                ("num_ops", C.c_uint32), ("ops", C.c_void_p)]


class _ColliderDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("params", C.c_float * 8), ("position", C.c_float * 3),
                ("rotation", C.c_float * 4), ("friction", C.c_float), ("restitution", C.c_float),
                ("world_bounds", C.c_float * 6), ("hull", C.POINTER(_ConvexHull)),
                ("motion_type", C.c_int32), ("inv_mass", C.c_float), ("inv_inertia_diagonal", C.c_float * 3),
                ("inertia_rotation", C.c_float * 4), ("linear_velocity", C.c_float * 3),
                ("angular_velocity", C.c_float * 3), ("is_sensor", C.c_int32), ("user_id", C.c_uint32),
                ("scale", C.c_float * 3), ("tree", C.POINTER(_TriangleTree)),
                ("num_parts", C.c_uint32), ("parts", C.POINTER(_ColliderPart))]


def _hull_struct(h, keep):
    a = (np.ascontiguousarray(h.points, np.float32), np.ascontiguousarray(h.planes, np.float32),
         np.ascontiguousarray(h.face_first, np.uint32), np.ascontiguousarray(h.face_count, np.uint32),
         np.ascontiguousarray(h.vertex_index, np.uint32))
    d = _ConvexHull(len(a[0]), a[0].ctypes.data, len(a[1]), a[1].ctypes.data, a[2].ctypes.data,
                    a[3].ctypes.data, len(a[4]), a[4].ctypes.data)
    keep.append((a, d))
    return C.pointer(d)


def _tree_struct(t, keep):
    a = (np.ascontiguousarray(t.nodes, TREE_NODE), np.ascontiguousarray(t.triangles, np.float32).reshape(-1, 9))
    d = _TriangleTree(int(t.root), len(a[0]), a[0].ctypes.data if len(a[0]) else None, len(a[1]),
                      a[1].ctypes.data if len(a[1]) else None, int(t.flags))
    keep.append((a, d))
    return C.pointer(d)


class _RolloutDesc(C.Structure):
    _fields_ = [("settings", C.c_void_p), ("num_bodies", C.c_uint32), ("num_iterations", C.c_uint32),
                ("linear_damping", C.c_float), ("max_linear_velocity", C.c_float), ("gravity_factor", C.c_float),
                ("dt", C.c_float), ("frames", C.c_uint32), ("checkpoint_interval", C.c_uint32), ("pressure", C.c_float)]


class _VertexAttributes(C.Structure):
    _fields_ = [("compliance", C.c_float), ("shear_compliance", C.c_float), ("bend_compliance", C.c_float),
                ("lra_type", C.c_int32), ("lra_max_distance_multiplier", C.c_float)]


class _BuildDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint32),
                ("faces", C.c_void_p), ("num_faces", C.c_uint32),
                ("edges", C.c_void_p), ("num_edges", C.c_uint32),
                ("volumes", C.c_void_p), ("num_volumes", C.c_uint32),
                ("lras", C.c_void_p), ("num_lras", C.c_uint32),
                ("attributes", C.c_void_p), ("num_attributes", C.c_uint32),
                ("bend_type", C.c_int32), ("angle_tolerance", C.c_float),
                ("lra_max_distance_multiplier", C.c_float), ("flags", C.c_uint32)]


# Exported symbols of the C ABI, in declaration order of include/jolt_xpbd.h.
EXPORTED_SYMBOLS = (
    "xpbd_last_error", "xpbd_context_create", "xpbd_context_destroy", "xpbd_context_stream",
    "xpbd_settings_create", "xpbd_settings_create_scheduled", "xpbd_settings_release", "xpbd_settings_stats",
    "xpbd_schedule_build_scheduled", "xpbd_schedule_build",
    "xpbd_schedule_sizes", "xpbd_schedule_read", "xpbd_settings_team_stats",
    "xpbd_settings_build", "xpbd_built_counts", "xpbd_built_read", "xpbd_built_free",
    "xpbd_build_last_error", "xpbd_body_create",
    "xpbd_set_colliders", "xpbd_set_gravity", "xpbd_update", "xpbd_context_set_team_min_windows", "xpbd_context_set_team_size",
    "xpbd_body_set_tail_mode", "xpbd_synchronize", "xpbd_body_read",
    "xpbd_body_write", "xpbd_body_state", "xpbd_body_placement", "xpbd_body_velocity_deltas", "xpbd_body_contacts", "xpbd_body_sensor_contacts",
    "xpbd_body_skin_vertices", "xpbd_body_rod_states", "xpbd_body_set_contact_settings",
    "xpbd_body_set_skin_options", "xpbd_body_save_state", "xpbd_body_restore_state", "xpbd_profile_enable",
    "xpbd_profile_read",
    "xpbd_rollout_create", "xpbd_rollout_destroy", "xpbd_rollout_sizes", "xpbd_rollout_forward",
    "xpbd_rollout_forward_f64", "xpbd_rollout_backward", "xpbd_rollout_forward_noalloc",
)

_lib = None


def load_library(path=LIB_PATH):
    """Loads libjolt_xpbd.so; raises if it has not been built.

    PyTorch ships its own HIP runtime under the same soname; importing it first
    makes the library bind to that one runtime, so torch device pointers and
    stream handles are valid for the library's kernels."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(path):
        raise XpbdError(f"HIP library {path} is missing; run joltphysics_amd/build.py "
                        "(there is no CPU fallback)")
    lib = C.CDLL(path)
    vp, i32, u32, f = C.c_void_p, C.c_int32, C.c_uint32, C.c_float
    lib.xpbd_last_error.restype = C.c_char_p
    lib.xpbd_context_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.xpbd_context_destroy.argtypes = [vp]
    lib.xpbd_context_stream.restype = vp
    lib.xpbd_context_stream.argtypes = [vp]
    lib.xpbd_settings_create.argtypes = [vp, C.POINTER(_SettingsDesc), C.POINTER(vp)]
    lib.xpbd_settings_release.argtypes = [vp]
    lib.xpbd_settings_stats.argtypes = [vp, C.POINTER(u32)]
    lib.xpbd_settings_team_stats.argtypes = [vp, C.POINTER(u32)]
    lib.xpbd_schedule_build.argtypes = [C.POINTER(_SettingsDesc), C.POINTER(vp)]
    lib.xpbd_schedule_build_scheduled.argtypes = [C.POINTER(_SettingsDesc), i32, C.POINTER(vp)]
    lib.xpbd_settings_create_scheduled.argtypes = [vp, C.POINTER(_SettingsDesc), i32, C.POINTER(vp)]
    lib.xpbd_schedule_sizes.argtypes = [vp, C.POINTER(u32)]
    lib.xpbd_schedule_read.argtypes = [vp] * 11
    lib.xpbd_context_set_team_min_windows.argtypes = [vp, u32]
    lib.xpbd_context_set_team_size.argtypes = [vp, u32]
    lib.xpbd_body_set_tail_mode.argtypes = [vp, i32, i32]
    lib.xpbd_body_placement.argtypes = [vp, i32, vp]
    lib.xpbd_settings_build.argtypes = [C.POINTER(_BuildDesc), C.POINTER(vp)]
    lib.xpbd_built_counts.argtypes = [vp, C.POINTER(u32)]
    lib.xpbd_built_read.argtypes = [vp] * 8
    lib.xpbd_built_free.argtypes = [vp]
    lib.xpbd_build_last_error.restype = C.c_char_p
    lib.xpbd_body_create.argtypes = [vp, C.POINTER(_BodyDesc), C.POINTER(i32)]
    lib.xpbd_set_colliders.argtypes = [vp, C.POINTER(_ColliderDesc), u32]
    lib.xpbd_set_gravity.argtypes = [vp, C.POINTER(f)]
    lib.xpbd_update.argtypes = [vp, f, i32, i32]
    lib.xpbd_synchronize.argtypes = [vp]
    lib.xpbd_body_read.argtypes = [vp, i32] + [vp] * 8
    lib.xpbd_body_write.argtypes = [vp, i32, vp, vp]
    lib.xpbd_body_state.argtypes = [vp, i32, vp]
    lib.xpbd_body_velocity_deltas.argtypes = [vp, i32, u32, vp, vp, vp]
    lib.xpbd_body_contacts.argtypes = [vp, i32, vp, vp]
    lib.xpbd_body_sensor_contacts.argtypes = [vp, i32, u32, vp]
    lib.xpbd_body_rod_states.argtypes = [vp, i32, u32, vp, vp]
    lib.xpbd_body_set_contact_settings.argtypes = [vp, i32, u32, vp]
    lib.xpbd_body_skin_vertices.argtypes = [vp, i32, vp, u32, i32]
    lib.xpbd_body_set_skin_options.argtypes = [vp, i32, i32, f]
    lib.xpbd_body_save_state.argtypes = [vp, i32, vp, C.c_size_t, C.POINTER(C.c_size_t)]
    lib.xpbd_body_restore_state.argtypes = [vp, i32, vp, C.c_size_t, vp, vp]
    lib.xpbd_profile_enable.argtypes = [vp, i32]
    lib.xpbd_profile_read.argtypes = [vp, C.POINTER(i32), C.POINTER(C.c_double)]
    lib.xpbd_rollout_create.argtypes = [vp, C.POINTER(_RolloutDesc), C.POINTER(vp)]
    lib.xpbd_rollout_destroy.argtypes = [vp]
    lib.xpbd_rollout_sizes.argtypes = [vp, C.POINTER(u32), C.POINTER(u32)]
    lib.xpbd_rollout_forward.argtypes = [vp] * 5
    lib.xpbd_rollout_forward_f64.argtypes = [vp] * 5
    lib.xpbd_rollout_forward_noalloc.argtypes = [vp] * 5
    lib.xpbd_rollout_backward.argtypes = [vp] * 5
    for name in EXPORTED_SYMBOLS:
        if name not in ("xpbd_last_error", "xpbd_context_stream", "xpbd_build_last_error"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _check(rc):
    if rc != XPBD_OK:
        raise XpbdError(f"xpbd error {rc}: {_lib.xpbd_last_error().decode()}")


def _ptr(a):
    if a is None or a.size == 0:
        return None
    return a.ctypes.data_as(C.c_void_p)


def _settings_desc(s: SettingsArrays):
    """xpbd_settings_desc of the arrays (and the contiguous copies it points into)."""
    v = np.ascontiguousarray(s.vertices, VERTEX)
    f = s.faces
    if f.dtype != FACE:
        f4 = np.zeros(len(f), FACE)
        f4["v"] = f
        f = f4
    f = np.ascontiguousarray(f, FACE)
    e = np.ascontiguousarray(s.edges, EDGE)
    b = np.ascontiguousarray(s.bends, BEND)
    t = np.ascontiguousarray(s.volumes, VOLUME)
    lr = np.ascontiguousarray(s.lras, LRA)
    g = np.ascontiguousarray(s.groups, GROUP)
    sk = np.ascontiguousarray(s.skinned, SKINNED)
    skn = np.ascontiguousarray(s.skinned_normals, np.uint32)
    ib = np.ascontiguousarray(s.inv_binds, INV_BIND)
    rs = np.ascontiguousarray(s.rod_stretch_shear, ROD_SS)
    rb = np.ascontiguousarray(s.rod_bend_twist, ROD_BT)
    d = _SettingsDesc(_ptr(v), len(v), _ptr(f), len(f), _ptr(e), len(e), _ptr(b), len(b),
                      _ptr(t), len(t), _ptr(lr), len(lr), _ptr(g), len(g), len(sk), len(rs), len(rb),
                      _ptr(sk), _ptr(skn), len(skn), _ptr(ib), len(ib), _ptr(rs), _ptr(rb))
    return d, (v, f, e, b, t, lr, g, sk, skn, ib, rs, rb)


SCHEDULES = {"exact": 0, "coloured_tail": 1, "coloured": 2}     # XPBD_SCHEDULE_*


def host_schedule(s: SettingsArrays, schedule="exact") -> dict:
    """The level programs the device would run for these settings, built on the host
    (xpbd_schedule_build_scheduled; no device needed): records (n, 4) uint32, their constraint
    indices, level offsets, windows (n, 4), the gathered windows' vertex list, the coloured
    non-parallel group and the identity program bundles merge."""
    lib = load_library()
    d, _keep = _settings_desc(s)
    h = C.c_void_p()
    rc = lib.xpbd_schedule_build_scheduled(C.byref(d), SCHEDULES[schedule], C.byref(h))
    if rc != XPBD_OK:
        raise XpbdError(f"xpbd error {rc}: {lib.xpbd_last_error().decode()}")
    try:
        n = (C.c_uint32 * 10)()
        _check(lib.xpbd_schedule_sizes(h, n))
        out = dict(recs=np.zeros((n[0], 4), np.uint32), rec_index=np.zeros(n[0], np.uint32),
                   lvl=np.zeros(n[1], np.uint32), windows=np.zeros((n[2], 4), np.uint32),
                   vlist=np.zeros(n[4], np.uint32), colour_recs=np.zeros((n[5], 8), np.uint32),
                   colour_begin=np.zeros(n[6], np.uint32), identity_recs=np.zeros((n[8], 4), np.uint32),
                   identity_rec_index=np.zeros(n[8], np.uint32), identity_level_begin=np.zeros(n[9], np.uint32))
        _check(lib.xpbd_schedule_read(h, *(_ptr(out[k]) for k in (
            "recs", "rec_index", "lvl", "windows", "vlist", "colour_recs", "colour_begin", "identity_recs",
            "identity_rec_index", "identity_level_begin"))))
        out["num_par_windows"] = int(n[3])
        out["resident"] = bool(n[7])
        return out
    finally:
        lib.xpbd_settings_release(h)


def build_settings(mesh) -> SettingsArrays:
    """Builds a SoftBodySharedSettings from a MeshSpec on the CPU (CreateConstraints,
    Calculate*, Optimize restated by the host library; no device needed)."""
    lib = load_library()
    v = np.ascontiguousarray(mesh.vertices, VERTEX)
    f = np.ascontiguousarray(mesh.faces, np.uint32).reshape(-1, 3)
    e = np.ascontiguousarray(mesh.edges, EDGE)
    t = np.ascontiguousarray(mesh.volumes, VOLUME)
    lr = np.ascontiguousarray(mesh.lras, LRA)
    a = np.asarray(mesh.attrs, np.float32).reshape(-1, 5)
    attrs = (_VertexAttributes * max(len(a), 1))()
    for i, row in enumerate(a):
        attrs[i].compliance, attrs[i].shear_compliance, attrs[i].bend_compliance = float(row[0]), float(row[1]), float(row[2])
        attrs[i].lra_type = int(row[3])
        attrs[i].lra_max_distance_multiplier = float(row[4])
    d = _BuildDesc(_ptr(v), len(v), _ptr(f), len(f), _ptr(e), len(e), _ptr(t), len(t), _ptr(lr), len(lr),
                   C.cast(attrs, C.c_void_p) if len(a) else None, len(a), int(mesh.bend_type),
                   float(mesh.angle_tolerance), float(mesh.lra_multiplier), int(mesh.flags))
    h = C.c_void_p()
    rc = lib.xpbd_settings_build(C.byref(d), C.byref(h))
    if rc != XPBD_OK:
        raise XpbdError(f"settings build error {rc}: {lib.xpbd_build_last_error().decode()}")
    try:
        counts = (C.c_uint32 * 7)()
        lib.xpbd_built_counts(h, counts)
        nv, nf, ne, nb, nt, nl, ng = (int(c) for c in counts)
        s = SettingsArrays(np.zeros(nv, VERTEX), np.zeros(nf, FACE), np.zeros(ne, EDGE), np.zeros(nb, BEND),
                           np.zeros(nt, VOLUME), np.zeros(nl, LRA), np.zeros(ng, GROUP))
        lib.xpbd_built_read(h, _ptr(s.vertices), _ptr(s.faces), _ptr(s.edges), _ptr(s.bends),
                            _ptr(s.volumes), _ptr(s.lras), _ptr(s.groups))
        return s
    finally:
        lib.xpbd_built_free(h)


class Settings:
    def __init__(self, ctx, handle, arrays):
        self.ctx = ctx
        self.handle = handle
        self.arrays = arrays

    def stats(self):
        out = (C.c_uint32 * 8)()
        _check(self.ctx.lib.xpbd_settings_stats(self.handle, out))
        team = (C.c_uint32 * 4)()
        _check(self.ctx.lib.xpbd_settings_team_stats(self.handle, team))
        return dict(parallel_groups=out[0], max_group_levels=out[1], tail_levels=out[2],
                    tail_constraints=out[3], constraints=out[4], resident=bool(out[5]),
                    parallel_windows=out[6], tail_windows=out[7], tail_colourable=bool(team[2]),
                    tail_colours=team[3])


class Context:
    """One device context: settings, bodies, colliders, one HIP stream."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.xpbd_context_create(device, C.byref(h)))
        self.h = h
        self.sizes = []
        self._keep = []
        self._body_rods = []

    def close(self):
        if self.h:
            self.lib.xpbd_context_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down: leave it to the process exit
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        return self.lib.xpbd_context_stream(self.h)

    def set_gravity(self, g):
        arr = (C.c_float * 3)(*map(float, g))
        _check(self.lib.xpbd_set_gravity(self.h, arr))

    def set_colliders(self, colliders):
        n = len(colliders)
        arr = (_ColliderDesc * max(n, 1))()
        keep = []   # hull arrays stay alive until the call has copied them
        for i, c in enumerate(colliders):
            if c.type == SHAPE_COMPOUND:
                if not c.parts:
                    raise XpbdError("compound collider without parts (take them from the reference's shape)")
                parts = (_ColliderPart * len(c.parts))()
                for k, pt in enumerate(c.parts):
                    parts[k].type = int(pt.type)
                    for j in range(8):
                        parts[k].params[j] = float(pt.params[j]) if j < len(pt.params) else 0.0
                    for j in range(3):
                        parts[k].scale[j] = float(pt.scale[j])
                    if pt.type == SHAPE_CONVEX_HULL:
                        parts[k].hull = _hull_struct(pt.hull, keep)
                    if pt.type == SHAPE_MESH:
                        parts[k].tree = _tree_struct(pt.tree, keep)
                    ops = np.ascontiguousarray(pt.ops, PART_OP)
                    keep.append(ops)
                    parts[k].num_ops = len(ops)
                    parts[k].ops = ops.ctypes.data if len(ops) else None
                keep.append(parts)
                arr[i].num_parts = len(c.parts)
                arr[i].parts = parts
            else:
                arr[i].type = int(c.type)
                if c.type == SHAPE_CONVEX_HULL:
                    if c.hull is None:
                        raise XpbdError("convex hull collider without hull data")
                    arr[i].hull = _hull_struct(c.hull, keep)
                if c.type == SHAPE_MESH:
                    if c.tree is None:
                        raise XpbdError("mesh collider without tree data")
                    arr[i].tree = _tree_struct(c.tree, keep)
                for k in range(3):
                    arr[i].scale[k] = float(c.scale[k])
                p = c.device_param_array()
                for k in range(8):
                    arr[i].params[k] = float(p[k])
            com = c.com_position()
            for k in range(3):
                arr[i].position[k] = float(com[k])
            for k in range(4):
                arr[i].rotation[k] = float(c.rot[k])
            arr[i].friction = float(c.friction)
            arr[i].restitution = float(c.restitution)
            if c.bounds is None:
                raise XpbdError("collider world bounds are required")
            for k in range(6):
                arr[i].world_bounds[k] = float(c.bounds[k])
            arr[i].motion_type = int(c.motion_type)
            arr[i].is_sensor = int(bool(c.is_sensor))
            # the collider's identity: its own id, else the Python object's
            arr[i].user_id = (int(c.user_id) if c.user_id is not None else id(c)) & 0xffffffff
            if c.motion_type == MOTION_DYNAMIC:
                arr[i].inv_mass = float(c.inv_mass)
                for k in range(3):
                    arr[i].inv_inertia_diagonal[k] = float(c.inv_inertia_diagonal[k])
                    arr[i].linear_velocity[k] = float(c.linear_velocity[k])
                    arr[i].angular_velocity[k] = float(c.angular_velocity[k])
                for k in range(4):
                    arr[i].inertia_rotation[k] = float(c.inertia_rotation[k])
        _check(self.lib.xpbd_set_colliders(self.h, arr, n))
        self._num_colliders = n

    def skin_vertices(self, body, joints, hard_skin_all=False):
        """SoftBodyMotionProperties::SkinVertices with joint matrices (n, 16), column major."""
        j = np.ascontiguousarray(joints, np.float32).reshape(-1, 16)
        _check(self.lib.xpbd_body_skin_vertices(self.h, int(body), j.ctypes.data, len(j), int(bool(hard_skin_all))))

    def set_skin_options(self, body, enable=True, max_distance_multiplier=1.0):
        _check(self.lib.xpbd_body_set_skin_options(self.h, int(body), int(bool(enable)), float(max_distance_multiplier)))

    def save_state(self, body) -> bytes:
        """SoftBodyMotionProperties::SaveState's byte stream of `body` (see the C ABI)."""
        n = C.c_size_t()
        _check(self.lib.xpbd_body_save_state(self.h, int(body), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        _check(self.lib.xpbd_body_save_state(self.h, int(body), buf, n.value, C.byref(n)))
        return buf.raw[:n.value]

    def restore_state(self, body, data: bytes, position=None, rotation=None):
        """SoftBodyMotionProperties::RestoreState from `data`; optionally the body's
        position and rotation too."""
        pos = None if position is None else np.ascontiguousarray(position, np.float32)
        rot = None if rotation is None else np.ascontiguousarray(rotation, np.float32)
        _check(self.lib.xpbd_body_restore_state(self.h, int(body), data, len(data), _ptr(pos), _ptr(rot)))

    def contacts(self, body):
        """(collider index per vertex, -1 without contact; local contact points (V, 3)) of the
        last collision step, as the contact listener's SoftBodyManifold reports them."""
        nv = self.sizes[body]
        col = np.zeros(max(nv, 1), np.int32)
        pts = np.zeros((max(nv, 1), 3), np.float32)
        _check(self.lib.xpbd_body_contacts(self.h, int(body), pts.ctypes.data, col.ctypes.data))
        return col[:nv], pts[:nv]

    def sensor_contacts(self, body):
        """Per collider: 1 where a sensor detected a vertex of `body` in the last collision step."""
        n = getattr(self, "_num_colliders", 0)
        hits = np.zeros(max(n, 1), np.int32)
        _check(self.lib.xpbd_body_sensor_contacts(self.h, int(body), n, hits.ctypes.data))
        return hits[:n]

    def set_contact_settings(self, body, settings=None):
        """Per-collider contact settings of `body` (n, 4): mInvMassScale1, mInvMassScale2,
        mInvInertiaScale2, accept (1 / 0), as a contact listener's OnSoftBodyContactValidate
        decides them; None removes the table."""
        if settings is None:
            _check(self.lib.xpbd_body_set_contact_settings(self.h, int(body), 0, None))
            return
        t = np.ascontiguousarray(settings, np.float32).reshape(-1, 4)
        _check(self.lib.xpbd_body_set_contact_settings(self.h, int(body), len(t), t.ctypes.data))

    def rod_states(self, body):
        """(rotations (n, 4) xyzw, angular velocities (n, 3)) of the Cosserat rods of `body`
        (SoftBodyMotionProperties::GetRodRotation / GetRodAngularVelocity)."""
        n = self._body_rods[int(body)]
        rot = np.zeros((max(n, 1), 4), np.float32)
        ang = np.zeros((max(n, 1), 3), np.float32)
        _check(self.lib.xpbd_body_rod_states(self.h, int(body), n, rot.ctypes.data, ang.ctypes.data))
        return rot[:n], ang[:n]

    def velocity_deltas(self, body):
        """(linear (n, 3), angular (n, 3), updated (n,)): the world-space velocity changes soft
        body `body` applied to each collider during the last update
        (SoftBodyMotionProperties::UpdateRigidBodyVelocities)."""
        n = getattr(self, "_num_colliders", 0)
        lin = np.zeros((max(n, 1), 3), np.float32)
        ang = np.zeros((max(n, 1), 3), np.float32)
        upd = np.zeros(max(n, 1), np.int32)
        _check(self.lib.xpbd_body_velocity_deltas(self.h, int(body), n, lin.ctypes.data, ang.ctypes.data,
                                                  upd.ctypes.data))
        return lin[:n], ang[:n], upd[:n]

    def create_settings(self, s: SettingsArrays, schedule="exact") -> Settings:
        """Uploads settings; `schedule` "exact" (the reference's constraint order), "coloured"
        or "coloured_tail" (the throughput schedules of xpbd_settings_create_scheduled)."""
        d, _keep = _settings_desc(s)
        h = C.c_void_p()
        _check(self.lib.xpbd_settings_create_scheduled(self.h, C.byref(d), SCHEDULES[schedule], C.byref(h)))
        st = Settings(self, h, s)
        self._keep.append(st)
        return st

    def create_body(self, settings: Settings, position=(0, 0, 0), rotation=(0, 0, 0, 1),
                    params: BodyParams | None = None) -> int:
        p = params or BodyParams()
        d = _BodyDesc()
        d.settings = settings.handle
        for k in range(3):
            d.position[k] = float(position[k])
        for k in range(4):
            d.rotation[k] = float(rotation[k])
        d.num_iterations = int(p.num_iterations)
        d.linear_damping = p.linear_damping
        d.max_linear_velocity = p.max_linear_velocity
        d.restitution = p.restitution
        d.friction = p.friction
        d.pressure = p.pressure
        d.gravity_factor = p.gravity_factor
        d.vertex_radius = p.vertex_radius
        d.update_position = int(p.update_position)
        d.make_rotation_identity = int(p.make_rotation_identity)
        d.allow_sleeping = int(p.allow_sleeping)
        out = C.c_int32()
        _check(self.lib.xpbd_body_create(self.h, C.byref(d), C.byref(out)))
        self.sizes.append(len(settings.arrays.vertices))
        self._body_rods.append(len(settings.arrays.rod_stretch_shear))
        return out.value

    def update(self, dt, collision_steps=1, frames=1):
        _check(self.lib.xpbd_update(self.h, float(dt), int(collision_steps), int(frames)))

    def synchronize(self):
        _check(self.lib.xpbd_synchronize(self.h))

    def read(self, body):
        n = self.sizes[body]
        out = dict(pos=np.zeros((n, 3), np.float32), prev=np.zeros((n, 3), np.float32),
                   vel=np.zeros((n, 3), np.float32), plane=np.zeros((n, 4), np.float32),
                   shape_idx=np.zeros(n, np.int32), contact=np.zeros(n, np.int32),
                   body_pos=np.zeros(3, np.float32), body_rot=np.zeros(4, np.float32))
        _check(self.lib.xpbd_body_read(self.h, body, *(_ptr(out[k]) for k in (
            "pos", "prev", "vel", "plane", "shape_idx", "contact", "body_pos", "body_rot"))))
        return out

    def write(self, body, pos=None, vel=None):
        pos = None if pos is None else np.ascontiguousarray(pos, np.float32)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32)
        _check(self.lib.xpbd_body_write(self.h, body, _ptr(pos), _ptr(vel)))

    def set_team_min_windows(self, windows):
        """Streamed bodies with at least `windows` parallel-group windows run on a team of
        workgroups (0: never)."""
        _check(self.lib.xpbd_context_set_team_min_windows(self.h, int(windows)))

    def set_team_size(self, workgroups):
        """Workgroups per team at least (0: one per parallel-group window)."""
        _check(self.lib.xpbd_context_set_team_size(self.h, int(workgroups)))

    def set_tail_mode(self, body, coloured):
        """The non-parallel group in the reference's order (False) or re-coloured (True)."""
        _check(self.lib.xpbd_body_set_tail_mode(self.h, body, 1 if coloured else 0))

    def placement(self, body):
        """(kind, threads per workgroup, bundle / team index, bodies in the bundle or team
        size) of the last update; kind 'bundle', 'streamed' or 'team'."""
        out = np.zeros(4, np.int32)
        _check(self.lib.xpbd_body_placement(self.h, body, _ptr(out)))
        return ("bundle", "streamed", "team")[int(out[0])], int(out[1]), int(out[2]), int(out[3])

    def body_state(self, body):
        out = np.zeros(14, np.float32)
        _check(self.lib.xpbd_body_state(self.h, body, _ptr(out)))
        return dict(linear_velocity=out[0:3], angular_velocity=out[3:6], bounds_min=out[6:9],
                    bounds_max=out[9:12], active=bool(out[12]), sleep_timer=float(out[13]))

    def create_rollout(self, settings, num_iterations=5, dt=1.0 / 60.0, frames=60, checkpoint_interval=10,
                       linear_damping=0.1, max_linear_velocity=500.0, gravity_factor=1.0, pressure=0.0) -> "Rollout":
        """Differentiable rollout of bodies with edge constraints only (see Rollout)."""
        return Rollout(self, settings, num_iterations, dt, frames, checkpoint_interval, linear_damping,
                       max_linear_velocity, gravity_factor, pressure)

    def profile_enable(self, on=True):
        _check(self.lib.xpbd_profile_enable(self.h, int(on)))

    def profile_read(self):
        n = C.c_int32()
        ms = C.c_double()
        _check(self.lib.xpbd_profile_read(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value


class Rollout:
    """Differentiable rollout (xpbd_rollout_* in include/jolt_xpbd.h).

    Bodies start at their settings' vertex positions with identity rotation and
    run `frames` frames of one collision step and `num_iterations` iterations.
    Parameters: per-edge compliance (bodies' mEdgeConstraints concatenated)
    and initial velocities (3 per vertex).  All array arguments are device
    pointers (ints, e.g. torch.Tensor.data_ptr()); `stream` is the hipStream_t
    handle the work is ordered on (None or 0: the default stream, where torch's
    default stream runs).
    """

    def __init__(self, ctx, settings, num_iterations, dt, frames, checkpoint_interval, linear_damping,
                 max_linear_velocity, gravity_factor, pressure=0.0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.settings = list(settings)
        handles = (C.c_void_p * len(self.settings))(*[s.handle.value for s in self.settings])
        d = _RolloutDesc(C.cast(handles, C.c_void_p), len(self.settings), int(num_iterations), float(linear_damping),
                         float(max_linear_velocity), float(gravity_factor), float(dt), int(frames),
                         int(checkpoint_interval), float(pressure))
        h = C.c_void_p()
        _check(self.lib.xpbd_rollout_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        ne, nv = C.c_uint32(), C.c_uint32()
        _check(self.lib.xpbd_rollout_sizes(h, C.byref(ne), C.byref(nv)))
        self.num_edges, self.num_vertices = ne.value, nv.value
        self.frames, self.num_iterations, self.dt = int(frames), int(num_iterations), float(dt)
        self.generation = 0     # float forward passes so far (the adjoint differentiates the last one)

    def close(self):
        if self.h:
            self.lib.xpbd_rollout_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down: leave it to the process exit
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def forward(self, compliance_ptr, velocity_ptr, positions_ptr, stream=None):
        _check(self.lib.xpbd_rollout_forward(self.h, compliance_ptr, velocity_ptr, positions_ptr, stream))
        self.generation += 1

    def forward_noalloc(self, compliance_ptr, velocity_ptr, positions_ptr, stream=None):
        """forward() through the allocation-free entry point (xpbd_rollout_forward_noalloc)."""
        _check(self.lib.xpbd_rollout_forward_noalloc(self.h, compliance_ptr, velocity_ptr, positions_ptr, stream))
        self.generation += 1

    def forward_f64(self, compliance_ptr, velocity_ptr, positions_ptr, stream=None):
        _check(self.lib.xpbd_rollout_forward_f64(self.h, compliance_ptr, velocity_ptr, positions_ptr, stream))

    def backward(self, grad_positions_ptr, grad_compliance_ptr, grad_velocity_ptr, stream=None):
        _check(self.lib.xpbd_rollout_backward(self.h, grad_positions_ptr, grad_compliance_ptr, grad_velocity_ptr,
                                              stream))
