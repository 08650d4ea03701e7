"""The Jolt-side adapter (joltphysics_amd/adapter/jolt_xpbd_adapter.cpp) drives a Jolt
PhysicsSystem's soft bodies through the C ABI.

Two identical reference worlds (the reference built with FMA contraction off, the adapter and
jolt_oracle.cpp in one library, oracle/Makefile target `adapter`): in one, PhysicsSystem::Update
simulates the soft bodies; in the other, JoltXpbd::SoftBodyUpdater takes them over (settings
from the body's SoftBodySharedSettings, state through SaveState, colliders from the broadphase
query, results back through RestoreState and BodyInterface::SetPosition) and
PhysicsSystem::Update runs with them inactive.  Bar: the vertex positions and velocities and
the body positions that Jolt's own API returns are float-equal in both worlds every checked
frame.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import scene_zoo
from helpers import float_equal, max_abs
from joltphysics_amd import scenes
from oracle import JoltReference, JoltWorld, have_reference

pytestmark = pytest.mark.gpu

SCENES = {
    "cloth32_on_box": lambda: scenes.config1_cloth_on_box(frames=20),
    "tetcube_plane_box": scene_zoo.tet_cube_on_plane,
    "lra_cloth": scene_zoo.lra_cloth,
    "cloth_cylinder_tapered": scene_zoo.cloth_cylinder_tapered,
    "cloth_hulls": scene_zoo.cloth_hulls,
    "pressure_sphere": scene_zoo.pressure_sphere,
    "batched": scene_zoo.batched,
    "cloth100_sphere": scene_zoo.big_cloth_on_sphere,
    "cloth_scaled_shapes": scene_zoo.cloth_scaled_shapes,
    "cloth_on_mesh": scene_zoo.cloth_on_mesh,
    "cloth_on_height_field": scene_zoo.cloth_on_height_field,
    "cloth_on_compounds": scene_zoo.cloth_on_compounds,
}


@pytest.fixture(scope="module")
def ref():
    if not have_reference("adapter"):
        pytest.skip("adapter library not built (make -C oracle adapter)")
    from joltphysics_amd import xpbd
    xpbd.load_library()             # the HIP runtime of torch first, as the binding does
    r = JoltReference("adapter")
    r.lib.ja_create.restype = C.c_void_p
    r.lib.ja_create.argtypes = [C.c_void_p, C.c_int]
    r.lib.ja_destroy.argtypes = [C.c_void_p]
    r.lib.ja_add.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    r.lib.ja_update.argtypes = [C.c_void_p, C.c_float, C.c_int]
    r.lib.ja_last_error.restype = C.c_char_p
    r.lib.ja_last_error.argtypes = [C.c_void_p]
    return r


def _world(ref, sc, settings):
    w = JoltWorld(ref, sc.gravity)
    for c in sc.colliders:
        w.add_collider(c)
    for b, s in zip(sc.bodies, settings):
        w.add_soft_body(s, b.position, b.rotation, b.params)
    w.optimize_broadphase()
    return w


@pytest.mark.parametrize("name", list(SCENES))
def test_adapter_drives_jolt_bit_exact(ref, name):
    sc = SCENES[name]()
    settings = [ref.build_settings(b.mesh) for b in sc.bodies]
    for b, s in zip(sc.bodies, settings):
        if b.post is not None:
            b.post(s)
    cpu, gpu = _world(ref, sc, settings), _world(ref, sc, settings)
    lib = ref.lib
    u = lib.ja_create(gpu.h, 0)
    assert u, "adapter context"
    try:
        for i in range(len(sc.bodies)):
            assert lib.ja_add(u, gpu.h, i) == 0, lib.ja_last_error(u).decode()
        frames = sorted({1, 2, sc.frames})
        print(f"adapter {name}: {len(sc.bodies)} soft bodies, {len(sc.colliders)} colliders, frames {frames}", flush=True)
        for f in range(1, sc.frames + 1):
            cpu.update(sc.dt, sc.collision_steps, 1)
            gpu.update(sc.dt, sc.collision_steps, 1)
            assert lib.ja_update(u, sc.dt, sc.collision_steps) == 0, lib.ja_last_error(u).decode()
            if f not in frames:
                continue
            for i in range(len(sc.bodies)):
                a, b = cpu.read(i), gpu.read(i)
                print(f"  frame {f} body {i}: max |dpos| {max_abs(a['pos'], b['pos']):.3e}", flush=True)
                assert float_equal(a["pos"], b["pos"]), f"{name}: positions differ at frame {f}, body {i}"
                assert float_equal(a["vel"], b["vel"]), f"{name}: velocities differ at frame {f}, body {i}"
                assert float_equal(a["body_pos"], b["body_pos"]), f"{name}: body position differs at frame {f}, body {i}"
    finally:
        lib.ja_destroy(u)
        cpu.close()
        gpu.close()
