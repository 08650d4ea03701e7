"""Save / restore of a soft body's state in the reference's byte stream
(SoftBodyMotionProperties::SaveState / RestoreState, SoftBodyMotionProperties.cpp:1446-1502).

* Restore: the reference's stream after 12 frames, restored into a fresh GPU body with
  the reference's body position and rotation, then 8 more frames: bit-exact with the
  reference's own continuation.
* Save: the GPU body's stream after 12 frames against the reference's: identical vertex
  positions and velocities, bounds, sleep timer and allow-sleeping; body velocities to
  float tolerance (the GPU sums the vertex velocities in tree order); force and torque
  zero.  A restore followed by a save returns the stream unchanged (sleep spheres aside,
  which soft bodies do not use and the GPU writes as zeros).
"""
from __future__ import annotations

import numpy as np
import pytest

import scene_zoo
from helpers import float_equal, reference_run
from oracle import have_reference

pytestmark = pytest.mark.gpu

HEADER = 101    # velocities 24, force 12, torque 12, spheres 48, timer 4, allow sleeping 1


def _split(blob, nv):
    h = np.frombuffer(blob[:100], np.float32)
    return dict(lin=h[0:3], ang=h[3:6], force=h[6:9], torque=h[9:12], spheres=h[12:24], timer=h[24],
                allow=blob[100], verts=np.frombuffer(blob[HEADER:HEADER + 24 * nv], np.float32).reshape(nv, 6),
                bounds=np.frombuffer(blob[HEADER + 24 * nv:], np.float32).reshape(4, 3))


def _reference_world(sc, settings):
    from oracle import JoltReference, JoltWorld
    w = JoltWorld(JoltReference("nofma"), sc.gravity)
    for c in sc.colliders:
        w.add_collider(c)
    for b, s in zip(sc.bodies, settings):
        w.add_soft_body(s, b.position, b.rotation, b.params)
    w.optimize_broadphase()
    return w


@pytest.mark.skipif(not have_reference("nofma"), reason="reference library not built")
def test_save_and_restore_match_the_reference_stream():
    from joltphysics_amd.xpbd import Context
    sc = scene_zoo.cloth_hulls()
    settings, bounds, _ = reference_run(sc, (1,))
    for c, b in zip(sc.colliders, bounds):
        c.bounds = tuple(float(x) for x in b)
    nv = len(settings[0].vertices)

    w = _reference_world(sc, settings)
    w.update(sc.dt, sc.collision_steps, 12)
    ref_blob = w.save_state(0)
    ref_at_12 = w.read(0)
    w.update(sc.dt, sc.collision_steps, 8)
    ref_at_20 = w.read(0)
    w.close()

    # save: the GPU body after the same 12 frames
    ctx = Context(0)
    ctx.set_gravity(sc.gravity)
    ctx.set_colliders(sc.colliders)
    body = ctx.create_body(ctx.create_settings(settings[0]), sc.bodies[0].position, sc.bodies[0].rotation, sc.bodies[0].params)
    ctx.update(sc.dt, sc.collision_steps, 12)
    gpu_blob = ctx.save_state(body)
    assert len(gpu_blob) == len(ref_blob)
    g, r = _split(gpu_blob, nv), _split(ref_blob, nv)
    assert float_equal(g["verts"], r["verts"]), "vertex positions / velocities differ"
    assert float_equal(g["bounds"], r["bounds"]), "bounds differ"
    assert g["timer"] == r["timer"] and g["allow"] == r["allow"]
    assert not g["force"].any() and not g["torque"].any()
    np.testing.assert_allclose(g["lin"], r["lin"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g["ang"], r["ang"], rtol=1e-5, atol=1e-6)
    ctx.close()

    # restore: the reference's stream into a fresh GPU body, then 8 more frames
    ctx = Context(0)
    ctx.set_gravity(sc.gravity)
    ctx.set_colliders(sc.colliders)
    body = ctx.create_body(ctx.create_settings(settings[0]), sc.bodies[0].position, sc.bodies[0].rotation, sc.bodies[0].params)
    ctx.restore_state(body, ref_blob, ref_at_12["body_pos"], ref_at_12["body_rot"])
    again = ctx.save_state(body)
    assert again[:48] == ref_blob[:48] and again[96:] == ref_blob[96:], "restore then save changed the stream"
    ctx.update(sc.dt, sc.collision_steps, 8)
    out = ctx.read(body)
    ctx.close()
    assert float_equal(out["pos"], ref_at_20["pos"]), "positions differ after the restored continuation"
    assert float_equal(out["vel"], ref_at_20["vel"]), "velocities differ after the restored continuation"
    assert float_equal(out["body_pos"], ref_at_20["body_pos"])
