"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container that has the reference sources (oracle/Makefile builds
oracle/_ref/libjolt_oracle_nofma.so from them):

    make -C oracle -j8 && python tests/golden/make_fixtures.py

Fixtures are data only: the settings arrays the reference's own
CreateConstraints + Optimize produce for a scene, and the vertex positions /
velocities / body positions its PhysicsSystem::Update produces, frame by frame.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from joltphysics_amd import scenes  # noqa: E402
from oracle import JoltReference, JoltWorld  # noqa: E402

CONFIG1_FRAMES = (1, 2, 3, 5, 10, 20, 50, 100, 200, 300)


def run_scene(sc, frames_to_keep, variant="nofma"):
    ref = JoltReference(variant)
    settings = [ref.build_settings(b.mesh) for b in sc.bodies]
    for b, s in zip(sc.bodies, settings):
        if b.post is not None:
            b.post(s)
    world = JoltWorld(ref, sc.gravity)
    ids = [world.add_collider(c) for c in sc.colliders]
    bounds = np.array([world.collider_bounds(i) for i in ids], np.float32).reshape(-1, 6)
    for b, s in zip(sc.bodies, settings):
        world.add_soft_body(s, b.position, b.rotation, b.params)
    world.optimize_broadphase()
    traj = {}
    last = max(frames_to_keep)
    for f in range(1, last + 1):
        world.update(sc.dt, sc.collision_steps, 1)
        if f in frames_to_keep:
            traj[f] = [world.read(i) for i in range(len(sc.bodies))]
    return settings, bounds, traj


def save_config1():
    sc = scenes.config1_cloth_on_box()
    settings, bounds, traj = run_scene(sc, CONFIG1_FRAMES)
    s = settings[0]
    out = dict(vertices=s.vertices, faces=s.faces, edges=s.edges, bends=s.bends, volumes=s.volumes,
               lras=s.lras, groups=s.groups, collider_bounds=bounds,
               frames=np.array(CONFIG1_FRAMES, np.int32))
    for f in CONFIG1_FRAMES:
        r = traj[f][0]
        out[f"pos_{f}"] = r["pos"]
        out[f"vel_{f}"] = r["vel"]
        out[f"body_pos_{f}"] = r["body_pos"]
    path = os.path.join(HERE, "config1_cloth32_box.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    save_config1()
