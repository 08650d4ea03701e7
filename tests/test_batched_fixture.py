"""The reference digests of the batched 2048-body scene (tests/golden/batched2048_sample_digests.json,
written by tests/golden/make_fixtures.py from the reference built with FMA contraction off).

bench.py compares the GPU's state of the 64 sampled bodies after its warm-up frames
with these digests.  Here they are pinned on the CPU: the sample is the one the
workload generator picks from the host builder's settings, and the own FP32
restatement of the update (bit-exact with the reference on every scene) reproduces
the digests of a spread of the sampled bodies.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, run_restatement
from joltphysics_amd import scenes
from joltphysics_amd.xpbd import build_settings


@pytest.fixture(scope="module")
def batched():
    with open(os.path.join(GOLDEN, "batched2048_sample_digests.json")) as f:
        fx = json.load(f)
    sc = scenes.config3_batched(frames=8)
    settings = scenes.resolve_meshes(sc, build_settings)
    return fx, sc, settings


def test_sample_is_the_generators(batched):
    fx, sc, settings = batched
    assert fx["scene"] == sc.name and fx["bodies"] == len(sc.bodies)
    assert fx["sample"] == scenes.batched_check_sample(sc, settings)
    kinds = {sc.bodies[k].mesh[0] for k in fx["sample"]}
    assert kinds == {"cloth", "cube", "ico"}
    sizes = [len(settings[k].vertices) for k in fx["sample"]]
    assert max(sizes) > 12288 and min(sizes) < 200


def test_restatement_reproduces_digests(batched):
    fx, sc, settings = batched
    sample = fx["sample"]
    # every fourth sampled body: small to the largest streamed cloths, cubes, icospheres
    # (the bodies whose cloths blow up in the first frames reach the ground plane's bounds,
    # so the plane's world bounds must be the reference's: scenes.plane_world_bounds)
    pick = sorted(set(sample[::4]) | {170, 286})
    frames = tuple(range(1, 9))
    sub = scenes.Scene(sc.name, [sc.bodies[k] for k in pick], sc.colliders, sc.gravity, sc.dt, sc.collision_steps, 8)
    bounds = np.array([c.bounds for c in sc.colliders], np.float32)
    out = run_restatement(sub, [settings[k] for k in pick], bounds, frames_to_keep=frames)
    for f in frames:
        for j, k in enumerate(pick):
            assert scenes.state_digest(out[f][j]) == fx["frames"][str(f)][sample.index(k)], (f, k)


def test_plane_world_bounds_are_the_references():
    """scenes.plane_world_bounds (the world AABB the update's DetermineCollidingShapes tests
    against) equals Body::GetWorldSpaceBounds of the reference for rotated, offset planes, to
    the last bit of float32 for axis-aligned ones and within 2 ulps for rotated ones (the
    rotation matrix is built in float64 here, in float32 by Mat44::sRotation there).  The GPU
    parity tests take the reference's own bounds; these serve the benchmark's workloads."""
    from oracle import JoltReference, JoltWorld, have_reference
    from joltphysics_amd.records import SHAPE_PLANE, Collider
    if not have_reference("nofma"):
        pytest.skip("reference library not built")
    rng = np.random.Generator(np.random.PCG64(9))
    world = JoltWorld(JoltReference("nofma"), (0.0, 0.0, -9.81))
    try:
        for k in range(12):
            n = rng.normal(size=3) if k else np.array([0.0, 0.0, 1.0])
            n = n / np.linalg.norm(n)
            c = Collider(SHAPE_PLANE, (float(n[0]), float(n[1]), float(n[2]), float(rng.uniform(-3, 3)),
                                       float(rng.uniform(5, 1000))),
                         pos=tuple(float(x) for x in rng.uniform(-10, 10, 3)), rot=scenes.random_quaternion(rng))
            i = world.add_collider(c)
            got, want = np.array(scenes.plane_world_bounds(c)), np.array(world.collider_bounds(i), np.float32)
            assert np.all(np.abs(got - want) <= 2 * np.spacing(np.abs(want))), (k, got - want)
        c = scenes.config3_batched(frames=1).colliders[0]
        assert scenes.plane_world_bounds(c) == world.collider_bounds(world.add_collider(c))
    finally:
        world.close()


def test_restatement_reproduces_coloured_digests(batched):
    """The coloured-schedule digests (batched2048_coloured_sample_digests.json, the reference on
    recoloured settings): the restatement on the same recoloured settings reproduces a spread of
    them, and they differ from the exact order's digests."""
    from helpers import recoloured_settings
    fx, sc, settings = batched
    with open(os.path.join(GOLDEN, "batched2048_coloured_sample_digests.json")) as f:
        cx = json.load(f)
    assert cx["sample"] == fx["sample"]
    sample = cx["sample"]
    pick = sorted(set(sample[1::8]))
    frames = (1, 2, 3)
    sub = scenes.Scene(sc.name, [sc.bodies[k] for k in pick], sc.colliders, sc.gravity, sc.dt, sc.collision_steps, 3)
    bounds = np.array([c.bounds for c in sc.colliders], np.float32)
    out = run_restatement(sub, [recoloured_settings(settings[k])[0] for k in pick], bounds, frames_to_keep=frames)
    for f in frames:
        for j, k in enumerate(pick):
            assert scenes.state_digest(out[f][j]) == cx["frames"][str(f)][sample.index(k)], (f, k)
    assert sum(a != b for a, b in zip(cx["frames"]["3"], fx["frames"]["3"])) > len(sample) // 2
