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
    sc = scenes.config3_batched(frames=2)
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
    pick = sample[::4]
    sub = scenes.Scene(sc.name, [sc.bodies[k] for k in pick], sc.colliders, sc.gravity, sc.dt, sc.collision_steps, 2)
    bounds = np.array([c.bounds for c in sc.colliders], np.float32)
    out = run_restatement(sub, [settings[k] for k in pick], bounds, frames_to_keep=(1, 2))
    for f in (1, 2):
        for j, k in enumerate(pick):
            assert scenes.state_digest(out[f][j]) == fx["frames"][str(f)][sample.index(k)], (f, k)
