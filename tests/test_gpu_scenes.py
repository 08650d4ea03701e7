"""GPU parity on scenes covering every constraint type, collider shape, the
LDS-resident and streamed storage modes, pressure, LRA, sleeping and batching.

The reference (compiled from its sources with FMA contraction off, loaded
through oracle/oracle.py) builds the settings with its own CreateConstraints
and Optimize and runs PhysicsSystem::Update; the HIP path runs the same
settings through the C ABI.  Bar: float-equal positions, velocities and body
positions at every checked frame.
"""
from __future__ import annotations

import pytest

import scene_zoo
from helpers import float_equal, max_abs, reference_run, run_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(scene_zoo.ALL))
def test_scene_bit_exact(name):
    sc = scene_zoo.ALL[name]()
    frames = (1, 2, sc.frames)
    print(f"scene {name}: {len(sc.bodies)} bodies, frames {frames}", flush=True)
    settings, bounds, ref = reference_run(sc, frames)
    gpu = run_gpu(sc, settings, bounds=bounds, frames_to_keep=frames)
    for f in frames:
        for i in range(len(sc.bodies)):
            r, g = ref[f][i], gpu[f][i]
            dp, dv = max_abs(g["pos"], r["pos"]), max_abs(g["vel"], r["vel"])
            print(f"  frame {f} body {i}: max |dpos| {dp:.3e} max |dvel| {dv:.3e}", flush=True)
            assert float_equal(g["pos"], r["pos"]), f"{name}: positions differ at frame {f}, body {i}"
            assert float_equal(g["vel"], r["vel"]), f"{name}: velocities differ at frame {f}, body {i}"
            assert float_equal(g["body_pos"], r["body_pos"]), f"{name}: body position differs at frame {f}, body {i}"
